"""The oracle's f64 restatement against the reference's own cmf_pfq.cuh
compiled for the host (oracle/_ref): identical bits on random trajectories.
Both are built with -ffp-contract=off against the same libm."""
import numpy as np
import pytest

CASES = [(3, 2, 0.5), (3, 2, 1.0), (4, 3, 0.5), (5, 4, 1.0), (6, 5, 1.0), (6, 5, 0.5),
         (7, 6, 0.5), (8, 7, 1.0), (8, 7, 0.5), (5, 2, 0.5), (2, 3, 0.5)]
FIELDS = ["limit", "delta", "p", "q", "precision", "log_scale"]


def _same(a, b):
    return bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


@pytest.mark.parametrize("p,q,z", CASES)
def test_oracle_matches_reference_build(orc, p, q, z):
    if not orc.ref_available():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(1000 * p + 10 * q + int(4 * z))
    n = 96
    # mix of signed, non-negative and unit directions; starts in {1..12}
    D = rng.integers(-2, 3, (n, p + q))
    D[: n // 3] = np.abs(D[: n // 3])
    D[n // 3: n // 2] = 0
    D[n // 3: n // 2, 0] = 1
    S = rng.integers(1, 13, (n, p + q)).astype(float)
    a = orc.walk_f64_batch(p, q, z, D, S, 120, 20)
    b = orc.ref_walk_f64_batch(p, q, z, D, S, 120, 20)
    assert (a["valid"] == b["valid"]).all()
    assert (a["n_limits"] == b["n_limits"]).all()
    for f in FIELDS:
        assert _same(a[f], b[f]), f
    assert a["valid"].sum() > n // 10


def test_oracle_snapshots_bitequal(orc):
    if not orc.ref_available():
        pytest.skip("oracle/_ref not built")
    import ctypes as C

    for (p, q, z, d, s) in [(6, 5, 1.0, [1, 0, 2, 0, 1, 0, 0, 0, 0, 0, 0], [1] * 6 + [2] * 5),
                            (4, 3, 0.5, [1, -1, 1, 0, 1, 0, -1], [5, 7, 3, 9, 4, 6, 8])]:
        ok, snaps, ls = orc.ref_walk_f64_snaps(p, q, z, d, s, 60, 10)
        dsc = orc.make_desc(p, q, z)
        sn = orc.Snaps()
        D = orc.pad16(d, np.int32)
        S = orc.pad16(s, np.float64)
        ok2 = orc.lib().orc_forward_walk(orc._iptr(D), orc._dptr(S), C.byref(dsc), 60, 10,
                                         C.byref(sn))
        assert bool(ok) == bool(ok2)
        m = np.frombuffer(sn.m, dtype=np.float64).reshape(3, 8, 8)
        r = dsc.rank
        assert np.array_equal(m[:, :r, :r], snaps[:, :r, :r])
        assert np.array_equal(np.array(sn.ls[:]), ls)
