"""GPU checks of mixed batches (cmfw_batch_*): trajectories of every CMF
from 3F2 to 8F7, z in {0.5, 1} and different depths in one graph launch
agree with the reference-order walk of each (CMF, z, N) group -- validity,
executed steps and limit presence identical, values at least as accurate
against the quad-precision oracle -- and graph replays are bit-identical."""
import numpy as np
import pytest

from test_gpu_walk_fast import check, _with_counts

pytestmark = pytest.mark.gpu

CMFS = [(3, 2), (4, 3), (5, 4), (6, 5), (7, 6), (8, 7)]


def _mixed(seed, n, depths):
    rng = np.random.default_rng(seed)
    k = rng.integers(0, len(CMFS), n)
    p = np.array([CMFS[i][0] for i in k], np.int32)
    q = np.array([CMFS[i][1] for i in k], np.int32)
    z = rng.choice([0.5, 1.0], n)
    N = rng.choice(depths, n).astype(np.int32)
    D = np.zeros((n, 16), np.int32)
    S = np.zeros((n, 16))
    for t in range(n):
        na = p[t] + q[t]
        D[t, :na] = rng.integers(0, 4, na)
        if not D[t, :na].any():
            D[t, 0] = 1
        S[t, :na] = rng.integers(1, 9, na)
    return p, q, z, N, D, S


@pytest.mark.parametrize("tune", [False, True])
def test_batch_matches_reference_order_per_group(engine, orc, tune):
    h = 20
    # depths include invalid ones (N - 2h <= 0)
    p, q, z, N, D, S = _mixed(7 + tune, 1500, [30, 41, 120, 200, 333])
    print(f"batch tune={tune}", flush=True)
    b = engine.Batch(p, q, z, N, h, D, S, tune=tune)
    info = b.info()
    print(f"  {info}", flush=True)
    assert sum(x["count"] for x in info["buckets"]) == len(p)
    g = engine.results_to_numpy(b.run())
    g2 = engine.results_to_numpy(b.run())
    for f in g.dtype.names:
        assert np.array_equal(g[f], g2[f], equal_nan=True), f
    groups = {}
    for t in range(len(p)):
        groups.setdefault((int(p[t]), int(q[t]), float(z[t]), int(N[t])), []).append(t)
    for (pp, qq, zz, NN), idx in sorted(groups.items()):
        idx = np.array(idx)
        na = pp + qq
        Dg, Sg = D[idx, :na], S[idx, :na]
        e = engine.results_to_numpy(engine.walk_f64(pp, qq, zz, Dg, Sg, NN, h))
        tag = f"{pp}F{qq} z={zz} N={NN}"
        if NN - 2 * h <= 0:
            assert (g["valid"][idx] == 0).all(), tag
            continue
        hp = orc.hp_walk_batch(pp, qq, zz, Dg, Sg, NN, h) if NN <= 120 else None
        check(g[idx], e, tag, hp)
    b.close()


def test_batch_empty_and_errors(engine):
    import cmfwalk

    b = engine.Batch([], [], [], [], 10, np.zeros((0, 16)), np.zeros((0, 16)))
    b.run()
    b.close()
    with pytest.raises(cmfwalk.CmfwError):
        engine.Batch([9], [1], [0.5], [100], 10, np.ones((1, 16)), np.ones((1, 16)))
