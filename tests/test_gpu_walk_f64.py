"""GPU parity of cmfw_walk_f64 (HIP, through the C ABI) with the oracle.

The kernel evaluates every double operation in the reference's order with no
FMA contraction, so walk matrices, p, q and limits are bit-identical to the
oracle; only log() comes from a different math library, so log_scale is
compared to 1e-13 relative and delta/precision to 1e-9 absolute."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same_bits(a, b):
    return bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def _close(a, b, rtol=0.0, atol=0.0):
    both_nan = np.isnan(a) & np.isnan(b)
    same_inf = np.isinf(a) & np.isinf(b) & (np.sign(a) == np.sign(b))
    fin = np.isfinite(a) & np.isfinite(b)
    ok = both_nan | same_inf | (fin & (np.abs(a - b) <= atol + rtol * np.abs(b)))
    return bool(ok.all())


def compare(gpu, ref, tag=""):
    assert (gpu["valid"] == ref["valid"]).all(), tag
    assert (gpu["n_limits"] == ref["n_limits"]).all(), tag
    for f in ("p", "q", "limit"):
        assert _same_bits(gpu[f], ref[f]), (tag, f)
    assert _close(gpu["log_scale"], ref["log_scale"], rtol=1e-13, atol=1e-13), tag
    assert _close(gpu["precision"], ref["precision"], atol=1e-9), tag
    assert _close(gpu["delta"], ref["delta"], atol=1e-9), tag


CASES = [(3, 2, 0.5), (3, 2, 1.0), (4, 3, 0.5), (5, 4, 1.0), (6, 5, 1.0), (6, 5, 0.5),
         (7, 6, 0.5), (8, 7, 1.0), (8, 7, 0.5), (5, 2, 0.5), (2, 3, 0.5)]


@pytest.mark.parametrize("p,q,z", CASES)
def test_walk_f64_matches_oracle_random(engine, orc, p, q, z):
    rng = np.random.default_rng(7 * p + 3 * q + int(z * 10))
    n = 333  # ragged: not a multiple of the wave size
    D = rng.integers(-2, 3, (n, p + q))
    D[: n // 2] = np.abs(D[: n // 2])
    D[n // 2: n // 2 + 40] = 0
    D[n // 2: n // 2 + 40, 0] = 1
    S = rng.integers(1, 13, (n, p + q)).astype(float)
    N, h = 150, 30
    print(f"walk_f64 {p}F{q} z={z} n={n} N={N}", flush=True)
    gpu = engine.results_to_numpy(engine.walk_f64(p, q, z, D, S, N, h))
    ref = orc.walk_f64_batch(p, q, z, D, S, N, h)
    compare(gpu, ref, f"{p}F{q} z={z}")
    assert gpu["valid"].sum() > 20


def test_walk_f64_snapshots_bit_identical(engine, orc):
    import torch

    p, q, z = 6, 5, 1.0
    rng = np.random.default_rng(3)
    n = 70
    D = rng.integers(0, 3, (n, p + q))
    D[:, p:] = 0
    D[:, 0] = np.maximum(D[:, 0], 1)
    S = np.tile(np.array([1] * p + [2] * q, float), (n, 1))
    N, h = 90, 20
    snaps = engine.alloc_snapshots(5, n)
    engine.walk_f64(p, q, z, D, S, N, h, snapshots=snaps)
    torch.cuda.synchronize()
    m = snaps["m"].cpu().numpy()  # [3][5][5][n]
    _, ref = orc.walk_f64_batch(p, q, z, D, S, N, h, want_snaps=True)
    for t in range(n):
        if not ref["valid"][t]:
            continue
        assert np.array_equal(m[:, :, :, t], ref["m"][t][:, :5, :5]), t


def test_published_runs_on_gpu(engine, published):
    """Legacy (stride-aliased) inputs of every published f64 run: validity
    and limit counts equal the published ones exactly."""
    import sys

    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
    import oracle

    for run in published:
        if run["variant"] != "f64":
            continue
        p, q, z = run["p"], run["q"], float(run["z"])
        D, S = oracle.legacy_inputs(p, q, run["n_traj"], run["mode"], run["suite"] == "zeta")
        print(f"published {run['suite']} {run['cmf']} {run['mode']}", flush=True)
        g = engine.results_to_numpy(engine.walk_f64(p, q, z, D[:, :p + q], S[:, :p + q],
                                                    run["N"], run["h"]))
        valid = g["valid"].astype(bool)
        assert int(valid.sum()) == run["n_valid"], run["cmf"]
        assert int((valid & ~np.isnan(g["limit"])).sum()) == run["n_with_limits"], run["cmf"]
        if run["suite"] == "zeta" and run["mode"] == "limit":
            rec = run["records"]
            for f in ("p", "q", "log_scale", "limit"):
                a = rec[f][0]
                assert abs(a - g[f][0]) <= 1e-9 * max(1, abs(a)), (run["cmf"], f)
            assert g["n_limits"][0] == rec["n_limits"][0]


def test_edge_cases(engine, orc):
    p, q, z = 3, 2, 0.5
    D = np.array([[1, 1, 1, 1, 1], [0, 0, 0, 0, 0], [-1, 0, 0, 0, 0], [1, 0, 0, 0, 0]])
    S = np.array([[2, 3, 1, 5, 7], [2, 3, 1, 5, 7], [1, 3, 2, 5, 4], [0, 3, 2, 5, 4]], float)
    # N - 2h <= 0: every trajectory invalid, as forward_walk returns false
    g = engine.results_to_numpy(engine.walk_f64(p, q, z, D, S, 100, 50))
    assert (g["valid"] == 0).all() and np.isnan(g["limit"]).all()
    # identity walk (all-zero direction), x- through zero, x0 = 0 start
    g = engine.results_to_numpy(engine.walk_f64(p, q, z, D, S, 40, 10))
    r = orc.walk_f64_batch(p, q, z, D, S, 40, 10)
    compare(g, r, "edge")
    assert list(g["valid"]) == [1, 1, 0, 0]
    # n = 0 is a no-op, n = 1 works
    import torch

    out = engine.walk_f64(p, q, z, torch.zeros((5, 0), dtype=torch.int32, device="cuda"),
                          torch.zeros((5, 0), dtype=torch.float64, device="cuda"), 40, 10)
    assert out["valid"].numel() == 0
    g = engine.results_to_numpy(engine.walk_f64(p, q, z, D[:1], S[:1], 40, 10))
    assert g["valid"][0] == 1


def test_rank8_large_batch_against_oracle_sample(engine, orc):
    """8F7 z=1/2 (rank 8, the largest state), 4096 trajectories; every 16th
    checked against the oracle."""
    p, q, z = 8, 7, 0.5
    rng = np.random.default_rng(11)
    n = 4096
    D = np.ones((n, p + q), int)
    S = rng.integers(1, 17, (n, p + q)).astype(float)
    N, h = 200, 10
    g = engine.results_to_numpy(engine.walk_f64(p, q, z, D, S, N, h))
    idx = np.arange(0, n, 16)
    r = orc.walk_f64_batch(p, q, z, D[idx], S[idx], N, h)
    compare(g[idx], r, "8F7")
