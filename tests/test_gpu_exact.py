"""GPU checks of the exact verification walk (cmfw_walk_exact): the
integer-scaled walk matrix reconstructed by CRT is proportional to the exact
rational product of the axis matrices (the oracle's Fraction walk), entry by
entry; validity matches; exactness is detected; and it measures the error of
the floating-point walks against exact p/q."""
from fractions import Fraction

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [  # p, q, z, direction, start (the last one fails at its first step)
    (3, 2, 1.0, [1, 0, 0, 0, 0], [1, 1, 1, 2, 2]),
    (5, 4, 1.0, [2, 1, 0, 0, 1, 0, 0, 0, 0], [1, 1, 1, 1, 1, 2, 2, 2, 2]),
    (8, 7, 1.0, [1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0], [1] * 8 + [2] * 7),
    (3, 2, 0.5, [1, 2, 0, 2, 0], [1, 6, 2, 4, 3]),
    (4, 3, 0.5, [0, 1, 0, 1, 0, 0, 1], [4, 6, 3, 3, 1, 1, 5]),
    (6, 5, 0.5, [0, 1, 1, 1, 0, 1, 0, 0, 1, 2, 1], [4, 1, 5, 3, 5, 3, 1, 7, 7, 7, 7]),
    (4, 3, 1.0, [0, 2, 1, 1, 2, 2, 0], [2, 3, 2, 2, 6, 5, 4]),
    (3, 3, 0.5, [0, 1, 2, 0, 0, 0], [6, 2, 4, 6, 5, 6]),
    (5, 3, 1.0, [1, 1, 1, 1, 1, 0, 0, 1], [5, 6, 4, 5, 4, 4, 1, 3]),
    (2, 1, 0.5, [0, 1, 1], [1, 7, 5]),
    (3, 2, 0.5, [1, 2, 0, 1, 0], [2, 3, 5, 2, 4]),
]


def _proportional(Wp, W, d):
    """Wp = lambda W with lambda != 0 (exact cross products)."""
    piv = max(((i, j) for i in range(d) for j in range(d)), key=lambda ij: abs(W[ij[0]][ij[1]]))
    a, b = Wp[piv], W[piv[0]][piv[1]]
    if b == 0 or a == 0:
        return False
    return all(Fraction(Wp[(i, j)]) * b == a * W[i][j] for i in range(d) for j in range(d))


@pytest.mark.parametrize("case", range(len(CASES)))
def test_exact_walk_proportional_to_fraction_product(engine, orc, case):
    from cmfwalk import exact

    p, q, z, d_, s_ = CASES[case]
    N, h = 12, 3
    print(f"exact walk {p}F{q} z={z} dir={d_}", flush=True)
    recs = exact.walk(p, q, z, [d_], [[float(v) for v in s_]], N, h, K=160)
    r = recs[0]
    W, ok = orc.exact_walk(p, q, Fraction(z), d_, s_, N)
    assert (r["status"] == "ok") == ok, r["status"]
    if not ok:
        return
    assert r["exact"]
    d = orc.make_desc(p, q, z).rank
    assert _proportional(r["W"][2], W, d)
    W6, ok6 = orc.exact_walk(p, q, Fraction(z), d_, s_, N - 2 * h)
    assert ok6 and _proportional(r["W"][0], W6, d)
    # the exact p/q
    if W[1][d - 1] != 0:
        assert exact.pq(r, 2, d) == W[0][d - 1] / W[1][d - 1]


def test_exact_walk_failures_match(engine, orc):
    """Walks that fail (zero lead on a y-first zeta step, zero denominators)
    fail at the oracle's step."""
    from cmfwalk import exact

    p, q, z = 4, 3, 1.0
    rng = np.random.default_rng(3)
    D = rng.integers(-1, 3, (40, p + q))
    D[np.abs(D).sum(1) == 0, 0] = 1
    S = rng.integers(-1, 4, (40, p + q)).astype(float)
    recs = exact.walk(p, q, z, D, S, 8, 2, K=96)
    for t in range(40):
        _, ok = orc.exact_walk(p, q, Fraction(z), D[t].tolist(), S[t].tolist(), 8)
        assert (recs[t]["status"] == "ok") == ok, (t, recs[t]["status"])


def test_exact_walk_detects_too_few_primes(engine):
    from cmfwalk import exact

    p, q, z = 6, 5, 0.5
    d_ = [1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0]
    s_ = [2.0, 3, 4, 5, 6, 7, 2, 3, 4, 5, 6]
    small = exact.walk(p, q, z, [d_], [s_], 40, 10, K=12, check=4)[0]
    big = exact.walk(p, q, z, [d_], [s_], 40, 10, K=256, check=8)[0]
    assert small["status"] == big["status"] == "ok"
    assert not small["exact"] and big["exact"]
    assert big["bits"] > 12 * 30 // 2


def test_exact_bit_growth_is_linear_in_depth(engine):
    """Bit length of W' for a 6F5 walk at increasing depth: about linear
    (each axis step multiplies by integers of bounded size)."""
    from cmfwalk import exact

    p, q, z = 6, 5, 0.5
    d_ = [1, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0]
    s_ = [2.0, 3, 4, 5, 6, 7, 2, 3, 4, 5, 6]
    bits = []
    for N in (10, 20, 40, 80):
        r = exact.walk(p, q, z, [d_], [s_], N, 5, K=1024, check=16)[0]
        assert r["status"] == "ok" and r["exact"], N
        bits.append(r["bits"])
    print("bits at N = 10, 20, 40, 80:", bits, flush=True)
    assert bits[0] < bits[1] < bits[2] < bits[3]
    per_step = [(b1 - b0) / (n1 - n0) for (b0, n0), (b1, n1) in
                zip(zip(bits, (10, 20, 40, 80)), zip(bits[1:], (20, 40, 80)))]
    assert max(per_step) < 1.6 * min(per_step)


@pytest.mark.parametrize("walk", ["walk_f64", "walk_f64_fast", "walk_f32x2"])
def test_float_walks_against_exact_pq(engine, walk):
    """Well-conditioned walks (4F3 z=1/2, x-only directions, starts 2..9):
    the reference-order and fast FP64 p/q within 1e-12 of the exact ratio,
    two-float within 1e-10."""
    from cmfwalk import exact

    p, q, z = 4, 3, 0.5
    rng = np.random.default_rng(11)
    n = 24
    D = np.zeros((n, p + q), np.int64)
    D[:, :p] = rng.integers(0, 3, (n, p))
    D[D.sum(1) == 0, 0] = 1
    S = rng.integers(2, 10, (n, p + q)).astype(float)
    N, h = 60, 10
    res = engine.results_to_numpy(getattr(engine, walk)(p, q, z, D, S, N, h))
    vals = res["p"] / res["q"]
    err, recs = exact.verify_pq(p, q, z, D, S, N, h, vals, K=512, check=8)
    ok = np.isfinite(err)
    assert ok.sum() >= n // 2
    tol = 1e-10 if walk == "walk_f32x2" else 1e-12
    print(f"{walk}: max error vs exact {np.nanmax(err):.2e} over {int(ok.sum())} walks", flush=True)
    assert np.nanmax(err) <= tol
