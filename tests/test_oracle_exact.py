"""Pins of the oracle that do not depend on the reference: exact rational
walks (dense Fraction products of I + C/a and its inverse, built from the
definition of the theta-companion matrix), closed forms, known answers."""
import math
from fractions import Fraction

import numpy as np
import pytest


def test_mt19937_64_known_answer(orc):
    # The C++ standard fixes the 10000th output of a default-seeded mt19937_64.
    g = orc.MT64(5489)
    x = None
    for _ in range(10000):
        x = g()
    assert x == 9981545732273789042


@pytest.mark.parametrize("p,q,z,rank", [(3, 2, 0.5, 3), (3, 2, 1.0, 2), (6, 5, 1.0, 5), (6, 5, 0.5, 6),
                                        (8, 7, 1.0, 7), (8, 7, 0.5, 8), (5, 2, 1.0, 5), (2, 3, 1.0, 4)])
def test_rank_rule(orc, p, q, z, rank):
    d = orc.make_desc(p, q, z)
    assert (d.rank, d.naxes) == (rank, p + q)


def test_poly_is_elementary_symmetric(orc):
    import ctypes as C
    import itertools

    vals = np.array([3.0, -2.0, 5.0, 7.0, 1.0, 4.0], dtype=np.float64)
    out = np.zeros(10)
    orc.lib().orc_poly_theta_plus(orc._dptr(vals), len(vals), orc._dptr(out))
    n = len(vals)
    for k in range(n + 1):
        # coefficient of theta^k = e_{n-k}(vals)
        e = sum(math.prod(c) for c in itertools.combinations(vals.tolist(), n - k))
        assert out[k] == e


@pytest.mark.parametrize("p,q,z", [(3, 2, 0.5), (6, 5, 1.0), (8, 7, 1.0), (4, 3, 0.5)])
def test_companion_column_closed_form(orc, p, q, z):
    rng = np.random.default_rng(p * 7 + q)
    d = orc.make_desc(p, q, z)
    for _ in range(20):
        pos = rng.integers(1, 9, p + q).astype(np.float64)
        col = np.zeros(8)
        P = orc.pad16(pos, np.float64)
        ok = orc.lib().orc_companion_column(orc._dptr(P), z, d.rank, p, q, orc._dptr(col))
        Cm = orc.exact_axis_matrix(list(pos), Fraction(z), d.rank, p, q)
        assert bool(ok) == (Cm is not None)
        if Cm is not None:
            for i in range(d.rank):
                assert abs(col[i] - float(Cm[i][d.rank - 1])) <= 1e-15 * max(1, abs(col[i]))


def test_zeta_lead_is_minus_one_and_yplus_dies(orc):
    """At z=1, p=q+1 the lead of D is -sum x + sum (y-1); at the zeta start it
    is -1 and a y+ first step (evaluated at y+1) makes it 0 -> invalid."""
    for p in (3, 4, 6, 8):
        q = p - 1
        s = [1] * p + [2] * q
        Cm = orc.exact_axis_matrix(s, Fraction(1), p - 1, p, q)
        assert Cm is not None
        ev = list(s)
        ev[p + q - 1] += 1
        assert orc.exact_axis_matrix(ev, Fraction(1), p - 1, p, q) is None
        d = [0] * (p + q)
        d[-1] = 1
        r = orc.walk_f64_batch(p, q, 1.0, [d], [s], 30, 5)[0]
        assert r["valid"] == 0
        d = [1] + [0] * (p + q - 1)
        assert orc.walk_f64_batch(p, q, 1.0, [d], [s], 30, 5)[0]["valid"] == 1


CASES = [
    (3, 2, 0.5, [1, 1, 1, 1, 1], [2, 3, 1, 4, 2]),
    (4, 3, 0.5, [1, 0, -1, 1, 0, 1, 0], [3, 5, 4, 2, 6, 3, 2]),
    (6, 5, 1.0, [1, 0, 2, 0, 1, 0, 0, 0, 0, 0, 0], [1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2]),
    (5, 4, 0.5, [2, 1, 0, 0, 1, -1, 0, 1, 0], [4, 2, 5, 3, 2, 6, 5, 4, 3]),
    (3, 2, 1.0, [2, 1, 1, 1, 0], [1, 1, 1, 2, 2]),
]


@pytest.mark.parametrize("p,q,z,d,s", CASES)
def test_walk_against_exact_product(orc, p, q, z, d, s):
    """W_N from the O(d^2) projective walk vs the exact dense product: every
    entry ratio W[i][c]/W[1][c] (projective invariant), validity, and
    log_scale + log|q| = log|q_exact|."""
    N, h = 12, 3
    W, ok = orc.exact_walk(p, q, Fraction(z), d, s, N)
    res, sn = orc.walk_f64_batch(p, q, z, [d], [s], N, h, want_snaps=True)
    assert bool(res["valid"][0]) == ok
    if not ok:
        return
    dim = orc.make_desc(p, q, z).rank
    M = sn["m"][0][2]
    ls = sn["ls"][0][2]
    c = dim - 1
    # projective invariance: compare to the exact matrix scaled by its max entry
    mx = max(abs(x) for row in W for x in row)
    for i in range(dim):
        for j in range(dim):
            assert abs(M[i][j] - float(W[i][j] / mx)) <= 1e-12, (i, j)
    assert abs(ls - orc.exact_log_abs(mx)) <= 1e-10 * max(1, abs(ls))
    if W[1][c] != 0:
        lq = math.log(abs(M[1][c])) + ls
        assert abs(lq - orc.exact_log_abs(W[1][c])) <= 1e-10 * max(1, abs(lq))


def test_llround_policy(orc):
    L = orc.lib().orc_llround
    for x in [0.5, -0.5, 1.5, -1.5, 2.4999999999999996, 1e15 + 0.5, -7.0, 4503599627370497.0]:
        r = math.floor(abs(Fraction(x)) + Fraction(1, 2))
        assert L(x) == (r if x >= 0 else -r)
    for x in [9.3e18, -9.3e18, float("inf"), float("nan"), 2.0 ** 63]:
        assert L(x) == -(2 ** 63)
    assert L(-(2.0 ** 63)) == -(2 ** 63)


def test_to_mod(orc):
    m = orc.modulus(0)
    for x in [0, 1, -1, 12345678901.0, -98765432109.0, 2.0 ** 62]:
        assert orc.lib().orc_to_mod(x, m) == int(x) % m
    # INT64_MIN from the out-of-range policy: (m - 2^63 mod m) mod m
    assert orc.lib().orc_to_mod(1e19, m) == (-(2 ** 63)) % m


def test_moduli_table_facts(orc):
    import sympy

    ms = [orc.modulus(k) for k in range(64)]
    assert len(set(ms)) == 64 and all(m < 2 ** 30 for m in ms)
    primes = [k for k, m in enumerate(ms) if sympy.isprime(m)]
    assert primes == [0, 1, 2, 3, 4, 19, 21, 23, 24, 26, 35, 56]
    assert ms[5] == 113 * 9502139 and ms[7] == 7 * 23 * 107 * 157 * 397


def test_crt_entry_matches_python_crt(orc):
    import ctypes as C

    m0, m1 = orc.modulus(0), orc.modulus(1)
    rng = np.random.default_rng(5)
    for _ in range(2000):
        v = int(rng.integers(0, m0 * m1 // 2 ** 4)) * 16 + int(rng.integers(0, 16))
        v %= m0 * m1
        r0, r1 = v % m0, v % m1
        if r0 > r1 + m1:  # the reference's unsigned wrap case; covered on the GPU
            continue
        out = C.c_uint64()
        dbl = orc.lib().orc_crt_entry(r0, r1, 2, C.byref(out))
        assert out.value == v
        want = -(m0 * m1 - v) if v > (m0 * m1) // 2 else v
        assert dbl == float(want)


@pytest.mark.parametrize("p,q,d,s,steps", [
    (3, 2, [1, 1, 1, 1, 1], [2, 3, 1, 5, 7], 9),
    (4, 3, [1, 0, 1, 0, 1, 0, 1], [3, 5, 4, 2, 9, 7, 11], 7),
    (6, 5, [1, 1, 0, 0, 1, 0, 1, 0, 0, 0, 1], [2, 3, 2, 4, 3, 5, 9, 10, 11, 12, 13], 5),
    (4, 3, [1, 0, 1, 0, -1, 0, 1], [3, 5, 4, 2, 9, 3, 2], 7),
])
def test_rns_prime_lanes_equal_exact_residues(orc, p, q, d, s, steps):
    """z=1/2, p=q+1: every column is an integer, so lane k holds the exact
    rational product reduced mod m_k whenever m_k is prime (lanes 0-4)."""
    ok, W = orc.rns_raw_walk(p, q, 0.5, d, s, steps, n_primes=8)
    E, ok_e = orc.exact_walk(p, q, Fraction(1, 2), d, s, steps)
    assert ok == ok_e  # the last case is singular (x_j == y at a y+ step)
    if not ok:
        return
    dim = orc.make_desc(p, q, 0.5).rank
    for k in range(5):
        m = orc.modulus(k)
        for i in range(dim):
            for j in range(dim):
                f = E[i][j]
                want = f.numerator % m * pow(f.denominator % m, -1, m) % m
                assert int(W[k][i][j]) == want, (k, i, j)


def test_rns_composite_lane_step_identities(orc):
    """On a composite modulus the 'inverse' a^(m-2) is only a fixed power s.
    The inverse step must still satisfy (out*(I+C*s))[j] = M[j] for j<d-1 and
    the last column to within the factor w*sd, with w = sd^(m-2)."""
    rng = np.random.default_rng(9)
    for k in (5, 6, 7, 8):
        m = orc.modulus(k)
        dim = 5
        for _ in range(20):
            M = rng.integers(0, m, (8, 8)).astype(np.uint32)
            col = rng.integers(0, m, 8).astype(np.uint32)
            den = int(rng.integers(2, 50))
            out = M.copy()
            ok = orc.lib().orc_rns_right_mul_lane(orc._uptr(out), dim, orc._uptr(col), den, 1, m)
            s = pow(den, m - 2, m)
            Mi = [[int(M[i][j]) for j in range(dim)] for i in range(dim)]
            Oi = [[int(out[i][j]) for j in range(dim)] for i in range(dim)]
            cl = [int(c) for c in col[:dim]]
            # beta (row independent) and sd from their definition
            beta = [0] * dim
            beta[dim - 1] = 1
            for j in range(dim - 2, -1, -1):
                beta[j] = (-beta[j + 1] * s) % m
            sd = (1 + sum(b * c for b, c in zip(beta, cl)) * s) % m
            assert bool(ok) == (sd != 0)
            if not ok:
                continue
            w = pow(sd, m - 2, m)
            for r in range(dim):
                o = Oi[r]
                for j in range(dim - 1):
                    assert (o[j] + o[j + 1] * s) % m == Mi[r][j]
                # last column: o[d-1] + (o.col)s = t*sd + da*s with t=(M-da*s)*w
                da = (sum(x * c for x, c in zip(o, cl)) - o[dim - 1] * 0) % m
                lhs = (o[dim - 1] + da * s) % m
                # lhs - (o[d-1]*sd - o[d-1]) ... verify via t: o[d-1] = t
                t = o[dim - 1]
                alpha_dot = (da - t * sum(b * c for b, c in zip(beta, cl))) % m
                assert t == ((Mi[r][dim - 1] - alpha_dot * s) * w) % m
                assert lhs == (t + da * s) % m


def test_legacy_inputs_aliasing_map(orc):
    """Limit mode writes 128-byte points at 192-byte struct stride: struct
    2k+1 holds point 3k+2 intact (cmf_walk_f64.cu:178)."""
    p, q, n = 3, 2, 30
    D, S = orc.legacy_inputs(p, q, n, "limit", False)
    g = orc.MT64(42)
    pts = [[2.0 + g() % 19 for _ in range(p + q)] for _ in range(n)]
    for k in range(n // 2 - 1):
        j = 2 * k + 1
        if 3 * k + 2 < n and j < n:
            assert list(S[j][:p + q]) == pts[3 * k + 2]
    assert (D[:, 0] == 1).all() and (D[:, 1:] == 0).all()
    # delta mode: struct j >= ceil(n/3) keeps an all-zero direction
    D, S = orc.legacy_inputs(6, 5, 30, "delta", True)
    assert (D[10:] == 0).all()
    assert (D[:10, :11] != 0).any(axis=1).all()


@pytest.mark.parametrize("p,q,z,d,s", CASES)
def test_quad_walk_against_exact_product(orc, p, q, z, d, s):
    """The quad-precision accuracy oracle against exact rationals: the
    primary ratio p/q and the log scale of the normalised product."""
    N, h = 12, 3
    W, ok = orc.exact_walk(p, q, Fraction(z), d, s, N)
    hp = orc.hp_walk_batch(p, q, z, [d], [s], N, h)
    if not ok:
        return
    dim = orc.make_desc(p, q, z).rank
    c = dim - 1
    if W[1][c] != 0:
        exact = W[0][c] / W[1][c]
        assert abs(hp["ratio"][0][2] - float(exact)) <= 1e-15 * max(1.0, abs(float(exact)))
    mx = max(abs(x) for row in W for x in row)
    assert abs(hp["log_scale"][0][2] - orc.exact_log_abs(mx)) <= 1e-13 * max(1, abs(hp["log_scale"][0][2]))
