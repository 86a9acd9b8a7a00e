"""The exact walk's prime list (host code of the C ABI, no GPU needed)."""
import numpy as np
import pytest


def _is_prime(n):
    if n < 2:
        return False
    if n % 2 == 0:
        return n == 2
    f = 3
    while f * f <= n:
        if n % f == 0:
            return False
        f += 2
    return True


def test_exact_primes_are_the_largest_below_2_30():
    from cmfwalk import exact

    pr = exact.primes(200)
    assert len(pr) == 200 and len(set(pr.tolist())) == 200
    assert (np.diff(pr.astype(np.int64)) < 0).all() and int(pr[0]) < 2 ** 30
    assert all(_is_prime(int(x)) for x in pr)
    # no prime is skipped between consecutive entries
    for a, b in zip(pr[:40].tolist(), pr[1:41].tolist()):
        assert not any(_is_prime(c) for c in range(b + 1, a))
    assert int(pr[0]) == 1073741789  # the largest prime below 2^30


def test_exact_primes_rejects_bad_counts():
    from cmfwalk import CmfwError, exact

    with pytest.raises(CmfwError):
        exact.primes(0)
    with pytest.raises(CmfwError):
        exact.primes(10 ** 6)


def test_crt_symmetric_range():
    from cmfwalk import exact

    pr = exact.primes(6)
    crt = exact.CRT(pr)
    for v in (0, 1, -1, 12345678901234567890, -98765432109876543210, crt.M // 2 - 3):
        assert crt([v % int(m) for m in pr]) == v
