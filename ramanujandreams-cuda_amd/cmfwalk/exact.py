"""Exact verification of walk results (cmfw_walk_exact).

The GPU walks the integer-scaled walk matrix W' (the exact walk matrix
times a nonzero scalar) mod K primes, one lane per (trajectory, prime);
this module reconstructs the entries it needs by CRT with Python integers
and turns them into exact rationals.  An entry is exact once the product
of the primes exceeds twice its magnitude: `walk` checks that by
reconstructing from all K primes and from the first K - `check` primes,
which agree only then (the bit length of the reconstructed entries is the
walk's measured bit growth).

    res = exact.walk(8, 7, 1.0, dirs, starts, N=30, h=5, K=256)
    res[t]["status"], res[t]["W"][2][(0, d - 1)]       # exact int at step N
    exact.pq(res[t])                                    # Fraction p/q at N
"""
import ctypes as C
from fractions import Fraction

import numpy as np

from . import (_check, _inputs_struct, _stream_ptr, load, load_library, make_desc,
               to_device_inputs)

STATUS = {1: "ok", 0: "fails", 2: "off-grid", 4: "prime-divides"}


def primes(K):
    """The K largest primes below 2^30, descending (the library's list)."""
    arr = (C.c_uint32 * K)()
    _check(load_library().cmfw_exact_primes(int(K), arr))
    return np.frombuffer(arr, dtype=np.uint32).copy()


def residues(p, q, z, dirs, starts, N, h, prime_list, stream=None):
    """W' mod each prime at the snapshots N-2h, N-h, N: (snap uint32
    [3][d][d][n][K], status int32 [n][K], fail_step int32 [n][K])."""
    import torch

    L = load()
    d = make_desc(p, q, z)
    dt, st = to_device_inputs(dirs, starts)
    n = dt.shape[1]
    K = len(prime_list)
    D = d.rank
    snap = torch.zeros(3 * D * D * n * K, dtype=torch.int32, device=dt.device)
    status = torch.empty(n * K, dtype=torch.int32, device=dt.device)
    fail = torch.empty(n * K, dtype=torch.int32, device=dt.device)
    pr = (C.c_uint32 * K)(*[int(x) for x in prime_list])
    _check(L.cmfw_walk_exact(C.byref(d), C.byref(_inputs_struct(dt, st)), int(N), int(h), K, pr,
                             C.c_void_p(snap.data_ptr()), C.c_void_p(status.data_ptr()),
                             C.c_void_p(fail.data_ptr()), _stream_ptr(stream)))
    snap_np = snap.cpu().numpy().view(np.uint32).reshape(3, D, D, n, K)
    return snap_np, status.cpu().numpy().reshape(n, K), fail.cpu().numpy().reshape(n, K)


class CRT:
    """Chinese remaindering onto the symmetric range (-M/2, M/2]."""

    def __init__(self, prime_list):
        self.m = [int(x) for x in prime_list]
        self.M = 1
        for x in self.m:
            self.M *= x
        self.c = []
        for x in self.m:
            Mk = self.M // x
            self.c.append(Mk * pow(Mk % x, -1, x))

    def __call__(self, r):
        v = sum(int(a) * c for a, c in zip(r, self.c)) % self.M
        return v - self.M if 2 * v > self.M else v


def _traj_status(st, fs):
    """One trajectory's outcome from its K lanes."""
    if (st == 2).any():
        return "off-grid", -1
    if (st == 1).all():
        return "ok", -1
    if (st == 0).all() and (fs == fs[0]).all():
        return "fails", int(fs[0])
    if (st == 4).all() and (fs == fs[0]).all():
        return "fails", int(fs[0])  # zero modulo every prime: the solve denominator vanishes
    return "prime-divides", int(fs[st != 1].min())


def walk(p, q, z, dirs, starts, N, h, K=256, entries=None, check=8, stream=None):
    """Exact snapshots of W' per trajectory.  entries: the (i, j) entries to
    reconstruct (default all).  Returns a list of dicts: status ("ok",
    "fails", "off-grid", "prime-divides"), fail_step, W = {snapshot index:
    {(i, j): int}} for the snapshots inside 1..N, exact = every
    reconstructed entry agrees between K and K - check primes, bits = the
    largest bit length among them."""
    D = make_desc(p, q, z).rank
    plist = primes(K)
    snap, status, fail = residues(p, q, z, dirs, starts, N, h, plist, stream)
    full, part = CRT(plist), CRT(plist[:K - check]) if check and K > check else None
    entries = entries or [(i, j) for i in range(D) for j in range(D)]
    cps = [N - 2 * h, N - h, N]
    out = []
    for t in range(status.shape[0]):
        s, f = _traj_status(status[t], fail[t])
        rec = {"status": s, "fail_step": f, "W": {}, "exact": True, "bits": 0}
        if s in ("ok", "fails"):
            for sn, cp in enumerate(cps):
                if not (1 <= cp <= N) or (s == "fails" and cp >= f):
                    continue
                Wsn = {}
                for (i, j) in entries:
                    r = snap[sn, i, j, t]
                    v = full(r)
                    if part is not None and part(r[:K - check]) != v:
                        rec["exact"] = False
                    rec["bits"] = max(rec["bits"], abs(v).bit_length())
                    Wsn[(i, j)] = v
                rec["W"][sn] = Wsn
        out.append(rec)
    return out


def pq(rec, sn=2, d=None):
    """Exact p/q = W[0][d-1] / W[1][d-1] at snapshot sn (None when q = 0 or
    the snapshot is missing)."""
    W = rec["W"].get(sn)
    if not W:
        return None
    d = d or (max(j for (_, j) in W) + 1)
    num, den = W.get((0, d - 1)), W.get((1, d - 1))
    if num is None or den is None or den == 0:
        return None
    return Fraction(num, den)


def verify_pq(p, q, z, dirs, starts, N, h, values, K=256, check=8, stream=None):
    """Relative error of each given p/q (at depth N) against the exact ratio:
    an array with NaN where no exact ratio exists (failed walk, q = 0, or K
    too small to be exact)."""
    D = make_desc(p, q, z).rank
    recs = walk(p, q, z, dirs, starts, N, h, K, entries=[(0, D - 1), (1, D - 1)], check=check,
                stream=stream)
    err = np.full(len(recs), np.nan)
    for t, r in enumerate(recs):
        e = pq(r, 2, D) if r["status"] == "ok" and r["exact"] else None
        if e is None or not np.isfinite(values[t]):
            continue
        ex = float(e)
        err[t] = abs(float(Fraction(values[t]) - e)) / max(abs(ex), 1e-300)
    return err, recs
