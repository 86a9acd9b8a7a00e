#!/usr/bin/env python3
"""Benchmark of the MI355X CMF walk engine: verified-delta trajectories/s on
8F7, z=1, N=1000 (zeta start, candidate directions in {0..4}^15, 4,000,000
candidates from std::mt19937_64(42)).

One step = the whole verified-delta pipeline over the candidate batch, inputs
already resident in HBM:
  1. screen: the first trajectory step of every candidate (exact: a
     trajectory that fails there fails the full walk the same way);
  2. compact the survivors and sort them by direction so the lanes of a wave
     follow the same axis sequence;
  3. deep walk: plain f64 (the fast engine: depth segments, FMA, power-of-two
     rescaling), N=1000, h=50, limits and delta on device;
  4. verify: the top 1% of candidates by delta walked again in RNS
     (8 moduli, N=1000) -- all survivors when fewer than 1% survive.
value = candidates resolved per second.  Kernel times are HIP-event
measurements on the stream the kernels run on.

python bench.py [--steps K] [--warmup W] [--gpus 1] [--n-cand N]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "ramanujandreams-cuda_amd")
sys.path.insert(0, PKG)

P, Q, Z, N_DEPTH, H = 8, 7, 1.0, 1000, 50
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector (and matrix) peak, AMD spec


def canonical_flops(d, p, q):
    """Per executed axis step, counting FMA as 2 and add/mul/div as 1 (the
    reference recurrence: column O(p^2+q^2), O(d^2) update, normalisation)."""
    c = 1.5 * p * (p + 1) + 1.5 * q * (q + 1) + p + q + 2
    return c + 5 * d * d + d + 2, c + 10 * d * d + 3 * d + 2


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-cand", type=int, default=4_000_000)
    ap.add_argument("--top-frac", type=float, default=0.01)
    ap.add_argument("--n-primes", type=int, default=8)
    ap.add_argument("--cpu-sample", type=int, default=1_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    return ap.parse_args()


class Pipeline:
    def __init__(self, torch, cw, n_cand, top_frac, n_primes):
        self.torch, self.cw = torch, cw
        self.desc = cw.make_desc(P, Q, Z)
        self.n = n_cand
        D, S = cw.synth_inputs(42, n_cand, P, Q, 1, 0, 0, 4)
        self.host_D, self.host_S = D, S
        naxes = P + Q
        self.dir = torch.from_numpy(D[:, :naxes].T.copy()).cuda()
        self.start = torch.from_numpy(S[:, :naxes].T.copy()).cuda()
        self.pre = cw.alloc_results(n_cand)
        self.n_top = max(1, int(round(top_frac * n_cand)))
        self.n_primes = n_primes
        torch.cuda.synchronize()

    def run(self, timers=None):
        torch, cw = self.torch, self.cw
        ev = timers if timers is not None else {}

        def mark(name):
            e = torch.cuda.Event(enable_timing=True)
            e.record()
            ev.setdefault(name, []).append(e)

        mark("t0")
        cw.walk_prefix_f64(P, Q, Z, self.dir, self.start, 1, out=self.pre)
        mark("screen")
        live = torch.nonzero(self.pre["valid"]).squeeze(1)
        # direction key: axes from the last to the first, 3 bits each, so
        # neighbours in the sorted order share their first axis steps
        dl = self.dir[:, live]
        key = torch.zeros(live.numel(), dtype=torch.int64, device="cuda")
        for a in range(P + Q - 1, -1, -1):
            key = key * 8 + dl[a].to(torch.int64)
        order = torch.argsort(key)
        live = live[order]
        dl = self.dir[:, live].contiguous()
        sl = self.start[:, live].contiguous()
        mark("compact")
        res = cw.walk_f64_fast(P, Q, Z, dl, sl, N_DEPTH, H)
        mark("deep")
        delta = res["delta"]
        ok = (res["valid"] == 1) & torch.isfinite(delta)
        score = torch.where(ok, delta, torch.full_like(delta, -float("inf")))
        # candidates outside the survivors rank below every survivor
        k = min(self.n_top, live.numel())
        top = torch.topk(score, k).indices if k > 0 else live[:0]
        dv = dl[:, top].contiguous()
        sv = sl[:, top].contiguous()
        mark("select")
        rns = cw.walk_rns(P, Q, Z, dv, sv, N_DEPTH, H, n_primes=self.n_primes)
        mark("verify")
        return {"live": live, "res": res, "top": top, "rns": rns}


def stage_ms(ev, names, i):
    out = {}
    for a, b in zip(names[:-1], names[1:]):
        out[b] = ev[a][i].elapsed_time(ev[b][i])
    return out


def cpu_baseline(D, S, sample, top_frac, n_primes):
    """The reference's own cmf_pfq.cuh (oracle/_ref, compiled for the host
    with OpenMP) walks a bounded prefix of the same candidates in f64; the
    top 1% of the sample by delta is then walked in RNS by the restatement
    (the reference's RNS driver needs the CUDA runtime)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle

    n = min(sample, D.shape[0])
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = min(threads, 16)
    kind = "reference" if oracle.ref_available() else "port"
    t0 = time.perf_counter()
    if kind == "reference":
        r = oracle.ref_walk_f64_batch(P, Q, Z, D[:n, :P + Q], S[:n, :P + Q], N_DEPTH, H,
                                      nthreads=threads)
    else:
        r = oracle.walk_f64_batch(P, Q, Z, D[:n, :P + Q], S[:n, :P + Q], N_DEPTH, H,
                                  nthreads=threads)
    t1 = time.perf_counter()
    delta = np.where((r["valid"] == 1) & np.isfinite(r["delta"]), r["delta"], -np.inf)
    alive = np.nonzero(r["valid"] == 1)[0]
    k = min(max(1, int(round(top_frac * n))), alive.size)
    top = alive[np.argsort(-delta[alive], kind="stable")[:k]]
    oracle.walk_rns_batch(P, Q, Z, D[top, :P + Q], S[top, :P + Q], N_DEPTH, H, n_primes,
                          nthreads=threads)
    t2 = time.perf_counter()
    return {"value": n / (t2 - t0), "unit": "traj/s", "cores": threads, "kind": kind,
            "sample": f"first {n} of the 4M candidates: f64 walk of all ({t1 - t0:.2f}s, "
                      f"{'reference cmf_pfq.cuh' if kind == 'reference' else 'restatement'}) "
                      f"+ RNS of the top {k} by delta ({t2 - t1:.2f}s, restatement)",
            "survivors": int(alive.size)}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1 "
                                   "(trajectories shard by range across GPUs, see DESIGN.md)"}))
        sys.exit(2)
    import numpy as np
    import torch
    import cmfwalk as cw

    cw.load()
    t_setup = time.perf_counter()
    pipe = Pipeline(torch, cw, args.n_cand, args.top_frac, args.n_primes)
    print(f"setup {time.perf_counter() - t_setup:.1f}s: {args.n_cand} candidates resident",
          file=sys.stderr, flush=True)
    for i in range(args.warmup):
        pipe.run()
        torch.cuda.synchronize()
        print(f"warmup {i} done", file=sys.stderr, flush=True)
    ev = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        out = pipe.run(ev)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    names = ["t0", "screen", "compact", "deep", "select", "verify"]
    stages = [stage_ms(ev, names, i) for i in range(args.steps)]
    avg = {k: sum(s[k] for s in stages) / len(stages) for k in stages[0]}
    print(f"stages (ms, avg of {args.steps}): " + json.dumps({k: round(v, 3) for k, v in avg.items()}),
          file=sys.stderr, flush=True)

    res = cw.results_to_numpy(out["res"])
    rns = cw.results_to_numpy(out["rns"])
    fwd_f, inv_f = canonical_flops(pipe.desc.rank, P, Q)
    deep_flops = float(res["n_fwd"].astype(np.float64).sum() * fwd_f +
                       res["n_inv"].astype(np.float64).sum() * inv_f)
    pre = cw.results_to_numpy(pipe.pre)
    screen_flops = float(pre["n_fwd"].astype(np.float64).sum() * fwd_f +
                         pre["n_inv"].astype(np.float64).sum() * inv_f)
    dominant = max(("screen", "deep", "verify"), key=lambda k: avg[k])
    ms = elapsed / args.steps * 1e3
    value = args.n_cand / (elapsed / args.steps)
    deep_tflops = deep_flops / (avg["deep"] * 1e-3) / 1e12 if avg["deep"] > 0 else 0.0
    roofline = {"kernel": "walk_f64 (deep)", "bound": "valu", "achieved": deep_tflops,
                "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": deep_tflops / FP64_PEAK_TFLOPS, "traffic": None,
                "algorithmic_flops_per_launch": deep_flops,
                "launch_ms": avg["deep"], "dominant_stage": dominant}
    line = {
        "metric": "verified-delta traj/s, 8F7 z=1 N=1000: proposal+2xFP32 vs plain f64",
        "value": value, "unit": "traj/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "8F7 z=1 N=1000 h=50, zeta start, dirs {0..4}^15, "
                               f"{args.n_cand} candidates, top {args.top_frac:.0%} RNS-verified "
                               f"({args.n_primes} moduli)",
                   "pipeline": "plain-f64 (screen, deep f64 walk, RNS verify)",
                   "parallelism": "single-gpu"},
        "stages_ms": {k: round(v, 3) for k, v in avg.items()},
        "survivors": int(out["live"].numel()),
        "finite_delta": int(((res["valid"] == 1) & np.isfinite(res["delta"])).sum()),
        "verified": int(out["top"].numel()),
        "rns_valid": int(rns["valid"].sum()),
        "screen_flops": screen_flops,
        "roofline": roofline,
    }
    if not args.no_cpu_baseline:
        print("cpu baseline ...", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(pipe.host_D, pipe.host_S, args.cpu_sample,
                                            args.top_frac, args.n_primes)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line))


if __name__ == "__main__":
    main()
