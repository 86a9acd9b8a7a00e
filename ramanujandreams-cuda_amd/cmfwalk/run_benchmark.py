#!/usr/bin/env python3
"""Benchmark matrix of the MI355X walk engine, in the reference runner's
format (benchmarks/run_benchmark.py): CMFs x variants (f64, rns, df64) x
modes (limit, delta), each by one run of this package's cmf_walk_<variant>
executable, which keeps the reference's flags and JSON keys.

Outputs, as the reference's runner writes them
(benchmarks/run_benchmark.py:96-172):
  <out>/individual/<cmf>_<variant>_<mode>.json   one run's JSON
  <out>/summary/all_results.json                 every run, plus wall_time_s
  <out>/summary/benchmark_table.csv              the reference's columns, then
                                                 the roofline columns
  <out>/summary/benchmark_table.md               the same as a table
The roofline columns come from the additional keys of the executables'
JSON: executed forward / inverse axis steps, the algorithmic work they stand
for (flops of the reference recurrence, or modular products for RNS), the
achieved rate and its fraction of the MI355X peak.

python -m cmfwalk.run_benchmark [--bin-dir DIR] [--output-dir DIR] [--N 1000]
    [--h 50] [--n-traj 10000] [--z 0.5] [--cmfs ...] [--variants ...]
    [--modes ...] [--n-primes 8] [--zeta]

Variant "f64_fast" (not in the reference) runs cmf_walk_f64 --fast, the
fast FP64 engine; its rows are labelled f64_fast.
"""
import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

CMF_TYPES = ["3F2", "4F3", "5F4", "6F5", "8F7"]
VARIANTS = ["f64", "rns", "df64"]
MODES = ["limit", "delta"]
HERE = Path(__file__).resolve().parent
DEFAULT_BIN = HERE.parent / "bin"

CSV_COLUMNS = ("cmf,variant,mode,rank,N,h,n_traj,gpu_time_ms,throughput_traj_s,"
               "mem_mb,n_valid,n_with_limits,avg_delta,wall_time_s")
ROOFLINE_COLUMNS = "n_fwd_steps,n_inv_steps,work_unit,canonical_work,achieved_per_s,roofline_frac"


def run_one(binary, cmf, z, N, h, n_traj, mode, output_path, n_primes=None, use_zeta=False,
            fast=False):
    """One executable run; its JSON plus wall_time_s, or None on failure."""
    cmd = [str(binary), "--cmf", cmf, "--z", str(z), "--N", str(N), "--h", str(h),
           "--n-traj", str(n_traj), "--mode", mode, "--output", str(output_path)]
    if fast:
        cmd.append("--fast")
    if n_primes and "rns" in Path(binary).name:
        cmd += ["--n-primes", str(n_primes)]
    if use_zeta:
        cmd.append("--zeta")
    print(f"  Running: {' '.join(cmd)}", flush=True)
    t0 = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    elapsed = time.time() - t0
    if r.returncode != 0:
        print(f"  ERROR: {r.stderr}", flush=True)
        return None
    try:
        with open(output_path) as f:
            data = json.load(f)
    except (OSError, ValueError) as e:
        print(f"  ERROR parsing output: {e}", flush=True)
        return None
    data["wall_time_s"] = elapsed
    return data


def _num(v, fmt):
    return format(v, fmt) if isinstance(v, (int, float)) else "nan"


def write_tables(results, out_dir):
    """benchmark_table.csv and benchmark_table.md (the reference's columns,
    then the roofline columns)."""
    out_dir = Path(out_dir)
    with open(out_dir / "benchmark_table.csv", "w") as f:
        f.write(CSV_COLUMNS + "," + ROOFLINE_COLUMNS + "\n")
        for r in results:
            f.write(f"{r.get('cmf', '')},{r.get('variant', '')},{r.get('mode', '')},"
                    f"{r.get('rank', 0)},{r.get('N', 0)},{r.get('h', 0)},"
                    f"{r.get('n_traj', 0)},{_num(r.get('gpu_time_ms', 0), '.3f')},"
                    f"{_num(r.get('throughput_traj_per_s', 0), '.1f')},"
                    f"{_num(r.get('mem_used_mb', 0), '.2f')},"
                    f"{r.get('n_valid', 0)},{r.get('n_with_limits', 0)},"
                    f"{r.get('avg_delta', 'nan')},{_num(r.get('wall_time_s', 0), '.3f')},"
                    f"{_num(r.get('n_fwd_steps'), '.0f')},{_num(r.get('n_inv_steps'), '.0f')},"
                    f"{r.get('work_unit', '')},{_num(r.get('canonical_work'), '.6g')},"
                    f"{_num(r.get('achieved_per_s'), '.6g')},"
                    f"{_num(r.get('roofline_frac'), '.4g')}\n")
    with open(out_dir / "benchmark_table.md", "w") as f:
        f.write("# MI355X CMF walk engine: benchmark results\n\n")
        f.write("| CMF | Variant | Mode | Rank | GPU Time (ms) | Throughput (traj/s) | Memory (MB) "
                "| Valid | Limits | Avg Delta | Work | Achieved/s | Roofline |\n")
        f.write("|-----|---------|------|------|---------------|---------------------|-------------"
                "|-------|--------|-----------|------|------------|----------|\n")
        for r in results:
            f.write(f"| {r.get('cmf', '')} | {r.get('variant', '')} | {r.get('mode', '')} "
                    f"| {r.get('rank', 0)} | {_num(r.get('gpu_time_ms', 0), '.1f')} "
                    f"| {_num(r.get('throughput_traj_per_s', 0), '.0f')} "
                    f"| {_num(r.get('mem_used_mb', 0), '.1f')} "
                    f"| {r.get('n_valid', 0)}/{r.get('n_traj', 0)} "
                    f"| {r.get('n_with_limits', 0)} | {r.get('avg_delta', 'nan')} "
                    f"| {_num(r.get('canonical_work'), '.3g')} {r.get('work_unit', '')} "
                    f"| {_num(r.get('achieved_per_s'), '.3g')} "
                    f"| {_num(r.get('roofline_frac'), '.2%')} |\n")


def main(argv=None):
    ap = argparse.ArgumentParser(description="MI355X CMF walk engine benchmark runner")
    ap.add_argument("--bin-dir", type=Path, default=DEFAULT_BIN)
    ap.add_argument("--output-dir", type=Path, default=Path("results"))
    ap.add_argument("--N", type=int, default=1000)
    ap.add_argument("--h", type=int, default=50)
    ap.add_argument("--n-traj", type=int, default=10000)
    ap.add_argument("--z", type=float, default=0.5)
    ap.add_argument("--cmfs", nargs="+", default=CMF_TYPES)
    ap.add_argument("--variants", nargs="+", default=VARIANTS)
    ap.add_argument("--modes", nargs="+", default=MODES)
    ap.add_argument("--n-primes", type=int, default=8)
    ap.add_argument("--zeta", action="store_true")
    args = ap.parse_args(argv)

    (args.output_dir / "individual").mkdir(parents=True, exist_ok=True)
    (args.output_dir / "summary").mkdir(parents=True, exist_ok=True)
    results = []
    total = len(args.cmfs) * len(args.variants) * len(args.modes)
    k = 0
    for cmf in args.cmfs:
        for variant in args.variants:
            fast = variant == "f64_fast"
            binary = args.bin_dir / ("cmf_walk_f64" if fast else f"cmf_walk_{variant}")
            if not binary.exists():
                print(f"  SKIP: {binary} not found", flush=True)
                continue
            for mode in args.modes:
                k += 1
                print(f"\n[{k}/{total}] {cmf} / {variant} / {mode}", flush=True)
                out = args.output_dir / "individual" / f"{cmf}_{variant}_{mode}.json"
                r = run_one(binary, cmf, args.z, args.N, args.h, args.n_traj, mode, out,
                            n_primes=args.n_primes, use_zeta=args.zeta, fast=fast)
                if r:
                    r["cmf"], r["variant"], r["mode"] = cmf, variant, mode
                    results.append(r)
                    print(f"  -> {r.get('throughput_traj_per_s', 0):.0f} traj/s, "
                          f"{r.get('gpu_time_ms', 0):.1f} ms, roofline "
                          f"{r.get('roofline_frac') or 0:.2%}", flush=True)
    with open(args.output_dir / "summary" / "all_results.json", "w") as f:
        json.dump(results, f, indent=2)
    write_tables(results, args.output_dir / "summary")
    print(f"\nBenchmark complete. {len(results)} configurations run.", flush=True)
    return 0 if len(results) == k else 1


if __name__ == "__main__":
    sys.exit(main())
