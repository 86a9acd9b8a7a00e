"""The benchmark runner writes the reference runner's files and columns
(benchmarks/run_benchmark.py:96-172) plus roofline columns.  The table
writer runs on the CPU; the end-to-end run (executables on the GPU) is
marked gpu."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ramanujandreams-cuda_amd")

REFERENCE_CSV = ("cmf,variant,mode,rank,N,h,n_traj,gpu_time_ms,throughput_traj_s,mem_mb,"
                 "n_valid,n_with_limits,avg_delta,wall_time_s")


def test_tables_keep_reference_columns(tmp_path):
    from cmfwalk import run_benchmark as rb

    rows = [{"cmf": "3F2", "variant": "f64", "mode": "limit", "rank": 3, "N": 100, "h": 10,
             "n_traj": 4, "gpu_time_ms": 1.5, "throughput_traj_per_s": 2666.7,
             "mem_used_mb": 10.0, "n_valid": 4, "n_with_limits": 4, "avg_delta": None,
             "wall_time_s": 0.2, "n_fwd_steps": 1200, "n_inv_steps": 800, "work_unit": "flop",
             "canonical_work": 1.0e6, "achieved_per_s": 6.7e8, "roofline_frac": 8.5e-6},
            {"cmf": "3F2", "variant": "rns", "mode": "delta", "rank": 3, "N": 100, "h": 10,
             "n_traj": 4, "gpu_time_ms": 2.0, "throughput_traj_per_s": 2000.0,
             "mem_used_mb": 10.0, "n_valid": 0, "n_with_limits": 0, "avg_delta": None,
             "wall_time_s": 0.3}]
    rb.write_tables(rows, tmp_path)
    lines = (tmp_path / "benchmark_table.csv").read_text().splitlines()
    assert lines[0].startswith(REFERENCE_CSV + ",")
    assert len(lines) == 3
    assert lines[1].split(",")[:7] == ["3F2", "f64", "limit", "3", "100", "10", "4"]
    md = (tmp_path / "benchmark_table.md").read_text()
    assert "| 3F2 | f64 | limit | 3 |" in md and "| 4/4 |" in md


def test_fast_variant_runs_f64_binary_with_fast_flag(tmp_path, monkeypatch):
    from cmfwalk import run_benchmark as rb

    calls = []

    def fake_run(cmd, capture_output, text):
        calls.append(cmd)
        out = cmd[cmd.index("--output") + 1]
        with open(out, "w") as f:
            json.dump({"gpu_time_ms": 1.0, "throughput_traj_per_s": 1.0}, f)
        return subprocess.CompletedProcess(cmd, 0, "", "")

    monkeypatch.setattr(rb.subprocess, "run", fake_run)
    bins = tmp_path / "bin"
    bins.mkdir()
    for v in ("f64", "rns", "df64"):
        (bins / f"cmf_walk_{v}").write_text("")
    assert rb.main(["--bin-dir", str(bins), "--output-dir", str(tmp_path / "o"), "--cmfs", "3F2",
                    "--variants", "f64", "f64_fast", "--modes", "limit"]) == 0
    assert [os.path.basename(c[0]) for c in calls] == ["cmf_walk_f64", "cmf_walk_f64"]
    assert "--fast" not in calls[0] and "--fast" in calls[1]
    res = json.load(open(tmp_path / "o" / "summary" / "all_results.json"))
    assert [r["variant"] for r in res] == ["f64", "f64_fast"]


@pytest.mark.gpu
def test_runner_end_to_end(tmp_path):
    r = subprocess.run([sys.executable, "-m", "cmfwalk.run_benchmark", "--output-dir",
                        str(tmp_path), "--cmfs", "3F2", "6F5", "--N", "200", "--h", "50",
                        "--n-traj", "2000", "--variants", "f64", "f64_fast", "rns", "df64"],
                       cwd=PKG, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    res = json.load(open(tmp_path / "summary" / "all_results.json"))
    assert len(res) == 2 * 4 * 2
    for x in res:
        assert {"cmf", "variant", "mode", "rank", "N", "h", "n_traj", "gpu_time_ms",
                "throughput_traj_per_s", "mem_used_mb", "n_valid", "n_with_limits",
                "avg_delta", "wall_time_s", "canonical_work", "roofline_frac"} <= set(x)
        assert (tmp_path / "individual" / f"{x['cmf']}_{x['variant']}_{x['mode']}.json").exists()
        assert x["canonical_work"] > 0 and 0 < x["roofline_frac"] < 1
    assert (tmp_path / "summary" / "benchmark_table.csv").read_text().startswith(REFERENCE_CSV)
