"""Extract the reference's published per-run records into a small fixture.

Run in a container where the reference tree is mounted (it is never read at
test time).  Reads results/{zeta,random}/summary/all_results.json and writes
tests/golden/published_records.json with, for every published run, its
summary fields and (f64 runs only) the per-trajectory records it holds.
"""
import json
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "published_records.json")
FIELDS = ["idx", "valid", "limit", "delta", "p", "q", "precision", "log_scale", "n_limits"]


def main():
    out = []
    for suite in ["zeta", "random"]:
        path = os.path.join(REF, "results", suite, "summary", "all_results.json")
        for run in json.load(open(path)):
            rec = {k: v for k, v in run.items() if k != "results"}
            rec["suite"] = suite
            rec["source"] = f"results/{suite}/summary/all_results.json"
            res = run.get("results") or []
            if res:
                rec["records"] = {f: [r.get(f) for r in res] for f in FIELDS}
            out.append(rec)
    with open(OUT, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(f"wrote {OUT}: {len(out)} runs")


if __name__ == "__main__":
    main()
