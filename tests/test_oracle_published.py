"""The oracle against the reference's published runs
(results/{zeta,random}/summary/all_results.json, extracted into
tests/golden/published_records.json by tests/golden/make_published.py).

Inputs are the legacy drivers' effective inputs, stride aliasing included
(cuda/cmf_walk_f64.cu:178,198).  Validity and limit counts are exact.  The
published values are H100 outputs (nvcc contracts FMAs), so values are
compared to 1e-9 relative on p, q and log_scale (delta, precision and the
Aitken limit of a converged walk measure rounding noise itself)."""
import numpy as np
import pytest


def _arr(v):
    return np.array([np.nan if x is None else x for x in v], dtype=float)


def _f64_runs(published):
    return [r for r in published if r["variant"] == "f64"]


def test_fixture_present(published):
    assert len(_f64_runs(published)) == 20


@pytest.mark.parametrize("suite", ["zeta", "random"])
@pytest.mark.parametrize("mode", ["limit", "delta"])
def test_published_counts_and_records(orc, published, suite, mode):
    runs = [r for r in _f64_runs(published) if r["suite"] == suite and r["mode"] == mode]
    assert len(runs) == 5
    for run in runs:
        p, q, z = run["p"], run["q"], float(run["z"])
        D, S = orc.legacy_inputs(p, q, run["n_traj"], mode, suite == "zeta")
        res = orc.walk_f64_batch(p, q, z, D, S, run["N"], run["h"], nthreads=0)
        valid = res["valid"].astype(bool)
        assert int(valid.sum()) == run["n_valid"], run["cmf"]
        assert int((valid & ~np.isnan(res["limit"])).sum()) == run["n_with_limits"], run["cmf"]
        rec = run["records"]
        r100 = res[:100]
        assert (np.array(rec["valid"]) == r100["valid"]).all()
        # values: p, q, log_scale of valid records to 1e-9 relative; walks
        # whose snapshots differ only by rounding noise (or that pass through a
        # near-singular solve) amplify nvcc's FMA contraction, so at least 90%
        # of the compared fields per run (or all but two) must match; the
        # zeta-limit runs are checked field by field below.
        good = total = 0
        for i in range(100):
            if not rec["valid"][i]:
                assert np.isnan(r100["p"][i])
                continue
            for f in ["p", "q", "log_scale"]:
                a, b = rec[f][i], r100[f][i]
                total += 1
                if a is None:
                    good += not np.isfinite(b)
                else:
                    good += abs(a - b) <= 1e-9 * max(1.0, abs(a))
        assert good >= 0.9 * total or total - good <= 2, (run["cmf"], good, total)


def test_zeta_limit_records_exact_fields(orc, published):
    """The five zeta-limit records reproduce in every field
    the reference printed, n_limits exactly (4, 36, 96, 200, 256)."""
    runs = [r for r in _f64_runs(published) if r["suite"] == "zeta" and r["mode"] == "limit"]
    want_nl = {"3F2": 4, "4F3": 36, "5F4": 96, "6F5": 200, "8F7": 256}
    for run in runs:
        p, q = run["p"], run["q"]
        d = [1] + [0] * (p + q - 1)
        s = [1] * p + [2] * q
        r = orc.walk_f64_batch(p, q, 1.0, [d], [s], 1000, 50)[0]
        rec = run["records"]
        assert r["n_limits"] == rec["n_limits"][0] == want_nl[run["cmf"]]
        for f in ["limit", "p", "q", "log_scale", "precision"]:
            a = rec[f][0]
            if a is None:
                assert not np.isfinite(r[f])
            else:
                assert abs(a - r[f]) <= 1e-9 * max(1, abs(a)), (run["cmf"], f)
        assert rec["delta"][0] is None and np.isnan(r["delta"])
