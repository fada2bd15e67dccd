"""The learned sizer's job stream (rows and distinct keys log-uniform, key
power law s in [0, 1.5], i32 / i64 / 98-bit composite keys, null keys up to
20%): every job under method "auto" without and with the shipped model,
bit for bit against the oracle, and one job of 1e8 rows through
size-independent properties."""
import numpy as np
import pytest

import oracle as orc
from helpers import assert_results_equal, run_gpu
from polars_metal import synth

pytestmark = pytest.mark.gpu

SPECS = synth.sizer_job_specs(60, seed=2027, rows_lo=1e3, rows_hi=2e6)


def test_job_sample_covers_the_stream():
    dtypes = {s["dtype"] for s in SPECS}
    assert dtypes == {"I32", "I64", "U128"}
    assert min(s["s"] for s in SPECS) < 0.2 and max(s["s"] for s in SPECS) > 1.3
    assert max(s["null_frac"] for s in SPECS) > 0.15 and max(s["rows"] for s in SPECS) > 1e6
    assert min(s["distinct"] for s in SPECS) < 10 and max(s["distinct"] / s["rows"] for s in SPECS) > 0.5


@pytest.mark.parametrize("chunk", range(6))
def test_sizer_jobs_with_and_without_model(native, chunk):
    from polars_metal import learned
    for spec in SPECS[chunk * 10:(chunk + 1) * 10]:
        frame, plan = synth.sizer_job(spec)
        keys = [k[0] for k in plan["keys"]]
        print(f"sizer job {spec}", flush=True)
        want = orc.groupby(plan, frame)
        p = dict(plan, options={"method": "auto"})
        learned.install(None)
        got = run_gpu(native, p, frame)
        assert native.last_stats()["sizer_choice"] is None
        assert_results_equal(got, want, keys, label=f"plain/{spec['seed']}")
        learned.install_shipped()
        try:
            got = run_gpu(native, p, frame)
            assert native.last_stats()["sizer_choice"] in ("hash", "sort")
        finally:
            learned.install(None)
        assert_results_equal(got, want, keys, label=f"learned/{spec['seed']}")


def test_sizer_job_at_1e8_rows(native):
    """1e8 rows, about 2e6 distinct 98-bit composite keys with nulls: the
    group count equals the distinct keys numpy.unique counts and the count
    and length totals are conserved, with and without the model."""
    from polars_metal import learned
    from polars_metal.buffer import device_columns
    spec = {"rows": 100_000_000, "distinct": 2_000_000, "s": 0.7, "dtype": "U128", "null_frac": 0.1, "seed": 99}
    print("sizer job 1e8 rows", flush=True)
    frame, plan = synth.sizer_job(spec)
    # k and k2 are both bijections of the row's rank, so a non-null k fixes
    # k2: the groups are the distinct non-null k plus, for the null k rows,
    # the distinct k2
    k, kv, k2 = frame["k"][1], frame["k"][2], frame["k2"][1]
    n_groups = len(np.unique(k[kv])) + len(np.unique(k2[~kv]))
    cols, bufs = device_columns(frame)
    for model in (False, True):
        learned.install_shipped() if model else learned.install(None)
        try:
            r = native.execute_filter_groupby_device(dict(plan, options={"method": "auto"}), spec["rows"], cols)
        finally:
            learned.install(None)
        assert r.n_rows == n_groups, (model, r.n_rows, n_groups)
        lens = np.frombuffer(r.fetch("n")[1], np.uint32)[:r.n_rows]
        cnts = np.frombuffer(r.fetch("v_count")[1], np.uint32)[:r.n_rows]
        sums = np.frombuffer(r.fetch("v_sum")[1], np.float64)[:r.n_rows]
        assert int(lens.sum()) == spec["rows"] and int(cnts.sum()) == spec["rows"] and lens.min() >= 1
        # every v is in [0, 1): each group's sum lies in [0, len)
        assert (sums >= 0).all() and (sums < lens).all()
        r.free()


def test_stream_of_jobs_batched_with_the_model(native):
    """stream.run: jobs of one plan template go through one batched launch
    (98-bit keys alone), with the shipped model installed; every job
    against the oracle, bit for bit, in job order."""
    from polars_metal import stream
    from polars_metal.buffer import device_columns
    specs = synth.sizer_job_specs(36, seed=2029, rows_lo=1e3, rows_hi=3e5)
    jobs, frames, bufs = [], [], []
    for i, spec in enumerate(specs):
        frame, plan = synth.sizer_job(spec)
        if i % 4 == 0:  # a predicate on some jobs: the batch takes each query's own
            plan = dict(plan, predicate=synth.pred_gt_f64("v", 0.25))
        cols, b = device_columns(frame)
        jobs.append((plan, spec["rows"], cols))
        frames.append(frame)
        bufs.append(b)
    from test_batch import _fetch
    seen = []

    def check(i, r):
        plan, n, _ = jobs[i]
        assert_results_equal(_fetch(r), orc.groupby(plan, frames[i]), [k[0] for k in plan["keys"]], label=f"stream{i}")
        seen.append(i)

    info = stream.run(jobs, consume=check, tune=False)
    assert info["batched"] >= 20 and info["single"] >= 1 and info["batches"] == 2
    assert sorted(seen) == list(range(len(jobs)))
    seen.clear()
    stream.run_plain(jobs, consume=check)
    assert sorted(seen) == list(range(len(jobs)))
