"""Benchmark of the filtered group-by hot path on one MI355X.

Workload (one "step" = one fused call over one batch of 100M rows, inputs
already resident in HBM, outputs left in HBM):
  key i64 bounded Zipf(s=1.1) over ranks 1..1e6, v f64 N(0,1) with 5% nulls,
  w i32 U{0..999}; filter (w >= 200) & (w < 500) (~30% of rows);
  group_by(key).agg(sum(v), count(v), min(w), max(w), mean(v)).

Prints one JSON line with value = input rows per second, the roofline of the
dominant kernel (HBM-bound; algorithmic bytes = 20.125 B per input row:
8 key + 8 v + 4 w + 1/8 validity), the CPU baseline (the C/OpenMP oracle
on every host core, on the same 100M-row frame) and the learned sizer's
live speedup over the plain method on a bounded stream of fresh jobs.

Other workloads (one JSON line each, not the default):
  --workload highcard   200M rows, composite key (a i32, b i64) over ~49M
                        groups, sum/count of v i64 with 10% nulls: the hash and
                        the u128 radix-sort engines, checked to agree.
  --workload batch      4096 filtered group-by queries (rows log-uniform in
                        [1e3, 1e6]) as one batch in a captured graph, against
                        the same queries run one call at a time.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for p in (ROOT / "polars-metal_amd", ROOT / "oracle", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

HBM_PEAK_GBS = 8000.0
BYTES_PER_ROW = 8 + 8 + 4 + 0.125
METRIC = "filtered group-by-agg rows/s + % HBM peak at 100M rows; learned-vs-plain speedup"


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(frame: dict, plan: dict, rows: int, warmup: int = 2, reps: int = 7) -> dict:
    """The oracle's all-core group-by (C + OpenMP, sequential per-group f64
    sums) on the same frame the GPU runs: the median of ``reps`` runs after
    ``warmup``, query only (predicate, key encoding, group-by)."""
    import statistics

    import oracle as orc
    orc.build()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    if rows < orc.frame_rows(frame):
        frame = {k: (t, v[:rows], None if ok is None else ok[:rows]) for k, (t, v, ok) in frame.items()}
    times = []
    for i in range(warmup + reps):
        t0 = time.perf_counter()
        res = orc.groupby(plan, frame, sum_mode=1, threads=threads)
        if i >= warmup:
            times.append(time.perf_counter() - t0)
    dt = statistics.median(times)
    return {"value": rows / dt, "unit": "rows/s", "cores": threads, "kind": "port",
            "sample": f"C++/OpenMP oracle (not Polars, which is absent) on all {threads} threads of the host "
                      f"({_cpu_model()}): the same {rows}-row zipf frame, median of {reps} after {warmup} "
                      f"warm-ups, {res['__n_groups__']} groups, {dt * 1e3:.0f} ms per query"}


# Row-scan kernels of the engines, in order of preference for the roofline:
# the scan reads every input byte once and dominates the step.
SCAN_KERNELS = ("hs_scan_fast", "hs_scan", "fgb_main_narrow")


def traffic_from_profiles(kernel: str) -> float | None:
    """HBM bytes per launch of the scan kernel from the committed rocprofv3
    PMC summary, when one exists."""
    f = ROOT / "profiles" / f"{kernel}_zipf100m_traffic.json"
    if not f.exists():
        return None
    try:
        return float(json.loads(f.read_text())["hbm_bytes_per_launch"])
    except Exception:
        return None


def log(msg: str) -> None:
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def stream_ab(n_jobs: int, seed: int = 777, rows_hi: float = 1e7) -> dict:
    """Learned sizing plus batching against the plain path on one stream of
    fresh jobs from the sizer's training distribution (another seed): the
    plain leg runs one call per job under method "auto" without a model;
    the other runs stream.run (jobs of one plan template batched into one
    captured launch, the rest alone, the shipped model installed).  Each leg
    is warmed once and then timed end to end; every job's groups are
    checked to be the same in both."""
    from polars_metal import _native, learned, stream, synth
    from polars_metal.buffer import device_columns
    specs = synth.sizer_job_specs(n_jobs, seed, rows_hi=rows_hi)
    meta = learned.install_shipped()
    learned.install(None)
    jobs, bufs, names = [], [], []
    for spec in specs:
        frame, plan = synth.sizer_job(spec)
        cols, b = device_columns(frame)
        jobs.append((plan, spec["rows"], cols))
        bufs.append(b)
        names.append([k[0] for k in plan["keys"]] + [a["output_alias"] for a in plan["aggs"]])
    n_rows = sum(s_["rows"] for s_ in specs)
    # warm both legs, then time each end to end; a device consumer would use
    # each result in place, so the timed legs hand results over and free them
    stream.run_plain(jobs)
    stream.run(jobs)
    t0 = time.perf_counter()
    stream.run_plain(jobs)
    t_plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    info = stream.run(jobs)
    t_fast = time.perf_counter() - t0
    # untimed: every job's groups, by an order-independent checksum, in both legs
    sums = {"plain": {}, "fast": {}}
    stream.run_plain(jobs, consume=lambda i, r: sums["plain"].__setitem__(i, group_checksum(r, names[i])))
    stream.run(jobs, consume=lambda i, r: sums["fast"].__setitem__(i, group_checksum(r, names[i])))
    if sums["plain"] != sums["fast"] or len(sums["plain"]) != n_jobs:
        raise AssertionError("stream: results differ between the plain and the batched legs")
    # the sizer alone: the same jobs one call each, with the model
    learned.install_shipped()
    try:
        t0 = time.perf_counter()
        picks = {"hash": 0, "sort": 0}
        for plan, n, cols in jobs:
            r = _native.execute_filter_groupby_device(dict(plan, options={"method": "auto"}), n, cols)
            picks[r.stats["method"]] += 1
            r.free()
        t_sizer = time.perf_counter() - t0
    finally:
        learned.install(None)
    del bufs
    return {"speedup": t_plain / t_fast, "jobs": n_jobs, "rows": n_rows, "plain_ms": t_plain * 1e3,
            "learned_batched_ms": t_fast * 1e3, "batched_jobs": info["batched"], "single_jobs": info["single"],
            "batches": info["batches"], "sizer_only_speedup": t_plain / t_sizer, "sizer_picks": picks,
            "sample": f"{n_jobs} fresh jobs (seed {seed}): rows log-uniform [1e3, {rows_hi:.0e}], distinct log-uniform "
                      f"[1, rows], key power law s~U[0,1.5], i32/i64/98-bit keys, nulls U[0,0.2]; wall time of the "
                      f"whole stream after one warm run of each leg; plain = one call per job, method auto, no "
                      f"model; learned = jobs of one plan template in one batched launch (98-bit keys alone), "
                      f"shipped model installed; every job's groups checked equal",
            "model_held_out_gain": meta.get("report", {}).get("held_out", {}).get("gain")}


def pinned_h2d_gbps(nbytes: int, reps: int = 3) -> float:
    """The PCIe ceiling: hipMemcpy from one pinned host buffer of nbytes to
    the device (best of reps), called on the HIP runtime directly."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    h, d = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipHostMalloc(ctypes.byref(h), ctypes.c_size_t(nbytes), 0) or hip.hipMalloc(ctypes.byref(d), ctypes.c_size_t(nbytes)):
        raise RuntimeError("pinned_h2d_gbps: allocation failed")
    try:
        ctypes.memset(h, 1, nbytes)
        best = float("inf")
        for _ in range(reps + 1):
            t = time.perf_counter()
            if hip.hipMemcpy(d, h, ctypes.c_size_t(nbytes), 1) or hip.hipDeviceSynchronize():  # 1: host to device
                raise RuntimeError("pinned_h2d_gbps: copy failed")
            best = min(best, time.perf_counter() - t)
        return nbytes / best / 1e9
    finally:
        hip.hipFree(d)
        hip.hipHostFree(h)


def host_path(frame: dict, plan: dict, n: int, reps: int = 2) -> dict:
    """The drop-in path through host buffers: execute_filter_groupby with
    Arrow bytes in (pageable) host memory, so the input crosses PCIe inside
    the call, through the library's pinned staging ring; the same query
    through the engine hook's Arrow dispatch with the device column cache
    (first call uploads, second finds every column resident); and the
    pinned-DMA ceiling the ring is measured against."""
    import numpy as np
    import pyarrow as pa
    from polars_metal import _native, colcache, udf
    cols = []
    for name, (tag, values, valid) in frame.items():
        vb = np.packbits(np.asarray(valid, bool), bitorder="little").tobytes() if valid is not None else None
        cols.append((name, tag, np.ascontiguousarray(values).tobytes(), vb))
    _native.execute_filter_groupby(plan, n, cols)
    t0 = time.perf_counter()
    for _ in range(reps):
        out = _native.execute_filter_groupby(plan, n, cols)
    dt = (time.perf_counter() - t0) / reps
    in_bytes = sum(len(c[2]) + (len(c[3]) if c[3] else 0) for c in cols)
    del cols
    gbps = pinned_h2d_gbps(1 << 30)
    res = {"ms_per_query": dt * 1e3, "rows_per_s": n / dt, "input_bytes": in_bytes,
           "groups": len(out[0][2]) // 8, "pinned_h2d_gbps": gbps,
           "pinned_floor_ms": in_bytes / gbps / 1e6,
           "note": "host (pageable) Arrow buffers in, host bytes out; PCIe copies inside the timed call, through "
                   "the pinned staging ring; pinned_floor_ms = input bytes / one pinned 1 GiB hipMemcpy's rate"}
    table = pa.table({name: pa.array(values, mask=None if valid is None else ~np.asarray(valid, bool))
                      for name, (tag, values, valid) in frame.items()})
    wire = {"kind": "GroupBy", "input": {"kind": "Scan", "selection": plan["predicate"]}, "keys": plan["keys"],
            "aggs": plan["aggs"]}
    colcache.clear()
    h0 = colcache.stats()["h2d_bytes"]
    t = time.perf_counter()
    udf.dispatch_groupby(table, wire, cache=True)
    t_first = time.perf_counter() - t
    h1 = colcache.stats()["h2d_bytes"]
    t = time.perf_counter()
    for _ in range(reps):
        udf.dispatch_groupby(table, wire, cache=True)
    t_second = (time.perf_counter() - t) / reps
    h2 = colcache.stats()["h2d_bytes"]
    colcache.clear()
    res["column_cache"] = {"first_ms": t_first * 1e3, "resident_ms": t_second * 1e3, "first_h2d_bytes": h1 - h0,
                           "resident_h2d_bytes": (h2 - h1) // reps,
                           "note": "udf.dispatch_groupby over a pyarrow table, cache=True; times include the "
                                   "Arrow result assembly"}
    return res


def group_checksum(res, names) -> tuple:
    """Order-independent checksum of a device result's groups: the wrapping
    sum over groups of a mix of the named columns' bits and validity."""
    import numpy as np
    ng = res.n_rows
    acc = np.zeros(ng, np.uint64)
    with np.errstate(over="ignore"):
        for k, name in enumerate(names):
            tag, data, valid = res.fetch(name)
            w = {"I64": 8, "F64": 8, "U64": 8, "I32": 4, "U32": 4}[tag]
            x = np.frombuffer(data, {8: np.uint64, 4: np.uint32}[w], ng).astype(np.uint64)
            if valid is not None:
                ok = np.unpackbits(np.frombuffer(valid, np.uint8), bitorder="little")[:ng].astype(np.uint64)
                x = x * ok + (ok ^ np.uint64(1)) * np.uint64(0x5BD1E995)
            acc = (acc ^ x) * np.uint64(0x9E3779B97F4A7C15) + np.uint64(k + 1)
            acc ^= acc >> np.uint64(29)
    return ng, int(acc.sum(dtype=np.uint64))


def bench_highcard(args) -> dict:
    """Both engines on the 200M-row composite-key workload."""
    from polars_metal import _native, synth
    from polars_metal.buffer import device_columns
    n = 200_000_000
    log(f"generating {n} rows (numpy, seed 3; 50M-pair pool) and uploading")
    frame, plan = synth.highcard(n, seed=3)
    cols, bufs = device_columns(frame)
    del frame
    names = [k[0] for k in plan["keys"]] + [a["output_alias"] for a in plan["aggs"]]
    engines = {}
    for method in ("hash", "sort"):
        p = dict(plan, options={"method": method})
        for _ in range(args.warmup):
            _native.execute_filter_groupby_device(p, n, cols).free()
        _native.set_profiling(True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            r = _native.execute_filter_groupby_device(p, n, cols)
            st = r.stats
            r.free()
        dt = (time.perf_counter() - t0) / args.steps
        kt = _native.kernel_times()
        _native.set_profiling(False)
        kernel_ms = sum(v[1] for v in kt.values()) / max(1, args.steps)
        r = _native.execute_filter_groupby_device(p, n, cols)
        chk = group_checksum(r, names)
        r.free()
        engines[method] = {"ms_per_query": dt * 1e3, "kernel_ms": kernel_ms, "groups": chk[0],
                           "checksum": chk[1], "method_run": st["method"]}
        log(f"{method}: {engines[method]}")
    agree = engines["hash"]["groups"] == engines["sort"]["groups"] and \
        engines["hash"]["checksum"] == engines["sort"]["checksum"]
    best = min(engines, key=lambda m: engines[m]["ms_per_query"])
    alg = (4 + 8 + 8 + 0.125) * n + engines[best]["groups"] * (4 + 8 + 8 + 4)
    achieved = alg / (engines[best]["kernel_ms"] * 1e-3) / 1e9
    return {
        "metric": "high-cardinality group-by rows/s (hash and u128 radix sort, bit-exact agreement)",
        "value": n / (engines[best]["ms_per_query"] * 1e-3), "unit": "rows/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": engines[best]["ms_per_query"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "i64",
        "data": "synthetic (numpy default_rng seed 3, uploaded before timing)",
        "config": {"workload": "highcard-200M: group_by([a i32 U[0,2^20), b i64 U[0,2^40)]) from a 50M-pair pool "
                               "(~49M groups).agg(sum(v), count(v)); v i64 U[-1e9,1e9] 10% nulls",
                   "rows": n, "engines": engines, "agree": agree, "faster": best},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": f"all kernels of the {best} engine",
                     "kernel_avg_ms": engines[best]["kernel_ms"], "algorithmic_bytes_per_launch": alg},
    }


def bench_batch(args) -> dict:
    """4096 queries batched into one captured graph vs one call each."""
    from polars_metal import _native, batch, synth
    from polars_metal.buffer import device_columns
    specs = synth.batch_specs(4096)
    log("generating and uploading 4096 queries (numpy, seed 5)")
    queries, bufs, plan = [], [], None
    for s_ in specs:
        f, p = synth.batch_query(s_)
        cols, b = device_columns(f)
        bufs.append(b)
        queries.append((s_["rows"], cols, p.get("predicate")))
        plan = {"keys": p["keys"], "aggs": p["aggs"]}
    rows = sum(s_["rows"] for s_ in specs)
    batch.sequential(plan, queries[:64])
    log("sequential pass over all queries")
    t0 = time.perf_counter()
    seq = batch.sequential(plan, queries)
    t_seq = time.perf_counter() - t0
    opts = batch.autotune(plan, queries)
    log(f"autotuned launch shape {opts}")
    for _ in range(args.warmup):
        res, st = batch.execute_batch(plan, queries, opts)
        for r in res:
            r.free()
    _native.set_profiling(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res, st = batch.execute_batch(plan, queries, opts)
        groups = [r.n_rows for r in res]
        for r in res:
            r.free()
    dt = (time.perf_counter() - t0) / args.steps
    kt = _native.kernel_times()
    _native.set_profiling(False)
    if groups != [q["n_groups"] for q in seq]:
        raise AssertionError("batch and sequential group counts differ")
    graph_ms = kt.get("bt_graph", (1, float("nan")))[1] / max(1, kt.get("bt_graph", (1, 0))[0])
    n_groups = sum(groups)
    alg = 16.0 * rows + 44.0 * n_groups  # key + v in; key, sum, count, min, max, mean out
    achieved = alg / (graph_ms * 1e-3) / 1e9
    return {
        "metric": "batched filtered group-by: queries/s of 4096 queries in one captured graph vs sequential calls",
        "value": 4096 / dt, "unit": "queries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+i64", "data": "synthetic (numpy default_rng, seeds 5*100003+i, uploaded before timing)",
        "config": {"workload": "batch-4096: rows log-uniform [1e3,1e6] per query, key i64 U{1..rows/10}, v f64 "
                               "U[0,1), 20% of queries filter v>0.5; sum/count/min/max/mean of v",
                   "rows": rows, "groups": n_groups, "sequential_ms": t_seq * 1e3,
                   "speedup_vs_sequential": t_seq / dt, "rows_per_s": rows / dt, "launch_shape": opts,
                   "device_ms": st["device_ms"], "n_rerun": st["n_rerun"],
                   "n_fallback_keys": st["n_fallback_keys"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "bt_graph (the captured chain)",
                     "kernel_avg_ms": graph_ms, "algorithmic_bytes_per_launch": alg},
        "kernel_ms": {k: v[1] / v[0] for k, v in kt.items() if v[0]},
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=("zipf100m", "highcard", "batch"), default="zipf100m")
    ap.add_argument("--rows", type=int, default=100_000_000)
    ap.add_argument("--cpu-rows", type=int, default=100_000_000)
    ap.add_argument("--sizer-jobs", type=int, default=120)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-sizer", action="store_true")
    ap.add_argument("--no-host-path", action="store_true")
    ap.add_argument("--no-highcard", action="store_true", help="skip the 200M-row composite-key leg")
    ap.add_argument("--options", type=str, default="", help="JSON plan options (launch shape overrides)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU; the multi-GPU sharding is a design only "
              "(see the sharding section of the design notes)", file=sys.stderr)
        return 2
    if args.workload == "highcard":
        print(json.dumps(bench_highcard(args)), flush=True)
        return 0
    if args.workload == "batch":
        print(json.dumps(bench_batch(args)), flush=True)
        return 0

    from polars_metal import _native, synth
    from polars_metal.buffer import device_columns

    n = args.rows
    print(f"bench: generating {n} rows (numpy, seed 2) and uploading", file=sys.stderr, flush=True)
    frame, plan = synth.zipf(n, seed=2)
    cols, bufs = device_columns(frame)
    if args.options:
        plan["options"] = json.loads(args.options)

    # Every call synchronises its stream before returning (the group count is
    # read back), so host timestamps around the loop bracket the device work.
    # The timed steps run without kernel timers (their events add gaps
    # between kernels); a separate profiled pass gives per-kernel times.
    for _ in range(args.warmup):
        r = _native.execute_filter_groupby_device(plan, n, cols)
        r.free()
    print(f"bench: timing {args.steps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    stats = None
    for _ in range(args.steps):
        r = _native.execute_filter_groupby_device(plan, n, cols)
        stats = r.stats
        r.free()
    dt = time.perf_counter() - t0
    _native.set_profiling(True)
    for _ in range(min(args.steps, 5)):
        _native.execute_filter_groupby_device(plan, n, cols).free()
    kt = _native.kernel_times()
    _native.set_profiling(False)

    main_name = next((k for k in SCAN_KERNELS if kt.get(k, (0, 0.0))[0]), SCAN_KERNELS[0])
    launches, ms = kt.get(main_name, (0, 0.0))
    avg_ms = ms / launches if launches else float("nan")
    alg_bytes = BYTES_PER_ROW * n
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9 if launches else None
    ms_step = dt * 1e3 / args.steps
    out = {
        "metric": METRIC,
        "value": n * args.steps / dt,
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+i64",
        "data": "synthetic (numpy default_rng seed 2, uploaded before timing)",
        "config": {"workload": "zipf-100M filtered group-by: filter (w>=200)&(w<500) on 100M rows; "
                               "group_by(key i64 Zipf s=1.1 over 1e6 ranks).agg(sum(v), count(v), min(w), max(w), "
                               "mean(v)); v f64 N(0,1) 5% nulls, w i32 U[0,1000)",
                   "rows": n, "groups": stats["n_groups"] if stats else None,
                   "selected_rows": stats["n_selected"] if stats else None,
                   "hash_passes": stats["n_passes"] if stats else None,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "traffic": traffic_from_profiles(main_name),
                     "kernel": main_name, "kernel_avg_ms": avg_ms,
                     "algorithmic_bytes_per_launch": alg_bytes},
        "kernel_ms": {k: v[1] / v[0] for k, v in kt.items() if v[0]},
    }
    del bufs, cols
    if not args.no_host_path:
        log("host-buffer (PCIe-inclusive) path")
        out["pcie_inclusive"] = host_path(frame, plan, n)
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle, all cores)")
        out["cpu_baseline"] = cpu_baseline(frame, plan, min(n, args.cpu_rows))
    del frame
    if not args.no_sizer:
        log(f"learned sizing + batching vs plain on {args.sizer_jobs} fresh jobs")
        out["learned_vs_plain"] = stream_ab(args.sizer_jobs)
    if not args.no_highcard:
        log("high-cardinality leg: 200M rows, ~49M groups, hash and sort engines")
        hc = bench_highcard(argparse.Namespace(steps=5, warmup=2))
        out["highcard"] = {"ms_per_query": {m: e["ms_per_query"] for m, e in hc["config"]["engines"].items()},
                           "kernel_ms": {m: e["kernel_ms"] for m, e in hc["config"]["engines"].items()},
                           "groups": hc["config"]["engines"]["hash"]["groups"], "agree": hc["config"]["agree"],
                           "rows_per_s": hc["value"], "roofline": hc["roofline"], "config": hc["config"]["workload"],
                           "steps": 5, "warmup": 2}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
