"""Benchmark of the filtered group-by hot path on one MI355X.

Workload (one "step" = one fused call over one batch of 100M rows, inputs
already resident in HBM, outputs left in HBM):
  key i64 bounded Zipf(s=1.1) over ranks 1..1e6, v f64 N(0,1) with 5% nulls,
  w i32 U{0..999}; filter (w >= 200) & (w < 500) (~30% of rows);
  group_by(key).agg(sum(v), count(v), min(w), max(w), mean(v)).

Prints one JSON line with value = input rows per second, the roofline of the
dominant kernel (HBM-bound; algorithmic bytes = 20.125 B per input row:
8 key + 8 v + 4 w + 1/8 validity) and the CPU baseline (the C oracle, one
core, on a bounded sample of the same distribution).
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


def cpu_baseline(rows: int) -> dict:
    import oracle as orc
    from polars_metal import synth
    orc.build()
    frame, plan = synth.zipf(rows, seed=12345)
    t0 = time.perf_counter()
    res = orc.groupby(plan, frame, sum_mode=1)
    dt = time.perf_counter() - t0
    return {"value": rows / dt, "unit": "rows/s", "cores": 1, "kind": "port",
            "sample": f"{rows} rows of the same zipf distribution (seed 12345), C oracle with sequential f64 "
                      f"sums, {res['__n_groups__']} groups, {dt:.2f} s"}


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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rows", type=int, default=100_000_000)
    ap.add_argument("--cpu-rows", type=int, default=20_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--options", type=str, default="", help="JSON plan options (launch shape overrides)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU; the multi-GPU sharding is a design only "
              "(see the sharding section of the design notes)", file=sys.stderr)
        return 2

    from polars_metal import _native, synth
    from polars_metal.buffer import device_columns

    n = args.rows
    print(f"bench: generating {n} rows (numpy, seed 2) and uploading", file=sys.stderr, flush=True)
    frame, plan = synth.zipf(n, seed=2)
    cols, bufs = device_columns(frame)
    del frame
    if args.options:
        plan["options"] = json.loads(args.options)

    # Every call synchronises its stream before returning (the group count is
    # read back), so host timestamps around the loop bracket the device work.
    for _ in range(args.warmup):
        r = _native.execute_filter_groupby_device(plan, n, cols)
        r.free()
    _native.set_profiling(True)
    print(f"bench: timing {args.steps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    stats = None
    for _ in range(args.steps):
        r = _native.execute_filter_groupby_device(plan, n, cols)
        stats = r.stats
        r.free()
    dt = time.perf_counter() - t0
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
    if not args.no_cpu_baseline:
        print("bench: CPU baseline (oracle, 1 core)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args.cpu_rows)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
