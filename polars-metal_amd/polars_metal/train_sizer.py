"""Train the learned sizer on the device from logs of the plain method.

    python -m polars_metal.train_sizer --jobs 2000 --out sizer_model.json

For each job of the synthetic stream (polars_metal.synth.sizer_job_specs)
the plain engines run on device-resident columns: the sample sketch gives
the features, the hash and sort engines give the true group count and both
device times.  The network is then trained by the library's own forward,
backward and Adam kernels (pmh_mlp_train) and written as JSON.  Needs an
MI355X (the library has no CPU path)."""
from __future__ import annotations

import argparse
import sys
import time

import numpy as np

from . import _native, learned, synth
from .buffer import device_columns


def run_query(plan: dict, n: int, cols: list) -> dict:
    r = _native.execute_filter_groupby_device(plan, n, cols)
    st = dict(r.stats)
    r.free()
    return st


def job_stream(specs):
    for spec in specs:
        frame, plan = synth.sizer_job(spec)
        cols, bufs = device_columns(frame)
        yield plan, spec["rows"], cols
        del bufs


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=2000)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--rows-hi", type=float, default=1e7)
    ap.add_argument("--epochs", type=int, default=400)
    ap.add_argument("--out", type=str, default=str(learned.MODEL_PATH))
    ap.add_argument("--logs", type=str, default="")
    args = ap.parse_args(argv)
    t0 = time.perf_counter()
    specs = synth.sizer_job_specs(args.jobs, args.seed, rows_hi=args.rows_hi)
    # grow every device buffer once, at the largest job, so no logged call
    # pays an allocation
    big = max(specs, key=lambda s: s["rows"])
    for plan, n, cols in job_stream([dict(big, dtype=d) for d in ("I32", "I64", "U128")]):
        for m in ("hash", "sort"):
            run_query(dict(plan, options={"method": m}), n, cols)
    x, y, meta = learned.collect_logs(job_stream(specs), run_query, log=lambda s: print(s, flush=True))
    print(f"collected {len(meta)} jobs in {time.perf_counter() - t0:.1f} s", flush=True)
    if args.logs:
        np.savez(args.logs, x=x, y=y, rows=[m["rows"] for m in meta], groups=[m["groups"] for m in meta],
                 hash_ms=[m["hash_ms"] for m in meta], sort_ms=[m["sort_ms"] for m in meta])
    # split by job: every fifth job is held out
    test = np.arange(len(meta)) % 5 == 0
    train = ~test
    model = learned.DeviceMLP(learned.init_params(args.seed))
    loss = model.fit(x[train], y[train], epochs=args.epochs, batch=64, lr=3e-3, lam=1.0, seed=args.seed)
    params = model.params()
    th = np.array([m["hash_ms"] for m in meta])
    ts = np.array([m["sort_ms"] for m in meta])
    report = {}
    for name, sel in (("train", train), ("held_out", test)):
        pred = model.forward(x[sel])
        pick_sort = pred[:, 1] > 0
        chosen = np.where(pick_sort, ts[sel], th[sel])
        report[name] = {
            "jobs": int(sel.sum()),
            "path_accuracy": float(np.mean(pick_sort == (y[sel, 1] > 0.5))),
            "median_log2_error": float(np.median(np.abs(pred[:, 0] - y[sel, 0]) * 32)),
            "plain_ms": float(th[sel].sum()), "learned_ms": float(chosen.sum()),
            "best_ms": float(np.minimum(th[sel], ts[sel]).sum()),
            "gain": float(th[sel].sum() / chosen.sum()),
        }
        print(f"{name}: {report[name]}", flush=True)
    learned.save_model(params, args.out, meta={"jobs": args.jobs, "seed": args.seed, "rows_hi": args.rows_hi,
                                                "epochs": args.epochs, "final_loss": loss, "report": report})
    return 0


if __name__ == "__main__":
    sys.exit(main())
