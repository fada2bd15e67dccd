"""rocprofv3 summaries: kernel-trace tables and the scan kernel's HBM traffic.

Traffic: run a workload under ``rocprofv3 --pmc FETCH_SIZE`` and again under
``--pmc WRITE_SIZE`` (one counter group per run), do the same for a plain
device copy of known size, then ``traffic`` reads the four result databases
and writes the bytes per launch, each counter scaled by what it reads for
the copy (which moves exactly its size each way):

    rocprofv3 --pmc FETCH_SIZE -d OUT/f -o run -- python3 rocprof_summary.py zipf100m
    rocprofv3 --pmc WRITE_SIZE -d OUT/w -o run -- python3 rocprof_summary.py zipf100m
    rocprofv3 --pmc FETCH_SIZE -d OUT/cf -o run -- python3 rocprof_summary.py copy
    rocprofv3 --pmc WRITE_SIZE -d OUT/cw -o run -- python3 rocprof_summary.py copy
    python3 rocprof_summary.py traffic OUT hs_scan_fast > traffic.json

The counters are KiB per dispatch (summed over the device's channels).

``traffic_all OUT "title" "commands"`` prints the same scaled counters for
every kernel of the run as a markdown table.  ``highcard --method hash|sort``
runs the 200M-row composite-key workload instead of ``zipf100m``.

Kernel traces: ``python3 rocprof_summary.py trace DB_DIR "title" "command"`` prints
a markdown table of a ``rocprofv3 --kernel-trace --stats`` result database;
``counters "title" "command" DIR...`` tabulates every counter of ``--pmc`` runs."""
from __future__ import annotations

import argparse
import collections
import glob
import json
import sqlite3
import sys
from pathlib import Path

_PKG = Path(__file__).resolve().parent / "polars-metal_amd"
if str(_PKG) not in sys.path:
    sys.path.insert(0, str(_PKG))

COPY_BYTES = 1 << 30
COPY_KERNEL = "__amd_rocclr_copyBuffer"


def run_zipf100m(steps: int, rows: int) -> None:
    from polars_metal import _native, synth
    from polars_metal.buffer import device_columns
    frame, plan = synth.zipf(rows, seed=2)
    cols, bufs = device_columns(frame)
    del frame
    for i in range(steps):
        r = _native.execute_filter_groupby_device(plan, rows, cols)
        print(f"step {i}: {r.stats['n_groups']} groups", flush=True)
        r.free()


def run_highcard(steps: int, method: str) -> None:
    from polars_metal import _native, synth
    from polars_metal.buffer import device_columns
    n = 200_000_000
    frame, plan = synth.highcard(n, seed=3)
    cols, bufs = device_columns(frame)
    del frame
    plan = dict(plan, options={"method": method})
    for i in range(steps):
        r = _native.execute_filter_groupby_device(plan, n, cols)
        print(f"step {i}: {r.stats['n_groups']} groups ({r.stats['method']})", flush=True)
        r.free()


def run_copy(steps: int) -> None:
    from polars_metal.buffer import DeviceBuffer
    a = DeviceBuffer.alloc(COPY_BYTES, zeroed=True)
    b = DeviceBuffer.alloc(COPY_BYTES)
    for i in range(steps):
        b.copy_from(a)
        print(f"copy {i}: {COPY_BYTES} bytes", flush=True)


def counter_kib(db_dir: Path, counter: str) -> dict:
    """Mean value of one counter per dispatch, per kernel name."""
    db = glob.glob(str(db_dir / "*.db"))[0]
    c = sqlite3.connect(db)
    cols = [r[1] for r in c.execute("pragma table_info(counters_collection)")]
    ki, ni, vi = cols.index("kernel_name"), cols.index("counter_name"), cols.index("value")
    di = cols.index("dispatch_id")
    tot = collections.defaultdict(float)
    disp = collections.defaultdict(set)
    for r in c.execute("select * from counters_collection"):
        if r[ni] != counter:
            continue
        name = r[ki].replace("(anonymous namespace)::", "").split("(")[0].split("::")[-1].split("<")[0].strip()
        tot[name] += r[vi]
        disp[name].add(r[di])
    return {k: tot[k] / len(disp[k]) for k in tot}


def summarize(out: Path, kernel: str) -> dict:
    f = counter_kib(out / "f", "FETCH_SIZE")
    w = counter_kib(out / "w", "WRITE_SIZE")
    cf = counter_kib(out / "cf", "FETCH_SIZE")
    cw = counter_kib(out / "cw", "WRITE_SIZE")
    fetch_factor = COPY_BYTES / (cf[COPY_KERNEL] * 1024.0)
    write_factor = COPY_BYTES / (cw[COPY_KERNEL] * 1024.0)
    fetch = f[kernel] * 1024.0 * fetch_factor
    write = w[kernel] * 1024.0 * write_factor
    return {"kernel": kernel, "fetch_bytes_per_launch": fetch, "write_bytes_per_launch": write,
            "hbm_bytes_per_launch": fetch + write, "fetch_kib_raw": f[kernel], "write_kib_raw": w[kernel],
            "calibration": {"copy_bytes": COPY_BYTES, "copy_fetch_kib": cf[COPY_KERNEL],
                            "copy_write_kib": cw[COPY_KERNEL], "fetch_factor": fetch_factor,
                            "write_factor": write_factor}}


def traffic_table(out: Path, steps: int, title: str, command: str) -> str:
    """Markdown table of every kernel's FETCH_SIZE and WRITE_SIZE per
    launch, each scaled by what the counter reads for the 1 GiB copy."""
    f = counter_kib(out / "f", "FETCH_SIZE")
    w = counter_kib(out / "w", "WRITE_SIZE")
    cf = counter_kib(out / "cf", "FETCH_SIZE")
    cw = counter_kib(out / "cw", "WRITE_SIZE")
    ff = COPY_BYTES / (cf[COPY_KERNEL] * 1024.0)
    wf = COPY_BYTES / (cw[COPY_KERNEL] * 1024.0)
    lines = [f"# {title}", "", "Commands (from `/tmp`, with `REPO` the repository root, on one MI355X; each counter "
             "in its own run):", "", "```", command, "```", "",
             f"Counters are KiB per dispatch; FETCH_SIZE is scaled by {ff:.3f} and WRITE_SIZE by {wf:.3f}, the "
             f"factors that make a {COPY_BYTES} B device copy read and write its size.", "",
             "| kernel | fetch MB / launch | write MB / launch | total MB |", "|---|---|---|---|"]
    tot_f = tot_w = 0.0
    for k in sorted(set(f) | set(w), key=lambda k: -(f.get(k, 0) + w.get(k, 0))):
        if k == COPY_KERNEL:
            continue
        fb, wb = f.get(k, 0.0) * 1024 * ff, w.get(k, 0.0) * 1024 * wf
        tot_f += fb
        tot_w += wb
        lines.append(f"| `{k}` | {fb / 1e6:.1f} | {wb / 1e6:.1f} | {(fb + wb) / 1e6:.1f} |")
    lines.append(f"| **all** | {tot_f / 1e6:.1f} | {tot_w / 1e6:.1f} | {(tot_f + tot_w) / 1e6:.1f} |")
    return "\n".join(lines) + "\n"


def counters_table(dirs: list, title: str, command: str) -> str:
    """Markdown table of every counter of the given ``--pmc`` runs, mean
    per dispatch, per kernel."""
    per = collections.defaultdict(dict)
    names = []
    for d in dirs:
        db = glob.glob(str(Path(d) / "*.db"))[0]
        c = sqlite3.connect(db)
        cols = [r[1] for r in c.execute("pragma table_info(counters_collection)")]
        ki, ni, vi = cols.index("kernel_name"), cols.index("counter_name"), cols.index("value")
        di = cols.index("dispatch_id")
        tot = collections.defaultdict(float)
        disp = collections.defaultdict(set)
        for r in c.execute("select * from counters_collection"):
            k = r[ki].replace("(anonymous namespace)::", "").split("(")[0].split("::")[-1].split("<")[0].strip()
            tot[(k, r[ni])] += r[vi]
            disp[(k, r[ni])].add(r[di])
        for (k, n), v in tot.items():
            per[k][n] = v / len(disp[(k, n)])
            if n not in names:
                names.append(n)
    lines = [f"# {title}", "", "Commands (from `/tmp`, with `REPO` the repository root, on one MI355X; each counter "
             "group in its own run):", "", "```", command, "```", "",
             "Mean per dispatch (rocprofv3 `--pmc`, summed over the device):", "",
             "| kernel | " + " | ".join(names) + " |", "|---|" + "---|" * len(names)]
    for k in sorted(per, key=lambda k: -per[k].get("SQ_WAVE_CYCLES", 0.0)):
        if k == COPY_KERNEL:
            continue
        lines.append(f"| `{k}` | " + " | ".join(f"{per[k].get(n, 0.0):.0f}" for n in names) + " |")
    return "\n".join(lines) + "\n"


def short_name(name: str) -> str:
    n = name.replace("pmh::(anonymous namespace)::", "").replace("pmh::", "").replace("void ", "")
    return n.split("(")[0].strip()


def trace_markdown(db_dir: Path, title: str, command: str) -> str:
    db = glob.glob(str(db_dir / "*.db"))[0]
    c = sqlite3.connect(db)
    shape = {}
    for r in c.execute("select name, grid_x, workgroup_x, lds_size, vgpr_count, sgpr_count, scratch_size "
                       "from kernels"):
        shape.setdefault(r[0], r[1:])
    lines = [f"# {title}", "", "Command (from the repository root, on one MI355X):", "", "```", command, "```", "",
             "rocprofv3 kernel-trace statistics (durations in microseconds; grid in work-items):", "",
             "| kernel | calls | total us | avg us | % | grid | block | LDS B | VGPR | SGPR | scratch |",
             "|---|---|---|---|---|---|---|---|---|---|---|"]
    for name, calls, tot, avg, pct in c.execute(
            "select name, total_calls, total_duration, average, percentage from top_kernels"):
        g = shape.get(name, (None,) * 6)
        lines.append(f"| `{short_name(name)}` | {calls} | {tot:.1f} | {avg:.1f} | {pct:.1f} | "
                     f"{g[0]} | {g[1]} | {g[2]} | {g[3]} | {g[4]} | {g[5]} |")
    return "\n".join(lines) + "\n"


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=("zipf100m", "highcard", "copy", "traffic", "traffic_all", "trace", "counters"))
    ap.add_argument("args", nargs="*")
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--rows", type=int, default=100_000_000)
    ap.add_argument("--method", default="hash", help="highcard: the engine")
    a = ap.parse_args(argv)
    if a.mode == "zipf100m":
        run_zipf100m(a.steps, a.rows)
    elif a.mode == "highcard":
        run_highcard(a.steps, a.method)
    elif a.mode == "copy":
        run_copy(a.steps)
    elif a.mode == "traffic_all":
        print(traffic_table(Path(a.args[0]), a.steps, a.args[1], a.args[2]), end="")
    elif a.mode == "counters":
        print(counters_table(a.args[2:], a.args[0], a.args[1]), end="")
    elif a.mode == "traffic":
        print(json.dumps(summarize(Path(a.args[0]), a.args[1] if len(a.args) > 1 else "hs_scan_fast"), indent=1))
    else:
        print(trace_markdown(Path(a.args[0]), a.args[1], a.args[2]), end="")
    return 0


if __name__ == "__main__":
    sys.exit(main())
