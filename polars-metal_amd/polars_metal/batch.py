"""Batched group-by: many filtered group-by queries of one shape run as one
sketch launch plus one captured graph (pmh_execute_batch), instead of one
call per query.  No counterpart in the reference, which runs each query
with its own dispatches and blocking waits
(crates/polars-metal-core/src/command.rs:137-254).

``queries`` are ``(n_rows, columns, predicate)``: device columns as for
``_native.execute_filter_groupby_device`` and a PredicateAst dict or None.
All queries share the template plan's keys and aggregates.

The autotuner picks the batch's launch shape -- rows per tile and the
per-partition LDS table size -- by timing a few candidates on a subset of
the batch, and remembers the winner per (plan shape, size class)."""
from __future__ import annotations

import ctypes
import json
import math
from typing import Iterable, Sequence

import numpy as np

from . import _lib
from ._native import PMH_DEVICE, DeviceResult, _plan_json, _tag, execute_filter_groupby_device

# numpy images of PmhColumn and PmhBatchQuery, so a batch's descriptor
# tables are filled column-wise instead of field by field through ctypes
_COL_DT = np.dtype([("name", "u8"), ("tag", "i4"), ("location", "i4"), ("data", "u8"), ("data_len", "u8"),
                    ("valid", "u8"), ("valid_len", "u8")])
_QRY_DT = np.dtype([("n_rows", "u8"), ("columns", "u8"), ("n_columns", "u4"), ("pad", "u4"), ("predicate", "u8")])
assert _COL_DT.itemsize == ctypes.sizeof(_lib.PmhColumn)
assert _QRY_DT.itemsize == ctypes.sizeof(_lib.PmhBatchQuery)

CANDIDATES = [{"tile": 4096, "tail_slots": 512}, {"tile": 8192, "tail_slots": 512},
              {"tile": 4096, "tail_slots": 1024}, {"tile": 8192, "tail_slots": 1024},
              {"tile": 8192, "tail_slots": 256}]
_tuned: dict = {}


def _stats_dict(st: _lib.PmhBatchStats) -> dict:
    return {"device_ms": st.device_ms, "n_queries": st.n_queries, "n_tiles": st.n_tiles,
            "n_partitions": st.n_partitions, "n_rerun": st.n_rerun, "tile_rows": st.tile_rows,
            "part_slots": st.part_slots, "graph_reused": bool(st.graph_reused), "learned": bool(st.learned),
            "n_fallback_keys": st.n_fallback_keys}


def execute_batch(template_plan: dict, queries: Sequence[tuple], options: dict | None = None):
    """Runs every query; returns ``(results, stats)`` with one DeviceResult
    per query (outputs in device memory)."""
    lib = _lib.load()
    plan = dict(template_plan)
    plan.pop("predicate", None)
    if options:
        plan["options"] = dict(plan.get("options", {}), **options)
    n = len(queries)
    keep = []
    strings: dict = {}  # text -> address of a NUL-terminated copy kept alive in `keep`

    def cstr(b: bytes) -> int:
        a = strings.get(b)
        if a is None:
            buf = ctypes.create_string_buffer(b)
            keep.append(buf)
            a = strings[b] = ctypes.addressof(buf)
        return a

    preds: dict = {}  # id(predicate dict) -> address of its JSON text
    tags: dict = {}
    rows, counts, pred_addr = [], [], []
    names, ctag, data, dlen, valid, vlen = [], [], [], [], [], []
    for n_rows, cols, pred in queries:
        rows.append(n_rows)
        counts.append(len(cols))
        if pred is None:
            pred_addr.append(0)
        else:
            a = preds.get(id(pred))
            if a is None:
                keep.append(pred)  # keeps id(pred) unique while cached
                a = preds[id(pred)] = cstr(json.dumps(pred).encode())
            pred_addr.append(a)
        for name, tag, dp, dl, vp, vl in cols:
            names.append(cstr(name.encode()))
            t = tags.get(tag)
            if t is None:
                t = tags[tag] = _tag(tag)
            ctag.append(t)
            data.append(dp)
            dlen.append(dl)
            valid.append(vp or 0)
            vlen.append(vl if vp else 0)
    ca = np.zeros(max(1, len(names)), _COL_DT)
    if names:
        ca["name"], ca["tag"], ca["location"] = names, ctag, PMH_DEVICE
        ca["data"], ca["data_len"], ca["valid"], ca["valid_len"] = data, dlen, valid, vlen
    qa = np.zeros(max(1, n), _QRY_DT)
    if n:
        offs = np.zeros(n, np.uint64)
        np.cumsum(counts[:-1], out=offs[1:])
        qa["n_rows"], qa["n_columns"], qa["predicate"] = rows, counts, pred_addr
        qa["columns"] = np.uint64(ca.ctypes.data) + offs * np.uint64(_COL_DT.itemsize)
    res = (ctypes.POINTER(_lib.PmhResult) * max(1, n))()
    st = _lib.PmhBatchStats()
    _lib.check(lib.pmh_execute_batch(_plan_json(plan), qa.ctypes.data_as(ctypes.POINTER(_lib.PmhBatchQuery)), n,
                                     _lib.PMH_OUT_DEVICE, res, ctypes.byref(st)))
    del keep, ca, qa
    return [DeviceResult(res[i]) for i in range(n)], _stats_dict(st)


def sequential(template_plan: dict, queries: Sequence[tuple], options: dict | None = None) -> list:
    """The plain method: one execute_filter_groupby_device call per query.
    Returns each query's stats (results are freed)."""
    out = []
    for n_rows, cols, pred in queries:
        plan = dict(template_plan)
        plan.pop("predicate", None)
        if pred is not None:
            plan["predicate"] = pred
        if options:
            plan["options"] = dict(options)
        r = execute_filter_groupby_device(plan, n_rows, cols)
        out.append(dict(r.stats))
        r.free()
    return out


def _shape_key(template_plan: dict, queries: Sequence[tuple]) -> tuple:
    rows = [q[0] for q in queries] or [1]
    mean = sum(rows) / len(rows)
    return (json.dumps({k: template_plan[k] for k in ("keys", "aggs")}, sort_keys=True),
            int(math.log2(max(1, len(queries)))), int(math.log2(max(1.0, mean))))


def autotune(template_plan: dict, queries: Sequence[tuple], sample: int = 512, repeats: int = 3,
             candidates: Iterable[dict] = CANDIDATES) -> dict:
    """The fastest candidate launch shape on a spread subset of the batch
    (cached per plan shape and size class), by the engine's own HIP-event
    time of each run, best of ``repeats``."""
    key = _shape_key(template_plan, queries)
    if key in _tuned:
        return _tuned[key]
    step = max(1, len(queries) // sample)
    sub = list(queries)[::step][:sample]
    best, best_t = None, float("inf")
    for cand in candidates:
        t = float("inf")
        for _ in range(repeats):
            results, st = execute_batch(template_plan, sub, cand)
            t = min(t, st["device_ms"])
            for r in results:
                r.free()
        if t < best_t:
            best, best_t = cand, t
    _tuned[key] = dict(best)
    return _tuned[key]


def run(template_plan: dict, queries: Sequence[tuple], tune: bool = True):
    """Autotuned batch run: ``(results, stats)``."""
    opts = autotune(template_plan, queries) if tune else None
    results, stats = execute_batch(template_plan, queries, opts)
    stats["options"] = opts
    return results, stats
