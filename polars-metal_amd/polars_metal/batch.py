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
import time
from typing import Iterable, Sequence

from . import _lib
from ._native import DeviceResult, _device_column_array, _plan_json, execute_filter_groupby_device

CANDIDATES = [{"tile": 4096, "tail_slots": 512}, {"tile": 8192, "tail_slots": 512},
              {"tile": 4096, "tail_slots": 1024}, {"tile": 8192, "tail_slots": 1024},
              {"tile": 8192, "tail_slots": 256}]
_tuned: dict = {}


def _stats_dict(st: _lib.PmhBatchStats) -> dict:
    return {"device_ms": st.device_ms, "n_queries": st.n_queries, "n_tiles": st.n_tiles,
            "n_partitions": st.n_partitions, "n_rerun": st.n_rerun, "tile_rows": st.tile_rows,
            "part_slots": st.part_slots, "graph_reused": bool(st.graph_reused), "learned": bool(st.learned)}


def execute_batch(template_plan: dict, queries: Sequence[tuple], options: dict | None = None):
    """Runs every query; returns ``(results, stats)`` with one DeviceResult
    per query (outputs in device memory)."""
    lib = _lib.load()
    plan = dict(template_plan)
    plan.pop("predicate", None)
    if options:
        plan["options"] = dict(plan.get("options", {}), **options)
    n = len(queries)
    arr = (_lib.PmhBatchQuery * max(1, n))()
    keep = []
    for i, (n_rows, cols, pred) in enumerate(queries):
        carr, k = _device_column_array(list(cols))
        keep.append((carr, k))
        arr[i].n_rows = n_rows
        arr[i].columns = carr
        arr[i].n_columns = len(cols)
        if pred is not None:
            pj = json.dumps(pred).encode()
            keep.append(pj)
            arr[i].predicate_json = pj
        else:
            arr[i].predicate_json = None
    res = (ctypes.POINTER(_lib.PmhResult) * max(1, n))()
    st = _lib.PmhBatchStats()
    _lib.check(lib.pmh_execute_batch(_plan_json(plan), arr, n, _lib.PMH_OUT_DEVICE, res, ctypes.byref(st)))
    del keep
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


def autotune(template_plan: dict, queries: Sequence[tuple], sample: int = 512, repeats: int = 2,
             candidates: Iterable[dict] = CANDIDATES) -> dict:
    """The fastest candidate launch shape on a spread subset of the batch
    (cached per plan shape and size class)."""
    key = _shape_key(template_plan, queries)
    if key in _tuned:
        return _tuned[key]
    step = max(1, len(queries) // sample)
    sub = list(queries)[::step][:sample]
    best, best_t = None, float("inf")
    for cand in candidates:
        t = float("inf")
        for _ in range(repeats):
            t0 = time.perf_counter()
            results, _ = execute_batch(template_plan, sub, cand)
            t = min(t, time.perf_counter() - t0)
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
