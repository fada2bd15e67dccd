"""A stream of uneven group-by jobs, run the fast way: jobs that share a
plan template (keys and aggregates) go through one batched launch
(``batch.run``: one sketch launch plus one captured graph for all of them),
the rest one call each, all with the learned sizer installed, which sizes
the batch's partitions on the device and picks the engine and sizing of
each single call.  No counterpart in the reference, which runs every query
with its own dispatches and blocking waits
(crates/polars-metal-core/src/command.rs:137-254).

``jobs`` are ``(plan, n_rows, columns)`` with device columns as for
``_native.execute_filter_groupby_device``; a plan's ``predicate`` is the
job's own.  Device outputs stay valid only until the library's next call
(a batch's until the next batch), so each job's result is handed to
``consume(job_index, result)`` while it is valid and freed afterwards;
jobs may complete out of order.  Results are bit-identical to the plain
path (``run_plain``): the sizer and the batching decide only how the work
is laid out.
"""
from __future__ import annotations

import json
from typing import Callable, Sequence

from . import _native, batch, learned

__all__ = ["run", "run_plain", "template_key"]


def template_key(plan: dict) -> str:
    """Jobs batch together when their keys and aggregates are the same."""
    return json.dumps({"keys": plan["keys"], "aggs": plan["aggs"]}, sort_keys=True)


def _single(plan: dict, n: int, cols) -> object:
    return _native.execute_filter_groupby_device(dict(plan, options={"method": "auto"}), n, cols)


def _hand_over(i: int, r, consume: Callable | None) -> None:
    if consume is not None:
        consume(i, r)
    r.free()


def run_plain(jobs: Sequence[tuple], consume: Callable | None = None) -> None:
    """The plain method: one call per job, method "auto" without a model."""
    learned.install(None)
    for i, (plan, n, cols) in enumerate(jobs):
        _hand_over(i, _single(plan, n, cols), consume)


# A single call carries a fixed cost of roughly 0.15 ms (its sample sketch,
# about ten launches, two host synchronisations) and then streams at about
# 11 us per million rows; the batch streams at about 44 us per million rows
# but shares one sketch launch and one graph among all its queries.  On 120
# jobs of the C4 distribution (rows log-uniform in [1e3, 1e7]) the stream
# measured 1.16x over one plain call per job with this limit at 2^18 rows,
# 1.25x at 2^19, 1.13x at 2^20 and 1.09x at 2^21.
BATCH_MAX_ROWS = 1 << 19


def run(jobs: Sequence[tuple], consume: Callable | None = None, min_batch: int = 2, tune: bool = True,
        batch_max_rows: int = BATCH_MAX_ROWS) -> dict:
    """Learned sizing plus batching.  Returns counts of the jobs that ran
    batched and alone.  The batches size their partitions from their own
    sample sketch: the model is trained on the single-call sketch's
    features, which the batch's smaller per-query sample does not match."""
    learned.install_shipped()
    try:
        groups: dict = {}
        for i, (plan, n, cols) in enumerate(jobs):
            key = template_key(plan) if n <= batch_max_rows else None
            groups.setdefault(key, []).append(i)
        info = {"batched": 0, "single": 0, "batches": 0, "batch_stats": []}
        for key, idxs in groups.items():
            if key is not None and len(idxs) >= min_batch:
                tmpl = {k: v for k, v in jobs[idxs[0]][0].items() if k not in ("predicate", "options")}
                tmpl["options"] = {"method": "hash"}
                queries = [(jobs[i][1], jobs[i][2], jobs[i][0].get("predicate")) for i in idxs]
                try:
                    res, bst = batch.run(tmpl, queries, tune=tune)
                    info["batch_stats"].append(bst)
                except NotImplementedError:
                    res = None  # the batch does not take this shape (e.g. keys wider than 64 bits)
                if res is not None:
                    for i, r in zip(idxs, res):
                        _hand_over(i, r, consume)
                    info["batched"] += len(idxs)
                    info["batches"] += 1
                    continue
            for i in idxs:
                plan, n, cols = jobs[i]
                _hand_over(i, _single(plan, n, cols), consume)
                info["single"] += 1
        return info
    finally:
        learned.install(None)
