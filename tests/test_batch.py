"""Batched group-by (pmh_execute_batch): every query of a batch against the
oracle, bit for bit, through the captured graph and its replays."""
import numpy as np
import pytest

import oracle as orc
from helpers import assert_results_equal

pytestmark = pytest.mark.gpu

AGGS = [{"kind": "Simple", "input_col": "v", "op": op, "output_alias": f"v_{op.lower()}"}
        for op in ("Sum", "Count", "Min", "Max", "Mean")] + [{"kind": "Length", "op": "Len", "output_alias": "n"}]
PRED = {"kind": "Compare", "op": "Gt", "dtype": "F64", "lhs": {"kind": "Column", "name": "v", "dtype": "F64"},
        "rhs": {"kind": "LiteralF64", "value": 0.5}}


def _frame(rng, rows, key_tag="I64", nulls=False):
    d = max(1, rows // 10)
    key = rng.integers(1, d + 1, rows)
    key_valid = None
    if nulls:
        key = np.where(rng.random(rows) < 0.05, -1, key)  # the all-ones key
        key_valid = rng.random(rows) >= 0.05
    v = rng.random(rows)
    v[rng.random(rows) < 0.01] = np.nan
    frame = {"key": (key_tag, key.astype(orc.NP_DTYPE[key_tag]), key_valid),
             "v": ("F64", v, (rng.random(rows) >= 0.05) if nulls else None)}
    return frame


def _fetch(res):
    ng = res.n_rows
    out = {"__n_groups__": ng}
    for c in res.columns:
        tag, data, valid = res.fetch(c["name"])
        vals = np.frombuffer(data, orc.NP_DTYPE[tag])[:ng] if tag != "Bool" else orc.unpack_bits(data, ng)
        ok = orc.unpack_bits(valid, ng) if valid is not None else np.ones(ng, bool)
        out[c["name"]] = (tag, vals, ok)
    return out


def _queries(seed, n, key_tag="I64", lo=1e3, hi=5e4, nulls=False):
    from polars_metal.buffer import device_columns
    rng = np.random.default_rng(seed)
    frames, queries, bufs = [], [], []
    for i in range(n):
        rows = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        f = _frame(rng, rows, key_tag, nulls=nulls and i % 3 == 0)
        pred = PRED if i % 5 == 0 else None
        cols, b = device_columns(f)
        frames.append(f)
        queries.append((rows, cols, pred))
        bufs.append(b)
    return frames, queries, bufs


def _plan(key_tag="I64"):
    return {"keys": [["key", key_tag]], "aggs": AGGS}


def _check(frames, queries, results, key_tag="I64", label=""):
    for i, (f, (rows, cols, pred), r) in enumerate(zip(frames, queries, results)):
        plan = dict(_plan(key_tag), predicate=pred) if pred else _plan(key_tag)
        assert_results_equal(_fetch(r), orc.groupby(plan, f), ["key"], label=f"{label}q{i}")


@pytest.mark.parametrize("key_tag,nulls", [("I64", False), ("I64", True), ("I32", True)])
def test_batch_matches_oracle(native, key_tag, nulls):
    from polars_metal import batch
    frames, queries, bufs = _queries(3, 48, key_tag, nulls=nulls)
    results, st = batch.execute_batch(_plan(key_tag), queries)
    assert st["n_queries"] == 48 and st["n_rerun"] == 0
    _check(frames, queries, results, key_tag, label=f"{key_tag}{nulls}/")


def test_batch_graph_replays_agree(native):
    """The captured graph is replayed for a second and third batch with the
    same buffers, with ordinary single-query calls in between: every replay
    matches the oracle."""
    from polars_metal import batch
    frames, queries, bufs = _queries(5, 40)
    plan = _plan()
    first, st0 = batch.execute_batch(plan, queries)
    _check(frames, queries, first, label="first/")
    for r in first:
        r.free()
    for rep in range(3):
        # eager work between replays
        res = native.execute_filter_groupby_device(dict(plan, predicate=PRED), queries[1][0], queries[1][1])
        res.free()
        again, st = batch.execute_batch(plan, queries)
        assert st["graph_reused"], rep
        _check(frames, queries, again, label=f"replay{rep}/")
        for r in again:
            r.free()


def test_batch_reruns_queries_the_sketch_underestimates(native):
    """Keys constant at every sampled row but unique elsewhere: the sketch
    sees one group, the output capacity is too small, and the query re-runs
    alone with the same exact result."""
    from polars_metal import batch
    from polars_metal.buffer import device_columns
    n = 131_072
    key = np.arange(n, dtype=np.int64) + 10
    key[::64] = 0  # the strided sample (2048 rows, stride 64) sees only key 0
    f = {"key": ("I64", key, None), "v": ("F64", np.random.default_rng(1).random(n), None)}
    cols, bufs = device_columns(f)
    frames, queries, more = _queries(7, 8)
    frames = [f] + frames
    queries = [(n, cols, None)] + queries
    results, st = batch.execute_batch(_plan(), queries)
    assert st["n_rerun"] >= 1
    assert st["n_fallback_keys"] > 0  # the keys beyond the one-group table went to the fallback table
    _check(frames, queries, results, label="rerun/")


def test_batch_autotune_and_sequential_agree(native):
    from polars_metal import batch
    frames, queries, bufs = _queries(9, 64, lo=1e3, hi=2e5)
    opts = batch.autotune(_plan(), queries, sample=32, repeats=1)
    assert opts in batch.CANDIDATES
    results, st = batch.run(_plan(), queries)
    _check(frames, queries, results, label="tuned/")
    seq = batch.sequential(_plan(), queries)
    assert [s["n_groups"] for s in seq] == [r.n_rows for r in results]


# One aggregate shape per compile-time signature of the batch fold (the
# first four) and one resolved at run time.
BATCH_SIG_AGGS = [
    [("v", "Sum"), ("v", "Mean"), ("w", "Min"), ("w", "Max")],
    [("a", "Sum")],
    [("v", "Sum"), ("v", "Min"), ("v", "Max")],
    [("v", "Sum")],
    [("a", "Sum"), ("a", "Min"), ("w", "Max")],
]


@pytest.mark.parametrize("sig", range(len(BATCH_SIG_AGGS)))
def test_batch_aggregate_signatures(native, sig):
    from polars_metal import batch
    from polars_metal.buffer import device_columns
    rng = np.random.default_rng(40 + sig)
    plan = {"keys": [["key", "I64"]],
            "aggs": [{"kind": "Simple", "input_col": c, "op": op, "output_alias": f"{c}_{op.lower()}"}
                     for c, op in BATCH_SIG_AGGS[sig]]}
    frames, queries, bufs = [], [], []
    for _ in range(24):
        rows = int(np.exp(rng.uniform(np.log(1e3), np.log(3e4))))
        f = {"key": ("I64", rng.integers(1, max(2, rows // 10) + 1, rows), None),
             "v": ("F64", rng.standard_normal(rows), rng.random(rows) >= 0.05),
             "w": ("I32", rng.integers(-1000, 1000, rows).astype(np.int32), None),
             "a": ("I64", rng.integers(-10**12, 10**12, rows), rng.random(rows) >= 0.1)}
        cols, b = device_columns(f)
        frames.append(f)
        queries.append((rows, cols, None))
        bufs.append(b)
    results, st = batch.execute_batch(plan, queries)
    assert st["n_queries"] == 24
    for i, (f, r) in enumerate(zip(frames, results)):
        assert_results_equal(_fetch(r), orc.groupby(plan, f), ["key"], label=f"sig{sig}/q{i}")


def test_batch_bool_key(native):
    """A Bool key, with and without nulls: the key column comes back as bits
    packed from the batch's byte rows, beside its validity."""
    from polars_metal import batch
    from polars_metal.buffer import device_columns
    rng = np.random.default_rng(77)
    plan = {"keys": [["flag", "Bool"]],
            "aggs": [{"kind": "Simple", "input_col": "v", "op": "Sum", "output_alias": "v_sum"},
                     {"kind": "Length", "op": "Len", "output_alias": "n"}]}
    frames, queries, bufs = [], [], []
    for i in range(16):
        rows = int(rng.integers(100, 20_000))
        f = {"flag": ("Bool", rng.random(rows) < 0.5, (rng.random(rows) >= 0.1) if i % 2 else None),
             "v": ("F64", rng.standard_normal(rows), None)}
        cols, b = device_columns(f)
        frames.append(f)
        queries.append((rows, cols, None))
        bufs.append(b)
    results, st = batch.execute_batch(plan, queries)
    assert st["n_queries"] == 16
    for i, (f, r) in enumerate(zip(frames, results)):
        got = _fetch(r)
        assert got["__n_groups__"] == (3 if i % 2 else 2)
        assert_results_equal(got, orc.groupby(plan, f), ["flag"], label=f"bool/q{i}")


def test_batch_refuses_oversized_or_short_queries(native):
    """Every query is checked on the host before any launch, not only the
    first: a query claiming 2^31 rows or more (the single-query limit, which
    a query the batch cannot finish is re-run on) and a query whose buffers
    are shorter than its row count are refused, and the next batch runs as
    usual."""
    from polars_metal import batch
    frames, queries, bufs = _queries(11, 3)
    rows, cols, pred = queries[1]
    for n in (1 << 31, (1 << 32) - 1, 1 << 32):
        with pytest.raises(ValueError, match=r"2\^31"):
            batch.execute_batch(_plan(), [queries[0], (n, cols, pred), queries[2]])
    with pytest.raises(ValueError, match="buffer too short"):
        batch.execute_batch(_plan(), [queries[0], (rows + 64, cols, pred), queries[2]])
    results, st = batch.execute_batch(_plan(), queries)
    _check(frames, queries, results, label="after/")


def test_batch_predicate_on_a_column_no_aggregate_reads(native):
    """filter(w > t).group_by(key).agg(v...): the template names key and v
    only; w comes from the queries' predicates."""
    from polars_metal import batch
    from polars_metal.buffer import device_columns
    rng = np.random.default_rng(17)
    frames, queries, bufs = [], [], []
    for i in range(6):
        rows = int(rng.integers(2000, 40000))
        f = _frame(rng, rows)
        f["w"] = ("I32", rng.integers(0, 100, rows).astype(np.int32), rng.random(rows) < 0.9)
        pred = {"kind": "Compare", "op": "Ge", "dtype": "I64", "lhs": {"kind": "Column", "name": "w", "dtype": "I64"},
                "rhs": {"kind": "LiteralI64", "value": int(10 * i)}} if i % 2 == 0 else None
        cols, b = device_columns(f)
        frames.append(f)
        queries.append((rows, cols, pred))
        bufs.append(b)
    results, st = batch.execute_batch(_plan(), queries)
    assert st["n_queries"] == 6
    _check(frames, queries, results, label="predw/")
