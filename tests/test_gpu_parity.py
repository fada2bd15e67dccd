"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Integer, key, count, min/max and null results must be bit-exact.  f64 sums
and means are compared by bits too: both sides compute the correctly rounded
exact per-group sum (the oracle with Shewchuk partials, the kernels with
fixed-point limbs), so any difference is a bug, not rounding noise."""
import numpy as np
import pytest

import oracle as orc
from golden_cases import CASES, expected_of, frame_of
from helpers import assert_results_equal, run_gpu
from polars_metal import synth

pytestmark = pytest.mark.gpu


def _keys(plan):
    return [k[0] for k in plan["keys"]]


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_golden_cases(native, case):
    frame = frame_of(case)
    print(f"golden case {case['name']}", flush=True)
    got = run_gpu(native, case["plan"], frame)
    assert_results_equal(got, expected_of(case, frame), _keys(case["plan"]), label=case["name"])
    if not case["plan"].get("predicate"):
        got2 = run_gpu(native, case["plan"], frame, entry="groupby")
        assert_results_equal(got2, expected_of(case, frame), _keys(case["plan"]), label=case["name"] + "/groupby")


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_golden_cases_sort(native, case):
    frame = frame_of(case)
    plan = dict(case["plan"], options={"method": "sort"})
    got = run_gpu(native, plan, frame)
    assert_results_equal(got, expected_of(case, frame), _keys(case["plan"]), label=case["name"] + "/sort")
    assert native.last_stats()["method"] == "sort"


GOLDEN_ENGINES = [
    {"method": "hash", "lds_slots": 0},                      # every group through the tail partitions
    {"method": "hash", "partitions": 2, "tail_slots": 16},   # tail tables overflow into the fallback table
    {"method": "hash", "fast_scan": 0},                      # generic scan
    {"method": "hash_global"}, {"method": "hash_global", "lds_slots": 0},
]


@pytest.mark.parametrize("opts", GOLDEN_ENGINES, ids=[str(o) for o in GOLDEN_ENGINES])
@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_golden_cases_every_engine(native, case, opts):
    frame = frame_of(case)
    plan = dict(case["plan"], options=opts)
    print(f"golden case {case['name']} {opts}", flush=True)
    got = run_gpu(native, plan, frame)
    assert_results_equal(got, expected_of(case, frame), _keys(case["plan"]), label=f"{case['name']}/{opts}")


def _conjunction(p):
    if p is None or p["kind"] == "Compare":
        return True
    return p["kind"] == "And" and _conjunction(p["lhs"]) and _conjunction(p["rhs"])


def _batchable(case):
    """The batch takes one key lane of at most 64 bits, at most two
    aggregated columns (no expression aggregates) and conjunctions of
    compares."""
    keys = case["plan"]["keys"]
    if any(a.get("kind") == "Expression" for a in case["plan"]["aggs"]):
        return False
    vals = {a["input_col"] for a in case["plan"]["aggs"] if a.get("op") != "Len"}
    return (len(keys) == 1 and keys[0][1] != "F32" and len(vals) <= 2 and case["n_rows"] > 0
            and _conjunction(case["plan"].get("predicate")))


@pytest.mark.parametrize("case", [c for c in CASES if _batchable(c)],
                         ids=[c["name"] for c in CASES if _batchable(c)])
def test_golden_cases_batch(native, case):
    from polars_metal import batch
    from polars_metal.buffer import device_columns
    from test_batch import _fetch
    frame = frame_of(case)
    plan = {k: v for k, v in case["plan"].items() if k != "predicate"}
    cols, bufs = device_columns(frame)
    pred = case["plan"].get("predicate")
    print(f"golden case {case['name']} batch", flush=True)
    # the case twice in one batch
    results, st = batch.execute_batch(plan, [(case["n_rows"], cols, pred), (case["n_rows"], cols, pred)])
    assert st["n_queries"] == 2
    for r in results:
        assert_results_equal(_fetch(r), expected_of(case, frame), _keys(case["plan"]), label=f"{case['name']}/batch")
    del bufs


def _sorted_bits(res, keys):
    c = orc.canonical(orc.from_wire(res), keys)
    return {k: (v[1].tobytes(), v[2].tobytes()) for k, v in c.items() if not k.startswith("__")}


@pytest.mark.parametrize("rows", [1_000, 2_000_000])
def test_sort_agrees_with_hash_bit_exact(native, rows):
    # hash-vs-sort agreement is the reference's own oracle for its sort path
    # (both must produce the same groups and the same bits)
    frame, plan = synth.zipf(rows, seed=rows + 7)
    wire = orc.to_wire(frame)
    a = native.execute_filter_groupby(dict(plan, options={"method": "hash"}), rows, wire)
    b = native.execute_filter_groupby(dict(plan, options={"method": "sort"}), rows, wire)
    st = native.last_stats()
    assert st["method"] == "sort" and 1 <= st["n_passes"] <= 3  # keys 1..1e6: at most 20 varying bits
    assert _sorted_bits(a, ["key"]) == _sorted_bits(b, ["key"])
    assert_results_equal(orc.from_wire(b), orc.groupby(plan, frame), ["key"], label=f"sort{rows}")


def test_sort_edge_shapes(native):
    rng = np.random.default_rng(5)
    n = 50_000
    aggs = [{"kind": "Simple", "input_col": "v", "op": op, "output_alias": op} for op in ("Sum", "Count", "Min", "Max", "Mean")]
    v = ("F64", rng.standard_normal(n), rng.random(n) < 0.9)
    cases = [
        # one group (no varying key bit: zero radix passes)
        ({"k": ("I64", np.full(n, 7, np.int64), None), "v": v}, [["k", "I64"]]),
        # compressed key wider than 64 bits: two random i64 halves
        ({"a": ("I64", rng.integers(-2**62, 2**62, n), rng.random(n) < 0.95),
          "b": ("I32", rng.integers(-2**31, 2**31 - 1, n).astype(np.int32), None), "v": v}, [["a", "I64"], ["b", "I32"]]),
        # all-ones and null i64 keys beside ordinary ones
        ({"k": ("I64", np.where(rng.random(n) < 0.1, -1, rng.integers(0, 50, n)), rng.random(n) < 0.9), "v": v},
         [["k", "I64"]]),
    ]
    for i, (frame, keys) in enumerate(cases):
        plan = {"keys": keys, "aggs": aggs, "options": {"method": "sort"}}
        assert_results_equal(run_gpu(native, plan, frame), orc.groupby(plan, frame), [k[0] for k in keys],
                             label=f"edge{i}")
    # a predicate that keeps nothing
    frame = {"k": ("I64", rng.integers(0, 9, n), None), "v": v}
    plan = {"keys": [["k", "I64"]], "aggs": aggs, "options": {"method": "sort"},
            "predicate": {"kind": "Compare", "op": "Gt", "dtype": "F64", "lhs": {"kind": "Column", "name": "v", "dtype": "F64"},
                          "rhs": {"kind": "LiteralF64", "value": 1e300}}}
    got = run_gpu(native, plan, frame)
    assert got["__n_groups__"] == 0


def test_small_config_full_size(native):
    frame, plan = synth.small(1_000_000, seed=11)
    got = run_gpu(native, plan, frame)
    want = orc.groupby(plan, frame)
    assert want["__n_groups__"] == 1000
    assert_results_equal(got, want, ["key"], label="small")


@pytest.mark.parametrize("rows", [1, 63, 64, 65, 4097, 2_000_000])
def test_zipf_scaled(native, rows):
    frame, plan = synth.zipf(rows, seed=rows)
    print(f"zipf rows={rows}", flush=True)
    got = run_gpu(native, plan, frame)
    assert_results_equal(got, orc.groupby(plan, frame), ["key"], label=f"zipf{rows}")


@pytest.mark.parametrize("method", ["hash", "hash_global", "sort"])
def test_highcard_scaled(native, method):
    frame, plan = synth.highcard(1_000_000, seed=5, pool=250_000)
    plan = dict(plan, options={"method": method})
    got = run_gpu(native, plan, frame)
    want = orc.groupby(plan, frame)
    assert want["__n_groups__"] > 200_000
    assert_results_equal(got, want, ["a", "b"], label="highcard")


def test_partitioned_stats_and_fallback(native):
    frame, plan = synth.zipf(1_000_000, seed=31)
    got = run_gpu(native, dict(plan, options={"method": "hash"}), frame)
    st = native.last_stats()
    assert st["hot_keys"] > 100 and st["sample_rows"] > 1000 and 0 < st["tail_rows"] < st["n_selected"]
    want = orc.groupby(plan, frame)
    assert_results_equal(got, want, ["key"], label="partitioned")
    # tables too small for the tail: under "auto" the fallback table
    # overflows and the query re-runs on the global-table engine
    got = run_gpu(native, dict(plan, options={"method": "auto", "partitions": 1, "tail_slots": 16}), frame)
    assert_results_equal(got, want, ["key"], label="auto-rerun")


def test_partitioned_fallback_table(native):
    # ~2000 keys, 8 hot slots, one 16-slot partition table: most keys take
    # the fallback global table inside the partitioned engine
    rng = np.random.default_rng(77)
    n = 100_000
    frame = {"key": ("I64", rng.integers(0, 2000, n), rng.random(n) < 0.99),
             "v": ("F64", rng.standard_normal(n), rng.random(n) < 0.9),
             "w": ("I32", rng.integers(0, 1000, n).astype(np.int32), None)}
    plan = {"keys": [["key", "I64"]], "aggs": synth.FIVE_AGGS,
            "options": {"method": "hash", "lds_slots": 16, "partitions": 1, "tail_slots": 16}}
    assert_results_equal(run_gpu(native, plan, frame), orc.groupby(plan, frame), ["key"], label="fallback")


def test_resize_on_overflow(native):
    frame, plan = synth.zipf(300_000, seed=9)
    plan = dict(plan, options={"capacity_hint": 8, "method": "hash_global"})
    got = run_gpu(native, plan, frame)
    st = native.last_stats()
    assert st["n_passes"] > 1 and st["final_capacity"] > st["initial_capacity"]
    assert_results_equal(got, orc.groupby(plan, frame), ["key"], label="resize")


ENGINE_OPTS = [
    {"method": "hash_global", "lds_slots": 0}, {"method": "hash_global", "lds_slots": 64},
    {"method": "hash_global", "block": 256, "tile": 1000}, {"method": "hash_global", "grid": 3},
    {"method": "hash_global", "block": 64, "grid": 1},
    {"method": "hash"}, {"method": "hash", "lds_slots": 0}, {"method": "hash", "sample": 0},
    {"method": "hash", "partitions": 1}, {"method": "hash", "partitions": 32768},
    {"method": "hash", "partitions": 1024}, {"method": "hash", "partitions": 2048},
    {"method": "hash", "partitions": 2, "tail_slots": 16},
    {"method": "hash", "lds_slots": 16, "sample": 1000},
    {"method": "hash", "scan_rows": 8}, {"method": "hash", "scan_rows": 8, "lds_slots": 0},
    {"method": "hash", "fast_scan": 0},
]


@pytest.mark.parametrize("opts", ENGINE_OPTS, ids=[str(o) for o in ENGINE_OPTS])
def test_launch_shapes_agree(native, opts):
    frame, plan = synth.zipf(200_000, seed=4)
    plan = dict(plan, options=opts)
    assert_results_equal(run_gpu(native, plan, frame), orc.groupby(plan, frame), ["key"], label=str(opts))


@pytest.mark.parametrize("parts", [1024, 256, 2048])
def test_partition_layouts_with_many_tail_tiles(native, parts):
    """Enough tail rows for hundreds of partition tiles, so every XCD slot of
    the one-pass layout (up to 1024 partitions) and the stable multi-pass
    layout (more) carry many runs per partition."""
    frame, plan = synth.zipf(3_000_000, seed=8)
    plan = dict(plan, options={"method": "hash", "partitions": parts})
    got = run_gpu(native, plan, frame)
    st = native.last_stats()
    assert st["partitions"] == parts and st["tail_rows"] > 200_000
    assert_results_equal(got, orc.groupby(plan, frame), ["key"], label=f"parts{parts}")


KEY_TAGS = ["I64", "I32", "I16", "I8", "U8", "U16", "U32", "U64", "F64", "F32", "Bool"]


@pytest.mark.parametrize("method", ["hash", "hash_global", "sort"])
@pytest.mark.parametrize("tag", KEY_TAGS)
def test_key_dtypes_with_nulls(native, tag, method):
    rng = np.random.default_rng(KEY_TAGS.index(tag))
    n = 50_000
    if tag == "Bool":
        k = rng.random(n) < 0.5
    elif tag in ("F64", "F32"):
        k = rng.integers(-20, 20, n).astype(orc.NP_DTYPE[tag]) / 4
        k[::97] = -0.0
    else:
        info = np.iinfo(orc.NP_DTYPE[tag])
        k = rng.integers(max(info.min, -300), min(info.max, 300), n).astype(orc.NP_DTYPE[tag])
        k[::101] = info.max  # all-ones patterns for unsigned types
        k[::103] = info.min
    frame = {"k": (tag, k, rng.random(n) < 0.95), "v": ("F64", rng.standard_normal(n), rng.random(n) < 0.9),
             "w": ("I32", rng.integers(-50, 50, n).astype(np.int32), None)}
    plan = {"keys": [["k", tag]], "aggs": synth.FIVE_AGGS + [{"kind": "Length", "op": "Len", "output_alias": "n"}],
            "options": {"method": method}}
    assert_results_equal(run_gpu(native, plan, frame), orc.groupby(plan, frame), ["k"], label=tag)


def test_i64_all_ones_and_null_side_slots(native):
    k = np.array([-1, -1, 5, 0, -1, 5, 7], np.int64)
    valid = np.array([1, 1, 1, 0, 1, 0, 1], bool)
    frame = {"k": ("I64", k, valid), "v": ("F64", np.arange(7, dtype=np.float64), None)}
    plan = {"keys": [["k", "I64"]], "aggs": [{"kind": "Simple", "input_col": "v", "op": "Sum", "output_alias": "s"}]}
    for opts in ({}, {"lds_slots": 0}, {"method": "hash_global"}, {"method": "hash_global", "lds_slots": 0},
                 {"method": "hash", "partitions": 4, "tail_slots": 16}):
        p = dict(plan, options=opts)
        assert_results_equal(run_gpu(native, p, frame), orc.groupby(p, frame), ["k"], label=f"side{opts}")


@pytest.mark.parametrize("method", ["hash", "hash_global", "sort"])
@pytest.mark.parametrize("keys", [["a", "b"], ["b", "c", "a"], ["c", "d"], ["a", "c", "d"]])
def test_composite_keys(native, keys, method):
    rng = np.random.default_rng(len(keys))
    n = 100_000
    frame = {"a": ("I32", rng.integers(-5, 5, n).astype(np.int32), rng.random(n) < 0.97),
             "b": ("I64", rng.integers(-3, 3, n) * (1 << 50), rng.random(n) < 0.97),
             "c": ("Bool", rng.random(n) < 0.5, rng.random(n) < 0.9),
             "d": ("U16", rng.integers(0, 4, n).astype(np.uint16), None),
             "v": ("I64", rng.integers(-1 << 62, 1 << 62, n), rng.random(n) < 0.8)}
    tags = {k: frame[k][0] for k in keys}
    plan = {"keys": [[k, tags[k]] for k in keys],
            "aggs": [{"kind": "Simple", "input_col": "v", "op": op, "output_alias": op} for op in
                     ("Sum", "Count", "Min", "Max", "Mean")], "options": {"method": method}}
    assert_results_equal(run_gpu(native, plan, frame), orc.groupby(plan, frame), keys, label=str(keys))


@pytest.mark.parametrize("vtag", ["F64", "F32", "I64", "I32"])
def test_value_dtypes_nan_and_nulls(native, vtag):
    rng = np.random.default_rng(3)
    n = 80_000
    if vtag in ("F64", "F32"):
        v = (rng.standard_normal(n) * 1e3).astype(orc.NP_DTYPE[vtag])
        v[rng.random(n) < 0.001] = np.nan
        v[rng.random(n) < 0.001] = np.inf
        v[::1000] = np.ldexp(1.0, -1070)  # subnormal-range values take the residual path
    else:
        info = np.iinfo(orc.NP_DTYPE[vtag])
        v = rng.integers(info.min, info.max, n, endpoint=True).astype(orc.NP_DTYPE[vtag])
    frame = {"k": ("I64", rng.integers(0, 300, n), None), "x": (vtag, v, rng.random(n) < 0.9)}
    plan = {"keys": [["k", "I64"]],
            "aggs": [{"kind": "Simple", "input_col": "x", "op": op, "output_alias": op} for op in
                     ("Sum", "Count", "Min", "Max", "Mean")]}
    want = orc.groupby(plan, frame)
    got = run_gpu(native, plan, frame)
    if vtag == "F64":
        # groups holding an out-of-window value (here 2^-1070) sum their
        # residual in f64 and may round twice; their sums and means compare
        # within 1 ulp (min and max still by bits)
        assert_results_equal(got, want, ["k"], float_exact=False, rtol=4e-16, label=vtag)
    else:
        assert_results_equal(got, want, ["k"], label=vtag)


def test_predicate_shapes(native):
    rng = np.random.default_rng(8)
    n = 60_000
    frame = {"k": ("I32", rng.integers(0, 50, n).astype(np.int32), None),
             "x": ("F64", rng.standard_normal(n), rng.random(n) < 0.9),
             "y": ("I16", rng.integers(-100, 100, n).astype(np.int16), rng.random(n) < 0.9),
             "f": ("Bool", rng.random(n) < 0.5, rng.random(n) < 0.9),
             "z": ("F32", rng.standard_normal(n).astype(np.float32), None)}
    frame["x"][1][::50] = np.nan
    col = lambda n_, d: {"kind": "Column", "name": n_, "dtype": d}  # noqa: E731
    preds = [
        {"kind": "Compare", "op": "Le", "dtype": "F64", "lhs": col("x", "F64"), "rhs": {"kind": "LiteralF64", "value": 0.25}},
        {"kind": "Compare", "op": "Ne", "dtype": "I64", "lhs": col("y", "I64"), "rhs": {"kind": "LiteralI64", "value": 3}},
        {"kind": "Or", "lhs": col("f", "Bool"),
         "rhs": {"kind": "Compare", "op": "Gt", "dtype": "F64", "lhs": col("x", "F64"), "rhs": col("z", "F64")}},
        {"kind": "And", "lhs": {"kind": "Or", "lhs": col("f", "Bool"), "rhs": {"kind": "LiteralBool", "value": False}},
         "rhs": {"kind": "Compare", "op": "Eq", "dtype": "F64", "lhs": col("x", "F64"), "rhs": col("x", "F64")}},
        {"kind": "Compare", "op": "Lt", "dtype": "I64", "lhs": {"kind": "LiteralI64", "value": -20}, "rhs": col("y", "I64")},
    ]
    for i, p in enumerate(preds):
        plan = {"keys": [["k", "I32"]], "predicate": p,
                "aggs": [{"kind": "Simple", "input_col": "x", "op": "Sum", "output_alias": "s"},
                         {"kind": "Length", "op": "Len", "output_alias": "n"}]}
        assert_results_equal(run_gpu(native, plan, frame), orc.groupby(plan, frame), ["k"], label=f"pred{i}")


I64_MIN, I64_MAX = -(1 << 63), (1 << 63) - 1


def _fast_scan_frame():
    rng = np.random.default_rng(9)
    n = 70_000
    x = rng.standard_normal(n)
    x[::37] = np.nan
    x[::41] = 0.0
    x[::43] = -0.0
    x[::47] = np.inf
    x[::53] = -np.inf
    key = rng.integers(0, 400, n)
    key[::101] = -1  # the all-ones key lane
    return {"k": ("I64", key, rng.random(n) < 0.98),
            "x": ("F64", x, rng.random(n) < 0.9),
            "y": ("I16", rng.integers(-100, 100, n).astype(np.int16), rng.random(n) < 0.9),
            "f": ("Bool", rng.random(n) < 0.5, rng.random(n) < 0.9)}


def _cmp(op, lhs, rhs, dtype):
    return {"kind": "Compare", "op": op, "dtype": dtype, "lhs": lhs, "rhs": rhs}


def _c(name, dtype):
    return {"kind": "Column", "name": name, "dtype": dtype}


def _lf(v):
    return {"kind": "LiteralF64", "value": v}


def _li(v):
    return {"kind": "LiteralI64", "value": v}


def _and(a, b):
    return {"kind": "And", "lhs": a, "rhs": b}


# (predicate, lowered to intervals for the specialised scan)
FAST_SCAN_PREDS = [
    (_cmp("Gt", _c("x", "F64"), _lf(float("nan")), "F64"), True),    # nothing sorts above NaN
    (_cmp("Lt", _c("x", "F64"), _lf(float("nan")), "F64"), True),    # every non-NaN value
    (_cmp("Ge", _c("x", "F64"), _lf(float("nan")), "F64"), True),    # NaN only
    (_cmp("Ge", _c("x", "F64"), _lf(-0.0), "F64"), True),            # both zeros and up
    (_cmp("Eq", _c("x", "F64"), _lf(0.0), "F64"), True),             # -0.0 == +0.0
    (_cmp("Lt", _c("x", "F64"), _lf(float("-inf")), "F64"), True),   # empty
    (_cmp("Le", _c("x", "F64"), _lf(float("inf")), "F64"), True),    # all but NaN
    (_cmp("Lt", _c("y", "I64"), _li(I64_MIN), "I64"), True),        # empty
    (_cmp("Gt", _c("y", "I64"), _li(I64_MAX), "I64"), True),        # empty
    (_cmp("Ge", _li(5), _c("y", "I64"), "I64"), True),              # literal on the left
    (_cmp("Gt", _c("y", "F64"), _lf(2.5), "F64"), True),            # integer column, float compare
    (_and(_and(_cmp("Gt", _c("x", "F64"), _lf(-1.0), "F64"), _cmp("Le", _c("x", "F64"), _lf(1.0), "F64")),
          _and(_cmp("Ge", _c("y", "I64"), _li(-50), "I64"), _cmp("Lt", _c("y", "I64"), _li(50), "I64"))), True),
    (_and(_c("f", "Bool"), _cmp("Gt", _c("x", "F64"), _lf(0.0), "F64")), True),
    (_and(_cmp("Ge", _c("y", "I64"), _li(10), "I64"), _cmp("Ge", _c("y", "I64"), _li(20), "I64")), True),
    ({"kind": "LiteralBool", "value": True}, True),
    (_and(_and(_c("f", "Bool"), _cmp("Gt", _c("x", "F64"), _lf(0.0), "F64")),
          _cmp("Lt", _c("y", "I64"), _li(0), "I64")), False),            # three columns
    (_cmp("Ne", _c("y", "I64"), _li(3), "I64"), False),
]


@pytest.mark.parametrize("i", range(len(FAST_SCAN_PREDS)))
def test_fast_scan_predicates(native, i):
    pred, lowered = FAST_SCAN_PREDS[i]
    frame = _fast_scan_frame()
    plan = {"keys": [["k", "I64"]], "predicate": pred,
            "aggs": [{"kind": "Simple", "input_col": "x", "op": op, "output_alias": op} for op in
                     ("Sum", "Count", "Min", "Max", "Mean")] + [{"kind": "Length", "op": "Len", "output_alias": "n"}]}
    want = orc.groupby(plan, frame)
    print(f"fast-scan predicate {i}", flush=True)
    for method in ("hash", "sort"):
        for fast in (1, 0):
            p = dict(plan, options={"method": method, "fast_scan": fast, "lds_slots": 64})
            got = run_gpu(native, p, frame)
            if method == "hash":
                assert native.last_stats()["scan_kernel"] == ("specialised" if fast and lowered else "generic")
            assert_results_equal(got, want, ["k"], label=f"pred{i}/{method}/fast{fast}")


@pytest.mark.parametrize("keys", [["a"], ["a", "b"], ["b", "c"]])
def test_fast_scan_lane_keys(native, keys):
    # narrow keys build a lane of at most 64 bits: the specialised scan's
    # lane path, against the generic scan and the oracle
    rng = np.random.default_rng(10)
    n = 90_000
    frame = {"a": ("I32", rng.integers(-30, 30, n).astype(np.int32), rng.random(n) < 0.95),
             "b": ("U8", rng.integers(0, 256, n).astype(np.uint8), rng.random(n) < 0.95),
             "c": ("I16", np.full(n, -1, np.int16), None),
             "v": ("I64", rng.integers(-1 << 40, 1 << 40, n), rng.random(n) < 0.9),
             "w": ("I32", rng.integers(0, 1000, n).astype(np.int32), None)}
    plan = {"keys": [[k, frame[k][0]] for k in keys],
            "predicate": _cmp("Lt", _c("w", "I64"), _li(700), "I64"),
            "aggs": [{"kind": "Simple", "input_col": "v", "op": op, "output_alias": op} for op in
                     ("Sum", "Count", "Min", "Max", "Mean")]}
    want = orc.groupby(plan, frame)
    for fast in (1, 0):
        got = run_gpu(native, dict(plan, options={"method": "hash", "fast_scan": fast}), frame)
        assert native.last_stats()["scan_kernel"] == ("specialised" if fast else "generic")
        assert_results_equal(got, want, keys, label=f"{keys}/fast{fast}")


def test_deterministic_bits(native):
    frame, plan = synth.zipf(500_000, seed=21)
    wire = orc.to_wire(frame)
    a = native.execute_filter_groupby(plan, 500_000, wire)
    b = native.execute_filter_groupby(plan, 500_000, wire)
    ca = orc.canonical(orc.from_wire(a), ["key"])
    cb = orc.canonical(orc.from_wire(b), ["key"])
    for name in ca:
        if not name.startswith("__"):
            assert ca[name][1].tobytes() == cb[name][1].tobytes(), name


def test_filter_compact(native):
    rng = np.random.default_rng(12)
    for n in [0, 1, 9, 4096, 4097, 300_001]:
        pd = rng.random(n) < 0.4
        pv = rng.random(n) < 0.9
        cols = {"a": ("I64", rng.integers(-1 << 60, 1 << 60, n), rng.random(n) < 0.8),
                "b": ("F64", rng.standard_normal(n), None),
                "c": ("Bool", rng.random(n) < 0.5, rng.random(n) < 0.7),
                "d": ("I32", rng.integers(-5, 5, n).astype(np.int32), None),
                "e": ("U8", rng.integers(0, 255, n).astype(np.uint8), rng.random(n) < 0.5),
                # every other fixed width (the reference compacts only i64 / f64 / bool)
                "f": ("F32", rng.standard_normal(n).astype(np.float32), rng.random(n) < 0.8),
                "g": ("I8", rng.integers(-128, 127, n).astype(np.int8), None),
                "h": ("I16", rng.integers(-30000, 30000, n).astype(np.int16), rng.random(n) < 0.6),
                "i": ("U16", rng.integers(0, 65535, n).astype(np.uint16), None),
                "j": ("U32", rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32), rng.random(n) < 0.9),
                "l": ("U64", rng.integers(0, 1 << 63, n, dtype=np.uint64) * 2 + 1, None)}
        out = native.execute_filter_compact(orc.pack_bits(pd).tobytes(), orc.pack_bits(pv).tobytes(), n,
                                            orc.to_wire(cols))
        keep = pd & pv
        m = int(keep.sum())
        for (name, (tag, vals, valid)), (data, vb, n_out) in zip(cols.items(), out):
            assert n_out == m
            if m == 0:
                assert data == b"" and vb == b""
                continue
            ev, evv = orc.compact(keep, tag, vals, valid)
            if tag == "Bool":
                np.testing.assert_array_equal(orc.unpack_bits(data, m), ev)
            else:
                np.testing.assert_array_equal(np.frombuffer(data, orc.NP_DTYPE[tag]), ev)
            np.testing.assert_array_equal(orc.unpack_bits(vb, m), evv)
            assert len(vb) == max(4, ((m + 7) // 8 + 3) & ~3)


@pytest.mark.parametrize("op", ["Eq", "Ne", "Lt", "Le", "Gt", "Ge"])
def test_cmp_and_bool(native, op):
    rng = np.random.default_rng(ord(op[0]))
    n = 10_007
    a = rng.integers(-3, 3, n)
    b = rng.integers(-3, 3, n)
    av = rng.random(n) < 0.9
    bv = rng.random(n) < 0.9
    x = rng.choice(np.array([np.nan, -0.0, 0.0, 1.0, -np.inf, 2.0]), n)
    y = rng.choice(np.array([np.nan, 0.0, 1.0, np.inf]), n)
    frame = {"a": ("I64", a, av), "b": ("I64", b, bv), "x": ("F64", x, av), "y": ("F64", y, bv)}
    pk = lambda m: orc.pack_bits(m).tobytes()  # noqa: E731
    col = lambda n_, d: {"kind": "Column", "name": n_, "dtype": d}  # noqa: E731
    checks = [
        (native.cmp_i64_col_scalar(a.tobytes(), pk(av), 1, op, n), {"kind": "Compare", "op": op, "dtype": "I64",
                                                                    "lhs": col("a", "I64"), "rhs": {"kind": "LiteralI64", "value": 1}}),
        (native.cmp_i64_col_col(a.tobytes(), pk(av), b.tobytes(), pk(bv), op, n),
         {"kind": "Compare", "op": op, "dtype": "I64", "lhs": col("a", "I64"), "rhs": col("b", "I64")}),
        (native.cmp_f64_col_scalar(x.tobytes(), pk(av), 0.0, op, n), {"kind": "Compare", "op": op, "dtype": "F64",
                                                                      "lhs": col("x", "F64"), "rhs": {"kind": "LiteralF64", "value": 0.0}}),
        (native.cmp_f64_col_col(x.tobytes(), pk(av), y.tobytes(), pk(bv), op, n),
         {"kind": "Compare", "op": op, "dtype": "F64", "lhs": col("x", "F64"), "rhs": col("y", "F64")}),
    ]
    for (gd, gv), pred in checks:
        d, v = orc.eval_predicate(pred, frame)
        np.testing.assert_array_equal(orc.unpack_bits(gv, n), v)
        np.testing.assert_array_equal(orc.unpack_bits(gd, n), d & v)
    for kind, fn in (("And", native.bool_and_dispatch), ("Or", native.bool_or_dispatch)):
        p = a > 0
        q = b > 0
        gd, gv = fn(pk(p), pk(av), pk(q), pk(bv), n)
        fb = {"p": ("Bool", p, av), "q": ("Bool", q, bv)}
        d, v = orc.eval_predicate({"kind": kind, "lhs": col("p", "Bool"), "rhs": col("q", "Bool")}, fb)
        np.testing.assert_array_equal(orc.unpack_bits(gv, n), v)
        np.testing.assert_array_equal(orc.unpack_bits(gd, n), d & v)


def test_device_buffers_round_trip_and_copy(native):
    """pmh_buffer_*: host -> device -> device copy -> host, zeroed
    allocation, and a copy larger than its buffer refused."""
    from polars_metal.buffer import DeviceBuffer
    rng = np.random.default_rng(31)
    a = rng.integers(0, 256, 1_000_003, dtype=np.uint8)
    src = DeviceBuffer.from_numpy(a)
    dst = DeviceBuffer.alloc(a.nbytes)
    dst.copy_from(src)
    assert np.array_equal(dst.to_numpy(np.uint8), a)
    z = DeviceBuffer.alloc(4096, zeroed=True)
    assert not z.to_numpy(np.uint8).any()
    with pytest.raises(ValueError):
        z.copy_from(src)
    for b in (src, dst, z):
        b.free()


# Aggregate shapes of the specialised scan's fold, one per compile-time
# signature (the first four) and one resolved at run time.
SCAN_SIG_AGGS = [
    [("v", "Sum"), ("v", "Mean"), ("w", "Min"), ("w", "Max")],  # f64 sum + i32 min/max
    [("a", "Sum")],                                              # i64 sum
    [("v", "Sum"), ("v", "Min"), ("v", "Max")],                  # f64 sum + min/max
    [("v", "Sum")],                                              # f64 sum
    [("a", "Sum"), ("a", "Min"), ("w", "Max")],                  # no signature
]


@pytest.mark.parametrize("key", ["I64", "I32"])
@pytest.mark.parametrize("two_pred", [False, True])
@pytest.mark.parametrize("sig", range(len(SCAN_SIG_AGGS)))
def test_scan_signatures_hot_and_tail(native, sig, two_pred, key):
    """Every instantiation of the specialised scan (raw 64-bit key or lane
    key, one or two predicate columns, each aggregate signature) on skewed
    keys, so rows fold into the hot table and spill to the partitions."""
    rng = np.random.default_rng(100 + sig)
    n = 300_000
    k = rng.zipf(1.2, n) % 40_000
    frame = {"k": (key, k.astype(np.int64 if key == "I64" else np.int32), None),
             "v": ("F64", rng.standard_normal(n), rng.random(n) < 0.95),
             "w": ("I32", rng.integers(0, 1000, n).astype(np.int32), None),
             "a": ("I64", rng.integers(-10**12, 10**12, n), rng.random(n) < 0.9)}
    pred = _cmp("Ge", _c("w", "I32"), _li(100), "I64")
    if two_pred:
        pred = _and(pred, _cmp("Gt", _c("v", "F64"), _lf(-1.5), "F64"))
    plan = {"keys": [["k", key]], "predicate": pred,
            "aggs": [{"kind": "Simple", "input_col": c, "op": op, "output_alias": f"{c}_{op}"}
                     for c, op in SCAN_SIG_AGGS[sig]],
            "options": {"method": "hash"}}
    want = orc.groupby(plan, frame)
    print(f"scan signature {sig} two_pred={two_pred} key={key}", flush=True)
    got = run_gpu(native, plan, frame)
    st = native.last_stats()
    assert st["scan_kernel"] == "specialised"
    assert st["hot_keys"] > 0 and st["tail_rows"] > 0
    assert_results_equal(got, want, ["k"], label=f"sig{sig}/{two_pred}/{key}")


EXPR_ENGINES = [{"method": "hash"}, {"method": "hash", "fast_scan": 0}, {"method": "sort"},
                {"method": "hash_global"}, {"method": "auto"}]


@pytest.mark.parametrize("opts", EXPR_ENGINES, ids=[str(o) for o in EXPR_ENGINES])
def test_expression_aggs_random(native, opts):
    """Expression aggregates over mixed dtypes with nulls, under a filter,
    on every engine: bit-exact against the oracle's evaluation."""
    rng = np.random.default_rng(21)
    n = 400_000
    frame = {"key": ("I64", rng.integers(0, 5000, n), rng.random(n) < 0.98),
             "a": ("F64", rng.standard_normal(n) * 100, rng.random(n) < 0.95),
             "b": ("F64", rng.random(n), None),
             "i": ("I32", rng.integers(-2**31, 2**31 - 1, n).astype(np.int32), rng.random(n) < 0.9),
             "j": ("I32", rng.integers(-50000, 50000, n).astype(np.int32), None),
             "f": ("F32", rng.standard_normal(n).astype(np.float32), rng.random(n) < 0.97),
             "h": ("I16", rng.integers(-30000, 30000, n).astype(np.int16), None),
             "u": ("U32", rng.integers(0, 2**32 - 1, n, dtype=np.uint64).astype(np.uint32), None)}
    col = lambda x: {"kind": "Column", "name": x}  # noqa: E731
    b = lambda op, x, y: {"kind": "Binary", "op": op, "lhs": x, "rhs": y}  # noqa: E731
    exprs = {
        "disc": b("Mul", col("a"), b("Sub", {"kind": "LiteralI64", "value": 1}, col("b"))),  # F64
        "iwrap": b("Add", b("Mul", col("i"), col("j")), {"kind": "LiteralI64", "value": 7}),  # I32, wraps
        "idiv": b("Div", col("i"), col("j")),                                                 # F64
        "f32": b("Add", b("Mul", col("f"), {"kind": "LiteralF64", "value": 0.1}), col("h")),  # F32
        "iu": b("Sub", col("i"), col("u")),                                                   # I64
        "deep": b("Div", b("Sub", b("Mul", b("Add", col("a"), col("f")), col("j")), col("b")),
                  {"kind": "LiteralI64", "value": 3}),                                        # F64, depth 4
    }
    aggs = []
    for name, e in exprs.items():
        for op in ("Sum", "Min", "Max", "Mean", "Count"):
            aggs.append({"kind": "Expression", "op": op, "output_alias": f"{name}_{op.lower()}", "expr": e})
    pred = {"kind": "Compare", "op": "Gt", "dtype": "F64", "lhs": {"kind": "Column", "name": "b", "dtype": "F64"},
            "rhs": {"kind": "LiteralF64", "value": 0.3}}
    want = orc.groupby({"keys": [["key", "I64"]], "aggs": aggs, "predicate": pred}, frame)
    print(f"expression aggs {opts}", flush=True)
    for lo in range(0, len(aggs), 10):  # at most eight aggregated columns per query
        part = aggs[lo:lo + 10]
        plan = {"keys": [["key", "I64"]], "aggs": part, "predicate": pred, "options": opts}
        got = run_gpu(native, plan, frame)
        sub = {k: v for k, v in want.items() if k == "key" or k.startswith("__") or k in {a["output_alias"] for a in part}}
        assert_results_equal(got, sub, ["key"], label=f"expr/{opts}/{lo}")


SORT_BUCKET_CASES = ["highcard", "wide", "zipf_skew", "i32_small"]


@pytest.mark.parametrize("case", SORT_BUCKET_CASES)
def test_sort_bucket_finish(native, case):
    """The sort engine's MSD finish (radix passes over the key's top bits,
    most significant digit first or stable LSD, then each bucket sorted in
    LDS) against the oracle and against whole-key LSD passes; skewed keys
    overflow a bucket and fall back to the LSD passes."""
    rng = np.random.default_rng(SORT_BUCKET_CASES.index(case) + 40)
    if case == "highcard":
        frame, plan = synth.highcard(1_500_000, seed=41, pool=700_000)
        keys = ["a", "b"]
    elif case == "wide":  # two random full-range halves: a compressed key of more than 64 bits
        n = 600_000
        frame = {"a": ("I64", rng.integers(-2**62, 2**62, n), rng.random(n) < 0.97),
                 "b": ("I32", rng.integers(-2**31, 2**31 - 1, n).astype(np.int32), None),
                 "v": ("F64", rng.standard_normal(n), rng.random(n) < 0.9)}
        plan = {"keys": [["a", "I64"], ["b", "I32"]], "aggs": [
            {"kind": "Simple", "input_col": "v", "op": op, "output_alias": op} for op in ("Sum", "Count", "Min", "Max")]}
        keys = ["a", "b"]
    elif case == "zipf_skew":
        frame, plan = synth.zipf(1_000_000, seed=42)
        keys = ["key"]
    else:
        n = 400_000
        # no null keys: the null group (one key, 5% of the rows) would fill
        # one bucket past the LDS sort, and the skew probe keeps whole-key passes
        frame = {"k": ("I32", rng.integers(-(1 << 30), 1 << 30, n).astype(np.int32), None),
                 "v": ("I64", rng.integers(-1 << 40, 1 << 40, n), None)}
        plan = {"keys": [["k", "I32"]], "aggs": [
            {"kind": "Simple", "input_col": "v", "op": op, "output_alias": op} for op in ("Sum", "Min", "Mean")]}
        keys = ["k"]
    want = orc.groupby(plan, frame)
    print(f"sort bucket finish {case}", flush=True)
    got = run_gpu(native, dict(plan, options={"method": "sort"}), frame)
    st_b = native.last_stats()
    got_lsd = run_gpu(native, dict(plan, options={"method": "sort", "sort_buckets": 0}), frame)
    st_l = native.last_stats()
    assert st_b["method"] == st_l["method"] == "sort"
    if case != "zipf_skew":
        assert st_b["n_passes"] < st_l["n_passes"], (st_b, st_l)  # fewer radix passes with the bucket finish
    else:
        assert st_b["n_passes"] == st_l["n_passes"]  # the skew probe keeps hot buckets out: whole-key passes
    assert_results_equal(got, want, keys, label=f"sortb/{case}")
    assert_results_equal(got_lsd, want, keys, label=f"sortl/{case}")
    # records that carry row indices (the fold gathers the values) instead of values
    got_rows = run_gpu(native, dict(plan, options={"method": "sort", "sort_carry": 0}), frame)
    assert native.last_stats()["n_passes"] == st_b["n_passes"]
    assert_results_equal(got_rows, want, keys, label=f"sortr/{case}")
    # buckets written back sorted and folded by a later pass, not in LDS
    got_wb = run_gpu(native, dict(plan, options={"method": "sort", "sort_fold": 0}), frame)
    assert native.last_stats()["n_passes"] == st_b["n_passes"]
    assert_results_equal(got_wb, want, keys, label=f"sortw/{case}")
    # top-bit passes as stable LSD passes (not most significant digit first)
    got_s = run_gpu(native, dict(plan, options={"method": "sort", "sort_msd": 0}), frame)
    assert native.last_stats()["n_passes"] == st_b["n_passes"]
    assert_results_equal(got_s, want, keys, label=f"sorts/{case}")


@pytest.mark.parametrize("case", ["high_bit", "low_bit", "null_key", "none"])
def test_sort_sampled_bits(native, case):
    """A large query without a predicate takes its varying key bits from an
    evenly spaced sample of 65536 rows; the encode checks every record and a
    key with a bit the sample did not see vary (one row off the sample grid:
    a high bit, a low bit, a null) sends the query through the exact pass.
    Results match the oracle and the exact pass either way."""
    rng = np.random.default_rng(["high_bit", "low_bit", "null_key", "none"].index(case) + 70)
    n = (1 << 21) + 3
    k = rng.integers(0, 1 << 18, n) * 2  # even keys: bit 0 of the data is constant
    kv = np.ones(n, bool)
    if case == "high_bit":
        k[5] = 1 << 40
    elif case == "low_bit":
        k[7] += 1
    elif case == "null_key":
        kv[9] = False
    frame = {"k": ("I64", k, kv), "v": ("I64", rng.integers(-1000, 1000, n), None)}
    plan = {"keys": [["k", "I64"]], "aggs": [{"kind": "Simple", "input_col": "v", "op": op, "output_alias": op}
                                             for op in ("Sum", "Count", "Max")]}
    want = orc.groupby(plan, frame)
    print(f"sort sampled bits {case}", flush=True)
    got = run_gpu(native, dict(plan, options={"method": "sort"}), frame)
    got_x = run_gpu(native, dict(plan, options={"method": "sort", "sort_sample_bits": 0}), frame)
    assert_results_equal(got, want, ["k"], label=f"sampled/{case}")
    assert_results_equal(got_x, want, ["k"], label=f"exact/{case}")


@pytest.mark.parametrize("carry", [1, 0])
def test_sort_bucket_fold_edges(native, carry):
    """The bucket fold (each LDS-sorted bucket's groups folded in place and
    stored once, with no initialised state) on the values its stores must
    carry: exact f64 sums with values outside the limb window (+-1e300,
    +-Inf, NaN, subnormals), signed zeros and NaN in min/max, all-null
    groups (count 0: sum 0, mean/min/max null), null keys, and groups of
    one row, against the oracle and the written-back fold."""
    rng = np.random.default_rng(91 + carry)
    n = 1_200_000
    k = rng.integers(0, 300_000, n)
    # huge values with normal ones, subnormals with eighths: the residual
    # then decides the correctly rounded sum whatever the order (a group
    # mixing 1e300, -1e300 and a subnormal would not be)
    v = rng.standard_normal(n)
    u = rng.random(n)
    for lo, hi, x in ((0.000, 0.004, 1e300), (0.004, 0.008, -1e300), (0.008, 0.009, np.inf),
                      (0.009, 0.010, -np.inf), (0.010, 0.011, np.nan), (0.011, 0.021, 0.0), (0.021, 0.031, -0.0)):
        v[(u >= lo) & (u < hi)] = x
    w = rng.integers(-1000, 1000, n).astype(np.float64) / 8
    u = rng.random(n)
    w[u < 0.01] = np.nan
    w[(u >= 0.01) & (u < 0.02)] = 5e-324 * rng.integers(1, 50, int(((u >= 0.01) & (u < 0.02)).sum()))
    w[(u >= 0.02) & (u < 0.025)] = -1e-310
    # keys >= 290000 never have a valid v: all-null groups
    vv = (rng.random(n) < 0.9) & (k < 290_000)
    # few null keys: the null group sorts into the first bucket, which must
    # stay within the LDS sort's capacity for the fold to run
    frame = {"k": ("I64", k, rng.random(n) < 0.999), "v": ("F64", v, vv), "w": ("F64", w, None)}
    aggs = [{"kind": "Simple", "input_col": "v", "op": op, "output_alias": f"v_{op}"}
            for op in ("Sum", "Mean", "Min", "Max", "Count")]
    aggs += [{"kind": "Simple", "input_col": "w", "op": op, "output_alias": f"w_{op}"} for op in ("Sum", "Max", "Mean")]
    aggs.append({"kind": "Length", "op": "Len", "output_alias": "n"})
    plan = {"keys": [["k", "I64"]], "aggs": aggs}
    want = orc.groupby(plan, frame)
    print(f"sort bucket fold edges carry={carry}", flush=True)
    got = run_gpu(native, dict(plan, options={"method": "sort", "sort_carry": carry}), frame)
    st = native.last_stats()
    got_wb = run_gpu(native, dict(plan, options={"method": "sort", "sort_carry": carry, "sort_fold": 0}), frame)
    st_wb = native.last_stats()
    assert st["method"] == st_wb["method"] == "sort" and st["n_passes"] == st_wb["n_passes"]
    assert st["n_passes"] <= 2  # the top-bit passes and the bucket finish, not the whole key's three
    assert_results_equal(got, want, ["k"], label=f"fold/{carry}")
    assert_results_equal(got_wb, want, ["k"], label=f"foldwb/{carry}")
