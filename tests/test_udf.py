"""The engine hook's group-by dispatch over Arrow tables (polars_metal.udf)
and polars_metal.install.  Host-only checks run without a GPU; the parity
checks (marked gpu) compare dispatch_groupby with the oracle."""
import types

import numpy as np
import pyarrow as pa
import pytest

import oracle as orc
from helpers import assert_results_equal
from polars_metal import udf

AGGS = [{"kind": "Simple", "input_col": "v", "op": op, "output_alias": f"v_{op.lower()}"}
        for op in ("Sum", "Count", "Min", "Max", "Mean")] + [{"kind": "Length", "op": "Len", "output_alias": "n"},
                                                            {"kind": "Simple", "input_col": "w", "op": "Max",
                                                             "output_alias": "w_max"}]


def _cmp(op, name, dtype, lit):
    kind = "LiteralF64" if dtype == "F64" else "LiteralI64"
    return {"kind": "Compare", "op": op, "dtype": dtype, "lhs": {"kind": "Column", "name": name, "dtype": dtype},
            "rhs": {"kind": kind, "value": lit}}


def test_resolve_upstream_shapes():
    p1, p2, p3 = _cmp("Gt", "v", "F64", 0.0), _cmp("Lt", "w", "I64", 5), _cmp("Ge", "w", "I64", 1)
    scan = {"kind": "Scan", "columns": []}
    assert udf.resolve_upstream(scan) == (None, None)
    assert udf.resolve_upstream(dict(scan, selection=p1)) == (p1, None)
    # Filter over a scan with a pushed-down selection: both hold
    got, _ = udf.resolve_upstream({"kind": "Filter", "predicate": p2, "input": dict(scan, selection=p1)})
    assert got == {"kind": "And", "lhs": p1, "rhs": p2}
    got, proj = udf.resolve_upstream({"kind": "Project", "columns": ["k", "v"],
                                      "input": {"kind": "Filter", "predicate": p3,
                                                "input": {"kind": "Filter", "predicate": p2, "input": scan}}})
    assert got == {"kind": "And", "lhs": p2, "rhs": p3} and proj == ["k", "v"]
    with pytest.raises(NotImplementedError):
        udf.resolve_upstream({"kind": "HStack", "input": scan})


DISC = {"kind": "Binary", "op": "Mul", "lhs": {"kind": "Column", "name": "v"},
        "rhs": {"kind": "Binary", "op": "Sub", "lhs": {"kind": "LiteralI64", "value": 1},
                "rhs": {"kind": "Column", "name": "w"}}}


def test_expression_aggs_stage_the_columns_they_read():
    # the AggExpr tree's columns are passed to the fused call, which
    # evaluates the tree on the device (no temporary host columns)
    plan = {"keys": [["k", "I64"]], "aggs": [{"kind": "Expression", "op": "Sum", "output_alias": "s", "expr": DISC},
                                             {"kind": "Simple", "input_col": "k", "op": "Count", "output_alias": "c"}]}
    assert udf._needed_columns(plan, _cmp("Gt", "x", "F64", 0.0)) == ["k", "v", "w", "x"]


@pytest.mark.gpu
def test_dispatch_groupby_expression_aggs(native):
    """sum(v * (1 - w)) and friends (the reference's Q1 shape,
    tests/python_integration/test_walker_expression_unfolding.py:38-56)
    under a pushed-down selection, against the oracle's evaluation."""
    rng = np.random.default_rng(11)
    n = 50_000
    frame = _table(rng, n)
    table = _pa(frame)
    aggs = [{"kind": "Expression", "op": op, "output_alias": f"d_{op.lower()}", "expr": DISC}
            for op in ("Sum", "Mean", "Min", "Max", "Count")] + AGGS[:2]
    sel = _cmp("Gt", "v", "F64", -0.5)
    plan = {"kind": "GroupBy", "input": {"kind": "Scan", "selection": sel}, "keys": [["k", "I64"]], "aggs": aggs}
    got = _from_table(udf.dispatch_groupby(table, plan))
    want = orc.groupby({"keys": [["k", "I64"]], "aggs": aggs, "predicate": sel}, frame)
    assert got["d_sum"][0] == "F64"
    assert_results_equal(got, want, ["k"], label="expr")


def test_install_rebinds_the_reference_hook():
    import polars_metal
    from polars_metal import _native
    calls = []

    def reference_dispatch(df_pydf, wire_plan):
        calls.append(wire_plan)
        return "reference"

    ref = types.SimpleNamespace(_native=object(), _dispatch_groupby=reference_dispatch)
    polars_metal.install(ref)
    assert ref._native is _native and ref._native_reference is not _native
    assert ref._dispatch_groupby is not reference_dispatch
    assert ref._dispatch_groupby_reference is reference_dispatch
    # a shape this backend does not lower goes to the reference's dispatch
    plan = {"kind": "GroupBy", "input": {"kind": "HStack", "input": {"kind": "Scan"}}, "keys": [], "aggs": []}
    assert ref._dispatch_groupby(None, plan) == "reference" and calls == [plan]
    before = ref._dispatch_groupby
    polars_metal.install(ref)  # idempotent
    assert ref._dispatch_groupby is before and ref._dispatch_groupby_reference is reference_dispatch
    with pytest.raises(TypeError):
        polars_metal.install(types.SimpleNamespace())


def test_arrow_tags():
    assert udf.arrow_tag(pa.int64()) == "I64" and udf.arrow_tag(pa.large_string()) == "Utf8"
    assert udf.arrow_tag(pa.date32()) == "I32" and udf.arrow_tag(pa.list_(pa.int8())) is None


def _table(rng, n, nulls=True):
    k = rng.integers(0, 300, n)
    v = rng.standard_normal(n)
    v[::97] = np.nan
    w = rng.integers(-50, 50, n).astype(np.int32)
    kv = rng.random(n) < 0.97 if nulls else None
    vv = rng.random(n) < 0.9 if nulls else None
    return {"k": ("I64", k, kv), "v": ("F64", v, vv), "w": ("I32", w, None)}


def _pa(frame, names=None):
    arrs = {}
    for name, (tag, vals, ok) in frame.items():
        mask = None if ok is None else ~np.asarray(ok)
        arrs[name] = pa.array(vals, type=None if tag != "I32" else pa.int32(), mask=mask)
    return pa.table(arrs)


def _from_table(t):
    out = {"__n_groups__": t.num_rows}
    for name in t.column_names:
        a = t.column(name).combine_chunks() if t.column(name).num_chunks != 1 else t.column(name).chunk(0)
        tag = udf.arrow_tag(a.type) if a.type != pa.uint32() else "U32"
        ok = ~np.asarray(a.is_null().to_numpy(zero_copy_only=False))
        if tag == "Utf8":
            vals = np.array([x if x is not None else "" for x in a.to_pylist()], object)
        else:
            vals = a.fill_null(0).to_numpy(zero_copy_only=False).astype(orc.NP_DTYPE[tag])
        out[name] = (tag, vals, ok)
    return out


def _slice_frame(frame, lo, hi):
    return {k: (t, v[lo:hi], None if ok is None else ok[lo:hi]) for k, (t, v, ok) in frame.items()}


@pytest.mark.gpu
@pytest.mark.parametrize("shape", ["scan", "selection", "filter_filter", "project_filter", "sliced", "chunked"])
def test_dispatch_groupby_matches_oracle(native, shape):
    rng = np.random.default_rng(["scan", "selection", "filter_filter", "project_filter", "sliced",
                                 "chunked"].index(shape))
    n = 60_000
    frame = _table(rng, n)
    p1, p2 = _cmp("Gt", "v", "F64", -0.5), _cmp("Lt", "w", "I64", 30)
    scan = {"kind": "Scan", "columns": [["k", "I64"], ["v", "F64"], ["w", "I32"]]}
    table = _pa(frame)
    pred = None
    if shape == "selection":
        inp, pred = dict(scan, selection=p1), p1
    elif shape == "filter_filter":
        inp = {"kind": "Filter", "predicate": p2, "input": {"kind": "Filter", "predicate": p1, "input": scan}}
        pred = {"kind": "And", "lhs": p1, "rhs": p2}
    elif shape == "project_filter":
        inp = {"kind": "Project", "columns": ["k", "v", "w"], "input": {"kind": "Filter", "predicate": p2,
                                                                        "input": scan}}
        pred = p2
    else:
        inp = scan
    if shape == "sliced":
        # Arrow slices keep the parent's buffers with an offset (13: inside a
        # validity byte), which the reference re-bases with a copy
        table = table.slice(13, n - 40)
        frame = _slice_frame(frame, 13, n - 27)
    if shape == "chunked":
        parts = [_pa(_slice_frame(frame, a, b)) for a, b in ((0, 7), (7, 20_001), (20_001, n))]
        table = pa.concat_tables(parts)
        assert table.column("k").num_chunks == 3
    plan = {"kind": "GroupBy", "input": inp, "keys": [["k", "I64"]], "aggs": AGGS}
    got = _from_table(udf.dispatch_groupby(table, plan))
    oplan = {"keys": [["k", "I64"]], "aggs": AGGS, **({"predicate": pred} if pred else {})}
    assert_results_equal(got, orc.groupby(oplan, frame), ["k"], label=shape)


@pytest.mark.gpu
def test_dispatch_groupby_utf8_keys(native):
    """Utf8 keys with nulls and the empty string: null is a group of its
    own and "" is a value (crates/polars-metal-buffer/src/dict.rs:40-66);
    results compare by string after sorting."""
    rng = np.random.default_rng(5)
    n = 20_000
    words = np.array(["", "a", "ab", "abc", "Zürich", "x" * 40, "ab "], object)
    keys = words[rng.integers(0, len(words), n)]
    kok = rng.random(n) < 0.9
    v = rng.standard_normal(n)
    table = pa.table({"s": pa.array(list(keys), pa.string(), mask=~kok), "v": pa.array(v),
                      "w": pa.array(rng.integers(0, 9, n).astype(np.int32))})
    plan = {"kind": "GroupBy", "input": {"kind": "Scan"}, "keys": [["s", "Utf8"]], "aggs": AGGS}
    got = udf.dispatch_groupby(table, plan).to_pydict()
    want = {}
    for i in range(n):
        key = keys[i] if kok[i] else None
        g = want.setdefault(key, {"n": 0, "vals": [], "w": []})
        g["n"] += 1
        g["vals"].append(v[i])
        g["w"].append(int(table.column("w")[i].as_py()))
    assert sorted(got["s"], key=lambda x: (x is not None, x or "")) == sorted(want, key=lambda x: (x is not None, x or ""))
    import math
    for i, key in enumerate(got["s"]):
        g = want[key]
        assert got["n"][i] == g["n"] and got["v_count"][i] == g["n"]
        assert got["v_sum"][i] == math.fsum(g["vals"]) and got["w_max"][i] == max(g["w"])
        assert got["v_min"][i] == min(g["vals"]) and got["v_max"][i] == max(g["vals"])
    # the _native bytes path (the reference's packed Utf8 wire form) agrees
    from polars_metal import _native
    import struct
    enc = pa.array(list(keys), pa.string(), mask=~kok)
    offs = np.frombuffer(enc.buffers()[1], np.int32, count=n + 1)
    wire = struct.pack("<I", n) + offs.tobytes() + bytes(enc.buffers()[2])[:offs[-1]]
    out = _native.execute_groupby({"keys": [["s", "Utf8"]], "aggs": AGGS[:2]},
                                  n, [("s", "Utf8", wire, bytes(enc.buffers()[0])),
                                      ("v", "F64", v.tobytes(), b"\xff" * ((n + 7) // 8))])
    assert out[0][1] == "Utf8"
    (m,) = struct.unpack_from("<I", out[0][2], 0)
    assert m == len(want)
