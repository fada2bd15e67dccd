"""C-ABI library: loads, exports every symbol include/pmh.h declares, and its
host-only logic (plan parser errors, router) behaves like the reference's.
No compute calls: these run without a GPU."""
import ctypes
import re
from pathlib import Path

import pytest

from polars_metal import _lib

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "pmh.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pmh_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 18
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(_lib.SIGNATURES), set(syms) ^ set(_lib.SIGNATURES)


def test_version_string():
    from polars_metal import _native
    assert _native.version_string().startswith("0.")


def _gb_plan(**kw):
    p = {"keys": [["k", "I64"]], "aggs": [{"kind": "Simple", "input_col": "v", "op": "Sum", "output_alias": "s"}]}
    p.update(kw)
    return p


COLS = [("k", "I64", b"\x00" * 8, b"\x01"), ("v", "F64", b"\x00" * 8, b"\x01")]


@pytest.mark.parametrize("plan,exc,msg", [
    ({"aggs": []}, ValueError, "plan parse error: missing field keys"),
    ({"keys": [["k", "I64"]]}, ValueError, "plan parse error: missing field aggs"),
    ({"keys": [["k", "XX"]], "aggs": []}, ValueError, "plan parse error: unknown dtype"),
    (_gb_plan(aggs=[{"kind": "Simple", "input_col": "v", "op": "Median", "output_alias": "m"}]), ValueError,
     "unknown agg op"),
    (_gb_plan(aggs=[{"kind": "Simple", "input_col": "v", "op": "Sum"}]), ValueError, "output_alias"),
    (_gb_plan(keys=[["missing", "I64"]]), KeyError, "column not found"),
    (_gb_plan(aggs=[{"kind": "Simple", "input_col": "k", "op": "Sum", "output_alias": "s"}],
              keys=[["v", "I64"]]), ValueError, "dtype mismatch"),
])
def test_groupby_plan_errors(plan, exc, msg):
    from polars_metal import _native
    with pytest.raises(exc, match=msg):
        _native.execute_groupby(plan, 1, COLS)


def test_unsupported_agg_dtype_is_value_error():
    from polars_metal import _native
    cols = [("k", "I64", b"\x00" * 8, b"\x01"), ("u", "U16", b"\x00" * 2, b"\x01")]
    plan = {"keys": [["k", "I64"]], "aggs": [{"kind": "Simple", "input_col": "u", "op": "Sum", "output_alias": "s"}]}
    with pytest.raises(ValueError, match=r"unsupported \(agg_op, dtype\) combination: Sum / U16"):
        _native.execute_groupby(plan, 1, cols)


def _xplan(expr, op="Sum"):
    return {"keys": [["k", "I64"]], "aggs": [{"kind": "Expression", "op": op, "output_alias": "e", "expr": expr}]}


def _col(n):
    return {"kind": "Column", "name": n}


def _bin(op, a, b):
    return {"kind": "Binary", "op": op, "lhs": a, "rhs": b}


XCOLS = COLS + [("i8", "I8", b"\x00", b"\x01"), ("u64", "U64", b"\x00" * 8, b"\x01"),
                ("b", "Bool", b"\x00", b"\x01"), ("i32", "I32", b"\x00" * 4, b"\x01")]


@pytest.mark.parametrize("expr,op,exc,msg", [
    # the reference's parser: depth cap of four binary levels (plan/mod.rs:169-178), op names
    (_bin("Add", _bin("Add", _bin("Add", _bin("Add", _bin("Add", _col("v"), {"kind": "LiteralF64", "value": 1.0}),
                                                  {"kind": "LiteralF64", "value": 1.0}),
                                      {"kind": "LiteralF64", "value": 1.0}),
                          {"kind": "LiteralF64", "value": 1.0}), {"kind": "LiteralF64", "value": 1.0}),
     "Sum", ValueError, "too deep"),
    (_bin("Pow", _col("v"), _col("v")), "Sum", ValueError, "unknown agg op binary op Pow"),
    ({"kind": "Sqrt"}, "Sum", ValueError, "expr.kind"),
    ({"kind": "LiteralI64", "value": 1.5}, "Sum", ValueError, "expr.value(i64)"),
    # dtype rules refused before any device work
    (_bin("Add", _col("b"), _col("v")), "Sum", NotImplementedError, "Bool"),
    (_bin("Add", _col("k"), _col("u64")), "Sum", NotImplementedError, "UInt64"),
    (_bin("Add", _col("i8"), {"kind": "LiteralI64", "value": 1000}), "Sum", NotImplementedError, "does not fit I8"),
    (_bin("Add", {"kind": "LiteralI64", "value": 1}, {"kind": "LiteralF64", "value": 2.0}), "Sum",
     NotImplementedError, "literals only"),
    # an I8 result is not a summable dtype (as a Simple agg over I8 is not)
    (_bin("Mul", _col("i8"), {"kind": "LiteralI64", "value": 2}), "Sum", ValueError, "Sum / I8"),
    (_col("nope"), "Sum", KeyError, "column not found"),
])
def test_expression_agg_errors(expr, op, exc, msg):
    from polars_metal import _native
    with pytest.raises(exc, match=re.escape(msg)):
        _native.execute_groupby(_xplan(expr, op), 1, XCOLS)


def test_legacy_len_without_kind():
    # groupby/mod.rs:420-434: a missing "kind" with op Len reads as Length;
    # the plan parses and fails only at the device (none here) or runs.
    from polars_metal import _native
    plan = {"keys": [["k", "I64"]], "aggs": [{"op": "Len", "output_alias": "n"}]}
    try:
        _native.execute_groupby(plan, 1, COLS[:1])
    except Exception as e:  # no GPU in this container
        assert "plan parse error" not in str(e)


def test_predicate_rejected_by_execute_groupby():
    from polars_metal import _native
    plan = _gb_plan(predicate={"kind": "LiteralBool", "value": True})
    with pytest.raises(ValueError, match="use the fused entry"):
        _native.execute_groupby(plan, 1, COLS)


def test_predicate_kind_not_implemented():
    from polars_metal import _native
    plan = _gb_plan(predicate={"kind": "IsNull", "input": {}})
    with pytest.raises(NotImplementedError):
        _native.execute_filter_groupby(plan, 1, COLS)


def test_router_matches_reference_rules():
    from polars_metal import _native
    scan = {"kind": "Scan", "n_rows": 200_000, "columns": [["k", "I64"], ["v", "F64"]]}
    out = _native.compute_lifting_plan({"kind": "GroupBy", "keys": [["k", "I64"]], "input": scan})
    assert out == {"Scan#0": "cpu_leave", "GroupBy#1": "gpu_lift"}
    small = dict(scan, n_rows=100_000)
    out = _native.compute_lifting_plan({"kind": "GroupBy", "keys": [["k", "I64"]], "input": small})
    assert out["GroupBy#1"] == "cpu_leave"  # GPU iff n_rows > 100_000 (cost.rs:46-52)
    wide = {"kind": "GroupBy", "keys": [["a", "I64"], ["b", "I64"]], "input": scan}
    assert _native.compute_lifting_plan(wide)["GroupBy#1"].startswith("fallback:composite key total 130 bits")
    nested = {"kind": "GroupBy", "keys": [["k", "I64"]],
              "input": {"kind": "Filter", "predicate": {}, "input": {"kind": "Project", "columns": ["k"], "input": scan}}}
    out = _native.compute_lifting_plan(nested)
    assert list(out) == ["Scan#0", "Project#1", "Filter#2", "GroupBy#3"]
    assert out["Project#1"] == "cpu_leave" and out["Filter#2"] == "gpu_lift"
    odd = _native.compute_lifting_plan({"kind": "Join", "input": scan})
    assert odd["Join#1"] == "fallback:unsupported IR node: Join"
    with pytest.raises(KeyError):
        _native.compute_lifting_plan({"n_rows": 3})


def test_json_reader_rejects_garbage():
    lib = _lib.load()
    out = ctypes.c_void_p()
    assert lib.pmh_compute_lifting_plan(b"{\"kind\": ", ctypes.byref(out)) == _lib.PMH_E_VALUE
    assert b"malformed JSON" in lib.pmh_last_error()


def test_no_gpu_means_compute_error_not_fallback():
    # Without a device the backend must fail loudly (no CPU fallback).
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from polars_metal import _native
    with pytest.raises(Exception) as ei:
        _native.execute_groupby(_gb_plan(), 1, COLS)
    assert "no HIP device" in str(ei.value)


def test_row_limit_is_2_31_before_any_device_work():
    # The reference takes up to u32::MAX rows (filter.rs:73-77); this backend
    # narrows that to 2^31 - 1 and refuses more with ValueError on the host,
    # before the device is touched (so the same error here without a GPU).
    from polars_metal import _native
    for n in (1 << 31, (1 << 32) - 1):
        with pytest.raises(ValueError, match=r"below 2\^31"):
            _native.execute_groupby(_gb_plan(), n, COLS)
        with pytest.raises(ValueError, match=r"below 2\^31"):
            _native.execute_filter_groupby(_gb_plan(), n, COLS)


def _fake_device_cols(rows, short=None):
    # device pointers that are never dereferenced: the batch checks every
    # query's buffers on the host first
    cols = []
    for name, tag, w in (("key", "I64", 8), ("v", "F64", 8), ("w", "I32", 4)):
        n = rows * w - (8 if name == short else 0)
        cols.append((name, tag, 0x10000, n, 0, 0))
    return cols


def test_batch_checks_every_query_before_any_device_work():
    from polars_metal import batch
    plan = {"keys": [["key", "I64"]], "aggs": [{"kind": "Simple", "input_col": "v", "op": "Sum", "output_alias": "s"}]}
    pred_w = {"kind": "Compare", "op": "Gt", "dtype": "I64", "lhs": {"kind": "Column", "name": "w", "dtype": "I64"},
              "rhs": {"kind": "LiteralI64", "value": 0}}
    good = (1000, _fake_device_cols(1000), None)
    # a later query whose value buffer is shorter than its row count
    with pytest.raises(ValueError, match="value buffer too short for v"):
        batch.execute_batch(plan, [good, (1000, _fake_device_cols(1000, short="v"), None)], options={"tile": 4096})
    # a column read only by a query's predicate is checked too
    with pytest.raises(ValueError, match="value buffer too short for w"):
        batch.execute_batch(plan, [good, (1000, _fake_device_cols(1000, short="w"), pred_w)], options={"tile": 4096})
    # 2^31 rows in any query
    with pytest.raises(ValueError, match=r"2\^31"):
        batch.execute_batch(plan, [good, ((1 << 31), _fake_device_cols(1000), None)], options={"tile": 4096})
    # a query without a column the template names
    with pytest.raises(KeyError, match="has no column v"):
        batch.execute_batch(plan, [good, (10, [c for c in _fake_device_cols(10) if c[0] != "v"], None)],
                            options={"tile": 4096})
