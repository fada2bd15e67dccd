"""Writes tests/golden/cases.json: small group-by / filter cases whose
expected outputs are computed here in pure Python (dicts, math.fsum, Python
integers), independently of the C oracle and of the HIP path.

Each case is named for what it pins:
  arith_mod7        keys i % 7, v = i: arithmetic-series sums, counts, min,
                    max, mean (closed form; reference shape
                    tests/python_integration/test_groupby.py:28-39)
  null_keys         a null key is its own group (test_groupby.py:114-125)
  null_values       nulls skipped; all-null group: sum 0 (valid), count 0,
                    min/max/mean null (groupby.rs:1268-1280)
  totalord_filter   v > 0.0 with -0.0, NaN, +-inf and a null predicate row
                    (docs/open-questions.md:55-66; test_filter_edges.py:46-92)
  kleene_filter     (a > 0) | (b > 0) and (a > 0) & (b > 0) with nulls
  composite_keys    (a i32, b i64) composite keys, i64 wrapping sum
  cancel_sums       f64 sums that cancel exactly (1e16 + 1 - 1e16) and sums
                    of 0.1, checked against math.fsum
  i32_sum_wrap      i32 sum wraps at 2^32 (groupby.rs i32 dispatch)
  all_unique        every row its own group (test_groupby.py:152)
  all_same          one group (test_groupby.py:166)
  bool_keys         Bool key column
  empty             zero rows
  all_filtered      predicate removes every row
  nan_minmax_f64    NaN poisons f64 min and max; a null NaN does not
                    (reference KATs crates/polars-metal-kernels/tests/
                    test_groupby_aggregate.rs:469-497: [1, NaN, 3] -> min
                    NaN, [10, NaN, 5] -> max NaN)
  signed_zero_f64   groups holding -0.0 and +0.0 in either order: min is
                    -0.0 and max is +0.0 whatever the row order
  f32_aggs          F32 sum / mean are the exact sum rounded to f64, then
                    to f32 (so [1, 2^-24, 2^-24] sums to 1 + 2^-23, where
                    an f32 running sum gives 1); min / max with NaN, +-0.0,
                    +-inf
  f32_nan_keys_f64  F32 values under a predicate on an F64 column
  expr_*            Expression aggs (AggExpr trees, crates/polars-metal-core/
                    src/plan/mod.rs:133-142): each Binary node of a case
                    carries its compute dtype ("ct"), assigned by hand from
                    Polars' arithmetic rules (the reference evaluates the
                    tree through Polars, python/polars_metal/_udf.py:
                    1164-1209), and the per-row values are computed here
                    with Python integers wrapped to that width, Python
                    floats, and f32 rounding through struct after every
                    operation; "ct" is stripped from the written plan.
                    expr_q1_disc and expr_ab_sum are the reference's
                    tests/python_integration/test_walker_expression_
                    unfolding.py:17-56 shapes, sum(a * (1 - b)) and
                    sum(a * b), with nulls added.

Float min/max rule: NaN anywhere in the group (among non-null rows) gives
the canonical quiet NaN; otherwise the extreme value, with -0.0 ordered
below +0.0 (the reference's Rayon fold keeps whichever zero it met first,
crates/polars-metal-kernels/src/groupby.rs:1471-1520, so no row order is
fixed there; this rule is order independent).

Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import math
import struct
from pathlib import Path

OUT = Path(__file__).resolve().parent / "cases.json"


def fenc(x):
    if x is None:
        return None
    if isinstance(x, float):
        if math.isnan(x):
            return "nan"
        return x.hex()
    return x


def col(tag, values):
    return {"tag": tag, "values": list(values)}


def f32r(x):
    """Round a Python float to the nearest f32 (ties to even), as a float."""
    if math.isnan(x) or math.isinf(x):
        return x
    return struct.unpack("<f", struct.pack("<f", x))[0]


def xsum(vals):
    """Correctly rounded sum; IEEE special values as a running sum gives
    them (NaN, or +inf and -inf together -> NaN; one infinity -> it)."""
    if any(math.isnan(x) for x in vals) or (inf_in(vals, 1) and inf_in(vals, -1)):
        return float("nan")
    if inf_in(vals, 1) or inf_in(vals, -1):
        return float("inf") if inf_in(vals, 1) else float("-inf")
    return math.fsum(vals) + 0.0


def inf_in(vals, sign):
    return any(math.isinf(x) and math.copysign(1.0, x) == sign for x in vals)


def fmin(vals):
    """Float min: NaN poisons; -0.0 below +0.0."""
    if any(math.isnan(x) for x in vals):
        return float("nan")
    return min(vals, key=lambda x: (x, math.copysign(1.0, x)))


def fmax(vals):
    if any(math.isnan(x) for x in vals):
        return float("nan")
    return max(vals, key=lambda x: (x, math.copysign(1.0, x)))


def tot_cmp(a, b):
    an, bn = math.isnan(a), math.isnan(b)
    if an and bn:
        return 0
    if an:
        return 1
    if bn:
        return -1
    return (a > b) - (a < b)


def eval_pred(p, cols, r):
    """(value, valid) for row r."""
    k = p["kind"]
    if k == "Column":
        v = cols[p["name"]]["values"][r]
        return v, v is not None
    if k.startswith("Literal"):
        return p["value"], True
    if k == "Compare":
        a, av = eval_pred(p["lhs"], cols, r)
        b, bv = eval_pred(p["rhs"], cols, r)
        if not (av and bv):
            return False, False
        if p.get("dtype") == "F64":
            c = tot_cmp(float(a), float(b))
        else:
            c = (a > b) - (a < b)
        return {"Eq": c == 0, "Ne": c != 0, "Lt": c < 0, "Le": c <= 0, "Gt": c > 0, "Ge": c >= 0}[p["op"]], True
    if k in ("And", "Or"):
        a, av = eval_pred(p["lhs"], cols, r)
        b, bv = eval_pred(p["rhs"], cols, r)
        dom = k == "Or"
        if (av and a == dom) or (bv and b == dom):
            return dom, True
        if av and bv:
            return (not dom), True
        return False, False
    raise ValueError(k)


def wrap(x, bits):
    x &= (1 << bits) - 1
    return x - (1 << bits) if x >= 1 << (bits - 1) else x


INT_BITS = {"I8": 8, "I16": 16, "I32": 32, "I64": 64, "U8": 8, "U16": 16, "U32": 32, "U64": 64}


def fdiv(a, b):
    """IEEE division (Python raises on a zero divisor)."""
    if b == 0.0:
        if a == 0.0 or math.isnan(a):
            return float("nan")
        return math.copysign(float("inf"), a) * math.copysign(1.0, b)
    return a / b


def eval_expr_row(e, cols, r):
    """Value of an annotated AggExpr tree at row r, or None when a column it
    reads is null there."""
    k = e["kind"]
    if k == "Column":
        return cols[e["name"]]["values"][r]
    if k.startswith("Literal"):
        return e["value"]
    a, b = eval_expr_row(e["lhs"], cols, r), eval_expr_row(e["rhs"], cols, r)
    if a is None or b is None:
        # null propagates; both sides are still evaluated (a column read
        # anywhere in the tree nulls the row)
        return None
    ct, op = e["ct"], e["op"]
    if ct in INT_BITS:
        a, b = int(a), int(b)
        x = {"Add": a + b, "Sub": a - b, "Mul": a * b}[op]
        bits = INT_BITS[ct]
        return wrap(x, bits) if ct.startswith("I") else x & ((1 << bits) - 1)
    a, b = float(a), float(b)
    if ct == "F32":
        a, b = f32r(a), f32r(b)
    x = {"Add": lambda: a + b, "Sub": lambda: a - b, "Mul": lambda: a * b, "Div": lambda: fdiv(a, b)}[op]()
    return f32r(x) if ct == "F32" else x


def strip_ct(e):
    e = {k: v for k, v in e.items() if k != "ct"}
    for side in ("lhs", "rhs"):
        if side in e:
            e[side] = strip_ct(e[side])
    return e


def expected(plan, cols, n):
    keys = [k[0] for k in plan["keys"]]
    groups = {}
    order = []
    for r in range(n):
        if plan.get("predicate"):
            v, ok = eval_pred(plan["predicate"], cols, r)
            if not (ok and v):
                continue
        kt = tuple(cols[k]["values"][r] for k in keys)
        ktn = tuple((x is None, 0 if x is None else (("f", float(x).hex()) if isinstance(x, float) else x)) for x in kt)
        if ktn not in groups:
            groups[ktn] = {"key": kt, "rows": []}
            order.append(ktn)
        groups[ktn]["rows"].append(r)
    out = {k: [] for k in keys}
    for a in plan["aggs"]:
        out[a["output_alias"]] = []
    for ktn in order:
        g = groups[ktn]
        for k, x in zip(keys, g["key"]):
            out[k].append(x)
        for a in plan["aggs"]:
            rows = g["rows"]
            if a["op"] == "Len":
                out[a["output_alias"]].append(len(rows))
                continue
            if a.get("kind") == "Expression":
                ev = [eval_expr_row(a["expr"], cols, r) for r in rows]
                vals = [x for x in ev if x is not None]
                tag = a["tag"]
            else:
                c = cols[a["input_col"]]
                vals = [c["values"][r] for r in rows if c["values"][r] is not None]
                tag = c["tag"]
            op = a["op"]
            if op == "Count":
                out[a["output_alias"]].append(len(vals))
            elif op == "Sum":
                if tag == "F64":
                    out[a["output_alias"]].append(xsum(vals))
                elif tag == "F32":
                    out[a["output_alias"]].append(f32r(xsum(vals)))
                else:
                    out[a["output_alias"]].append(wrap(sum(vals), 64 if tag == "I64" else 32))
            elif op == "Mean":
                if not vals:
                    out[a["output_alias"]].append(None)
                elif tag == "F64":
                    out[a["output_alias"]].append(xsum(vals) / len(vals))
                elif tag == "F32":
                    out[a["output_alias"]].append(f32r(xsum(vals) / len(vals)))
                else:
                    out[a["output_alias"]].append(float(wrap(sum(vals), 64)) / len(vals))
            elif op in ("Min", "Max"):
                if not vals:
                    out[a["output_alias"]].append(None)
                elif tag in ("F64", "F32"):
                    out[a["output_alias"]].append(fmin(vals) if op == "Min" else fmax(vals))
                else:
                    out[a["output_alias"]].append(min(vals) if op == "Min" else max(vals))
    return {k: [fenc(x) for x in v] for k, v in out.items()}, len(order)


def agg(col_, op, alias=None):
    return {"kind": "Simple", "input_col": col_, "op": op, "output_alias": alias or f"{col_}_{op.lower()}"}


FIVE = [agg("v", "Sum"), agg("v", "Count"), agg("w", "Min"), agg("w", "Max"), agg("v", "Mean"),
        {"kind": "Length", "op": "Len", "output_alias": "len"}]


def gt0(name, dtype="F64"):
    lit = {"kind": "LiteralF64", "value": 0.0} if dtype == "F64" else {"kind": "LiteralI64", "value": 0}
    return {"kind": "Compare", "op": "Gt", "dtype": dtype, "lhs": {"kind": "Column", "name": name, "dtype": dtype},
            "rhs": lit}


def cases():
    cs = []
    n = 70
    cs.append(("arith_mod7", {"keys": [["key", "I64"]], "aggs": FIVE},
               {"key": col("I64", [i % 7 for i in range(n)]), "v": col("F64", [float(i) for i in range(n)]),
                "w": col("I32", list(range(n)))}, n))
    keys = [1, None, 2, None, 1, 3, None, 2]
    cs.append(("null_keys", {"keys": [["key", "I64"]], "aggs": FIVE},
               {"key": col("I64", keys), "v": col("F64", [0.5, 1.5, 2.5, 3.5, 4.5, 5.5, 6.5, 7.5]),
                "w": col("I32", [10, 20, 30, 40, 50, 60, 70, 80])}, len(keys)))
    cs.append(("null_values", {"keys": [["key", "I64"]], "aggs": FIVE},
               {"key": col("I64", [1, 1, 2, 2, 3, 3]), "v": col("F64", [None, None, 1.25, None, -2.0, 4.0]),
                "w": col("I32", [None, None, 7, None, -3, 9])}, 6))
    vals = [-0.0, 0.0, float("nan"), 1.0, -1.0, float("inf"), None, float("-inf"), 2.5, float("nan")]
    cs.append(("totalord_filter", {"keys": [["key", "I64"]], "aggs": [agg("v", "Count"), agg("w", "Sum")],
                                   "predicate": gt0("v")},
               {"key": col("I64", [0, 1, 2, 3, 0, 1, 2, 3, 0, 1]), "v": col("F64", vals),
                "w": col("I32", list(range(10)))}, 10))
    a = [1, -1, None, 1, None, -1, 1, None, -1]
    b = [1, 1, 1, -1, -1, -1, None, None, None]
    for kind in ("And", "Or"):
        pred = {"kind": kind, "lhs": gt0("a", "I64"), "rhs": gt0("b", "I64")}
        cs.append((f"kleene_{kind.lower()}", {"keys": [["k", "I64"]], "aggs": [agg("a", "Count"),
                                                                                {"kind": "Length", "op": "Len", "output_alias": "len"}],
                                              "predicate": pred},
                   {"k": col("I64", list(range(9))), "a": col("I64", a), "b": col("I64", b)}, 9))
    ca = [1, 2, 1, 2, 1, None, 1, 2, (1 << 20) - 1, 0]
    cb = [5, 5, 5, 6, 1 << 40, 7, 5, 6, (1 << 40) - 1, -9]
    cv = [1 << 62, 1 << 62, 3, -4, 5, 6, None, 8, -(1 << 63), 1]
    cs.append(("composite_keys", {"keys": [["a", "I32"], ["b", "I64"]],
                                  "aggs": [agg("v", "Sum"), agg("v", "Count"), agg("v", "Min"), agg("v", "Max"),
                                           agg("v", "Mean")]},
               {"a": col("I32", ca), "b": col("I64", cb), "v": col("I64", cv)}, 10))
    cv2 = [1e16, 1.0, -1e16, 0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 1e-300, -1e-300, 2.0 ** -1074, 3.0]
    ck = [0, 0, 0] + [1] * 10 + [2, 2, 2, 3]
    cs.append(("cancel_sums", {"keys": [["key", "I64"]], "aggs": [agg("v", "Sum"), agg("v", "Mean")]},
               {"key": col("I64", ck), "v": col("F64", cv2)}, len(ck)))
    big = (1 << 31) - 5
    cs.append(("i32_sum_wrap", {"keys": [["key", "I32"]], "aggs": [agg("v", "Sum"), agg("v", "Mean"), agg("v", "Min"),
                                                                    agg("v", "Max")]},
               {"key": col("I32", [7, 7, 7, -7]), "v": col("I32", [big, big, 100, -(1 << 31)])}, 4))
    cs.append(("all_unique", {"keys": [["key", "I64"]], "aggs": FIVE},
               {"key": col("I64", [10 * i - 300 for i in range(64)]), "v": col("F64", [i * 0.25 for i in range(64)]),
                "w": col("I32", [64 - i for i in range(64)])}, 64))
    cs.append(("all_same", {"keys": [["key", "I64"]], "aggs": FIVE},
               {"key": col("I64", [-1] * 50), "v": col("F64", [(-1) ** i * i / 3.0 for i in range(50)]),
                "w": col("I32", [i * 37 % 101 for i in range(50)])}, 50))
    cs.append(("bool_keys", {"keys": [["flag", "Bool"], ["key", "I32"]], "aggs": [agg("v", "Sum"), agg("v", "Count")]},
               {"flag": col("Bool", [True, False, None, True, False, True]), "key": col("I32", [1, 1, 1, 2, 1, 1]),
                "v": col("F64", [1.0, 2.0, 3.0, 4.0, 5.0, 6.0])}, 6))
    cs.append(("empty", {"keys": [["key", "I64"]], "aggs": FIVE},
               {"key": col("I64", []), "v": col("F64", []), "w": col("I32", [])}, 0))
    cs.append(("all_filtered", {"keys": [["key", "I64"]], "aggs": FIVE, "predicate": gt0("v")},
               {"key": col("I64", [1, 2, 3]), "v": col("F64", [-1.0, -2.0, None]), "w": col("I32", [1, 2, 3])}, 3))
    nan = float("nan")
    inf = float("inf")
    # groups 0 and 1 are the reference's KATs; 2: NaN only under a null;
    # 3: all null; 4: NaN alone; 5: no NaN
    nk = [0, 0, 0, 1, 1, 1, 2, 2, 3, 4, 5, 5, 5]
    nv = [1.0, nan, 3.0, 10.0, nan, 5.0, None, 7.0, None, nan, -inf, 2.0, inf]
    cs.append(("nan_minmax_f64", {"keys": [["key", "I64"]],
                                  "aggs": [agg("v", "Min"), agg("v", "Max"), agg("v", "Count"), agg("v", "Sum")]},
               {"key": col("I64", nk), "v": col("F64", nv)}, len(nk)))
    # nulls hide a NaN in group 2: a null NaN row must not poison
    cs[-1][2]["v"]["values"][6] = None
    zk = [0, 0, 1, 1, 2, 2, 2, 3, 3, 4, 5, 5, 6, 6]
    zv = [0.0, -0.0, -0.0, 0.0, 0.0, 0.0, 1.0, -0.0, -1.0, -0.0, 0.0, None, -0.0, -0.0]
    cs.append(("signed_zero_f64", {"keys": [["key", "I64"]],
                                   "aggs": [agg("v", "Min"), agg("v", "Max"), agg("v", "Sum"), agg("v", "Mean")]},
               {"key": col("I64", zk), "v": col("F64", zv)}, len(zk)))
    t = 2.0 ** -24
    fk = [0, 0, 0, 1, 1, 1, 2, 2, 3, 3, 3, 4, 4, 5, 5, 6, 6, 6, 7]
    fv = [1.0, t, t, 1e8, 1.0, -1e8, nan, 2.5, 0.0, -0.0, 0.0, -inf, 3.0, None, None, 0.1, 0.2, 0.3, 16777217.0]
    fv = [None if x is None else f32r(x) for x in fv]
    cs.append(("f32_aggs", {"keys": [["key", "I32"]],
                            "aggs": [agg("v", "Sum"), agg("v", "Mean"), agg("v", "Min"), agg("v", "Max"),
                                     agg("v", "Count"), {"kind": "Length", "op": "Len", "output_alias": "len"}]},
               {"key": col("I32", fk), "v": col("F32", fv)}, len(fk)))
    gv = [f32r(x) for x in [0.5, -1.5, nan, 4.0, -0.0, 0.0, 1e-3, 7.0]]
    gp = [1.0, 2.0, -1.0, nan, 0.0, 3.0, None, 5.0]
    cs.append(("f32_under_f64_pred", {"keys": [["key", "I64"]], "predicate": gt0("p"),
                                      "aggs": [agg("v", "Sum"), agg("v", "Min"), agg("v", "Max"), agg("v", "Mean")]},
               {"key": col("I64", [1, 1, 1, 2, 2, 2, 1, 2]), "v": col("F32", gv), "p": col("F64", gp)}, 8))
    cs.extend(expression_cases())
    return cs


def col_(name):
    return {"kind": "Column", "name": name}


def lit(v):
    return {"kind": "LiteralF64" if isinstance(v, float) else "LiteralI64", "value": v}


def binop(op, lhs, rhs, ct):
    return {"kind": "Binary", "op": op, "lhs": lhs, "rhs": rhs, "ct": ct}


def xagg(expr, op, alias, tag):
    """An Expression agg; `tag` (the tree's result dtype) is for this
    generator only and is stripped from the written plan."""
    return {"kind": "Expression", "op": op, "output_alias": alias, "expr": expr, "tag": tag}


def expression_cases():
    cs = []
    # sum(a * (1 - b)) and friends on the reference's Q1 shape: F64 columns,
    # an integer literal 1 taking Float64 beside b
    k = [0, 0, 1, 1] * 6 + [2, 2]
    a = [10.0, 20.0, 30.0, 40.0] * 6 + [None, 5.5]
    b = [0.05, 0.1, 0.15, 0.2] * 6 + [0.5, None]
    disc = binop("Mul", col_("a"), binop("Sub", lit(1), col_("b"), "F64"), "F64")
    cs.append(("expr_q1_disc", {"keys": [["k", "I64"]],
                                "aggs": [xagg(disc, "Sum", "disc", "F64"), xagg(disc, "Mean", "disc_mean", "F64"),
                                         xagg(disc, "Count", "disc_n", "F64"), xagg(disc, "Min", "disc_min", "F64"),
                                         xagg(disc, "Max", "disc_max", "F64"), agg("a", "Sum")]},
               {"k": col("I64", k), "a": col("F64", a), "b": col("F64", b)}, len(k)))
    k = [0, 0, 1, 1, 2, 2] * 5
    a = [1.0, 2.0, 3.0, 4.0, 5.0, 6.0] * 5
    b = [0.1, 0.2, 0.3, 0.4, 0.5, 0.6] * 5
    cs.append(("expr_ab_sum", {"keys": [["k", "I64"]],
                               "aggs": [xagg(binop("Mul", col_("a"), col_("b"), "F64"), "Sum", "sum_ab", "F64")]},
               {"k": col("I64", k), "a": col("F64", a), "b": col("F64", b)}, len(k)))
    # i32 * i32 wraps in 32 bits, then the i32 sum wraps too; + literal 7
    x = [70000, -70000, 46341, 1 << 30, None, 3, -(1 << 31), 12345]
    y = [70000, 70000, 46341, 4, 9, None, -1, -3]
    prod = binop("Add", binop("Mul", col_("x"), col_("y"), "I32"), lit(7), "I32")
    cs.append(("expr_i32_wrap", {"keys": [["k", "I32"]],
                                 "aggs": [xagg(prod, "Sum", "s", "I32"), xagg(prod, "Min", "mn", "I32"),
                                          xagg(prod, "Max", "mx", "I32"), xagg(prod, "Mean", "m", "I32"),
                                          xagg(prod, "Count", "c", "I32")]},
               {"k": col("I32", [1, 1, 2, 2, 1, 2, 3, 3]), "x": col("I32", x), "y": col("I32", y)}, 8))
    # true division: two integers give Float64 (x / 0 -> +-inf, 0 / 0 -> NaN)
    p = [7, -7, 0, 1 << 62, 9, None, 5, -12]
    q = [2, 0, 0, 3, -4, 1, 0, 5]
    cs.append(("expr_int_div", {"keys": [["k", "I64"]],
                                "aggs": [xagg(binop("Div", col_("p"), col_("q"), "F64"), "Sum", "s", "F64"),
                                         xagg(binop("Div", col_("p"), col_("q"), "F64"), "Max", "mx", "F64"),
                                         xagg(binop("Div", col_("w"), lit(3), "F64"), "Sum", "w3", "F64"),
                                         xagg(binop("Div", col_("w"), lit(3), "F64"), "Min", "w3min", "F64")]},
               {"k": col("I64", [0, 0, 1, 1, 2, 2, 3, 3]), "p": col("I64", p), "q": col("I64", q),
                "w": col("I32", [1, 2, 3, 4, -5, 6, 7, None])}, 8))
    # Float32: a float literal beside F32 is F32; I16 with F32 gives F32;
    # every operation rounds to f32
    f = [f32r(v) for v in [0.1, 1.0, 3.3, -2.5, 1e7, 0.7, 16777216.0, 2.0]]
    h = [3, -4, 100, 32767, -32768, 0, 1, None]
    fexpr = binop("Add", binop("Mul", col_("f"), lit(0.1), "F32"), col_("h"), "F32")
    cs.append(("expr_f32", {"keys": [["k", "I64"]],
                            "aggs": [xagg(fexpr, "Sum", "s", "F32"), xagg(fexpr, "Min", "mn", "F32"),
                                     xagg(fexpr, "Max", "mx", "F32"), xagg(fexpr, "Mean", "m", "F32"),
                                     xagg(binop("Div", col_("f"), lit(3), "F32"), "Sum", "f3", "F32")]},
               {"k": col("I64", [0, 0, 0, 1, 1, 1, 2, 2]), "f": col("F32", f), "h": col("I16", h)}, 8))
    # promotions: I32 + U32 -> I64; (U8 - 1) wraps in U8, then with I32 -> I32;
    # I32 * F32 -> F64; I8 * U16 -> I32
    u = [0, 4294967295, 7, 100, 2, 3]
    s32 = [-1, -1, 5, -200, None, 8]
    u8 = [0, 1, 255, 9, 3, None]
    f2 = [f32r(v) for v in [0.3, -1.5, 2.25, 1e-3, 8.0, 3.0]]
    i8 = [-128, 127, -1, 5, 0, 9]
    u16 = [65535, 65535, 2, 3, 1, 0]
    cs.append(("expr_promote", {"keys": [["k", "I64"]],
                                "aggs": [xagg(binop("Add", col_("s32"), col_("u"), "I64"), "Sum", "su", "I64"),
                                         xagg(binop("Add", binop("Sub", col_("u8"), lit(1), "U8"), col_("s32"), "I32"),
                                              "Sum", "u8s", "I32"),
                                         xagg(binop("Mul", col_("s32"), col_("f2"), "F64"), "Sum", "sf", "F64"),
                                         xagg(binop("Mul", col_("i8"), col_("u16"), "I32"), "Max", "iu", "I32")]},
               {"k": col("I64", [0, 1, 0, 1, 0, 1]), "u": col("U32", u), "s32": col("I32", s32),
                "u8": col("U8", u8), "f2": col("F32", f2), "i8": col("I8", i8), "u16": col("U16", u16)}, 6))
    # four binary levels (the reference's depth cap), nulls of w propagate
    deep = binop("Div", binop("Sub", binop("Mul", binop("Add", col_("v"), lit(1.0), "F64"), lit(2), "F64"),
                                       col_("w"), "F64"), lit(4), "F64")
    cs.append(("expr_depth4", {"keys": [["k", "I64"]], "predicate": gt0("v"),
                               "aggs": [xagg(deep, "Sum", "s", "F64"), xagg(deep, "Count", "c", "F64"),
                                        xagg(deep, "Mean", "m", "F64"), agg("v", "Count")]},
               {"k": col("I64", [1, 2, 1, 2, 1, 2, 1, 2]), "v": col("F64", [0.5, 1.5, -2.0, 3.25, 9.0, None, 0.125, 7.0]),
                "w": col("I32", [1, None, 3, 4, -5, 6, 7, 8])}, 8))
    return cs


def main():
    out = []
    for name, plan, cols, n in cases():
        exp, ng = expected(plan, cols, n)
        enc_cols = {k: {"tag": c["tag"], "values": [fenc(v) for v in c["values"]]} for k, c in cols.items()}
        expr_tags = {}
        aggs = []
        for a in plan["aggs"]:
            if a.get("kind") == "Expression":
                expr_tags[a["output_alias"]] = a["tag"]
                a = {k: v for k, v in a.items() if k != "tag"}
                a["expr"] = strip_ct(a["expr"])
            aggs.append(a)
        plan = dict(plan, aggs=aggs)
        case = {"name": name, "plan": plan, "n_rows": n, "columns": enc_cols, "expected": exp, "n_groups": ng}
        if expr_tags:
            case["expr_tags"] = expr_tags
        out.append(case)
    OUT.write_text(json.dumps(out, indent=1))
    print(f"wrote {len(out)} cases to {OUT}")


if __name__ == "__main__":
    main()
