"""Python side of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  It wraps oracle/pmh_oracle.c (built by
oracle/Makefile into oracle/_build/libpmh_oracle.so) and assembles group-by
outputs with the reference's dtype rules:

  * sum: I64 -> I64 (wrapping), I32 -> I32 (wrapping), F64 -> F64, F32 -> F32;
    all-null groups sum to 0 and are valid
    (crates/polars-metal-kernels/src/groupby.rs:2199-2215)
  * mean: F64 for integer and F64 inputs, F32 for F32 inputs; null when the
    group has no non-null value (groupby.rs:1641-1670; Polars' Float64 mean of
    integers, tests/python_integration/test_groupby.py:188-193)
  * count / len: U32 (crates/polars-metal-core/src/udf/groupby/mod.rs:695-706)
  * min / max: input dtype, null for all-null groups, NaN poisons floats
    (groupby.rs:1467-1469)
  * keys: the value at the group's first row (groupby.rs:1833-1985)
  * Utf8 keys: grouped on dictionary codes in first-seen row order, null
    rows code 0 and null, "" a value of its own (build_dict_nullable,
    crates/polars-metal-buffer/src/dict.rs:40-66; ``utf8_dict``); the key
    column comes back as the group's first string, and on the wire in the
    packed form u32 n, i32 offsets[n + 1], bytes, a null row empty
    (crates/polars-metal-core/src/udf/groupby/mod.rs:646-670)

Frames are dicts ``name -> (tag, values, valid)``; ``values`` is a numpy
array (bool array for Bool, an object array of str for Utf8), ``valid`` a
bool array or None (all valid).
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "libpmh_oracle.so"

TAGS = {"I64": 0, "F64": 1, "Bool": 2, "I32": 3, "F32": 4, "I8": 5, "I16": 6, "U8": 7,
        "U16": 8, "U32": 9, "U64": 10}
NP_DTYPE = {"I64": np.int64, "F64": np.float64, "I32": np.int32, "F32": np.float32, "I8": np.int8,
            "I16": np.int16, "U8": np.uint8, "U16": np.uint16, "U32": np.uint32, "U64": np.uint64,
            "Bool": np.bool_}
CMP = {"Eq": 0, "Ne": 1, "Lt": 2, "Le": 3, "Gt": 4, "Ge": 5}
N_COLUMN, N_LIT_I64, N_LIT_F64, N_LIT_BOOL, N_COMPARE, N_AND, N_OR = range(7)


class OrcCol(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_int32), ("pad", ctypes.c_int32), ("data", ctypes.c_void_p),
                ("valid", ctypes.c_void_p)]


class OrcNode(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("op", ctypes.c_int32), ("col", ctypes.c_int32),
                ("dtype", ctypes.c_int32), ("ival", ctypes.c_int64), ("fval", ctypes.c_double)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        L.orc_eval_predicate.argtypes = [ctypes.POINTER(OrcNode), ctypes.c_int, ctypes.POINTER(OrcCol),
                                         ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_compact.restype = ctypes.c_uint64
        L.orc_compact.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_void_p]
        L.orc_exact_sum.restype = ctypes.c_double
        L.orc_exact_sum.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
        L.orc_encode_keys.argtypes = [ctypes.c_int, ctypes.POINTER(OrcCol), ctypes.c_uint64, ctypes.c_void_p,
                                      ctypes.c_void_p]
        L.orc_groupby.restype = ctypes.c_void_p
        L.orc_groupby.argtypes = [ctypes.c_uint64, ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(OrcCol),
                                  ctypes.c_int, ctypes.POINTER(OrcCol), ctypes.c_int]
        L.orc_groupby_par.restype = ctypes.c_void_p
        L.orc_groupby_par.argtypes = [ctypes.c_uint64, ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(OrcCol),
                                      ctypes.c_int, ctypes.POINTER(OrcCol), ctypes.c_int, ctypes.c_int]
        L.orc_gb_n_groups.restype = ctypes.c_uint64
        L.orc_gb_n_groups.argtypes = [ctypes.c_void_p]
        L.orc_gb_lanes.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 4
        L.orc_gb_val.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 7
        L.orc_gb_free.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


# ---------------------------------------------------------------------------
# frame <-> buffers
# ---------------------------------------------------------------------------
def pack_bits(b: np.ndarray) -> np.ndarray:
    return np.packbits(np.asarray(b, dtype=bool), bitorder="little")


def unpack_bits(buf, n: int) -> np.ndarray:
    a = np.frombuffer(bytes(buf), dtype=np.uint8) if not isinstance(buf, np.ndarray) else buf
    return np.unpackbits(a, bitorder="little")[:n].astype(bool)


class _Keep:
    """Keeps numpy buffers alive while ctypes holds raw pointers."""

    def __init__(self):
        self.refs = []

    def col(self, tag: str, values: np.ndarray, valid) -> OrcCol:
        c = OrcCol()
        c.tag = TAGS[tag]
        if tag == "Bool":
            data = pack_bits(values)
        else:
            data = np.ascontiguousarray(values, dtype=NP_DTYPE[tag])
        self.refs.append(data)
        c.data = data.ctypes.data if data.size else None
        if valid is None:
            c.valid = None
        else:
            vb = pack_bits(valid)
            self.refs.append(vb)
            c.valid = vb.ctypes.data
        return c


def frame_rows(frame: dict) -> int:
    return len(next(iter(frame.values()))[1])


def utf8_dict(values, valid) -> tuple:
    """The reference's build_dict_nullable (dict.rs:40-66): (distinct strings
    in first-seen order, u32 codes, validity); a null row has code 0."""
    n = len(values)
    ok = np.ones(n, bool) if valid is None else np.asarray(valid, bool)
    dictionary: list = []
    seen: dict = {}
    codes = np.zeros(n, np.uint32)
    for i in range(n):
        if not ok[i]:
            continue
        s = values[i]
        c = seen.get(s)
        if c is None:
            c = seen[s] = len(dictionary)
            dictionary.append(s)
        codes[i] = c
    return dictionary, codes, ok


def utf8_packed(values, valid) -> bytes:
    """Strings -> the packed wire form (a null row has length 0)."""
    n = len(values)
    ok = np.ones(n, bool) if valid is None else np.asarray(valid, bool)
    blobs = [values[i].encode("utf-8") if ok[i] else b"" for i in range(n)]
    offs = np.zeros(n + 1, np.int32)
    np.cumsum([len(b) for b in blobs], out=offs[1:])
    return np.uint32(n).tobytes() + offs.tobytes() + b"".join(blobs)


def utf8_unpacked(data: bytes) -> np.ndarray:
    n = int(np.frombuffer(data, np.uint32, 1)[0])
    offs = np.frombuffer(data, np.int32, n + 1, 4)
    blob = data[4 + 4 * (n + 1):]
    return np.array([blob[offs[i]:offs[i + 1]].decode("utf-8") for i in range(n)], object)


def to_wire(frame: dict, names=None) -> list:
    """Frame -> the reference's ``(name, tag, data, valid)`` tuples."""
    out = []
    for name in (names or frame.keys()):
        tag, values, valid = frame[name]
        if tag == "Utf8":
            n = len(values)
            out.append((name, tag, utf8_packed(values, valid),
                        pack_bits(np.ones(n, bool) if valid is None else valid).tobytes()))
            continue
        if tag == "Bool":
            data = pack_bits(values).tobytes()
        else:
            data = np.ascontiguousarray(values, dtype=NP_DTYPE[tag]).tobytes()
        n = len(values)
        vb = pack_bits(np.ones(n, bool) if valid is None else valid).tobytes()
        out.append((name, tag, data, vb))
    return out


# ---------------------------------------------------------------------------
# predicate
# ---------------------------------------------------------------------------
def _lower_pred(p: dict, frame: dict, names: list, nodes: list) -> None:
    k = p["kind"]
    if k == "Column":
        name = p["name"]
        if name not in names:
            names.append(name)
        tag = frame[name][0]
        dtype = "Bool" if tag == "Bool" else ("F64" if tag in ("F64", "F32") else "I64")
        nodes.append(OrcNode(N_COLUMN, 0, names.index(name), TAGS[dtype], 0, 0.0))
    elif k == "LiteralI64":
        nodes.append(OrcNode(N_LIT_I64, 0, 0, 0, int(p["value"]), float(p["value"])))
    elif k == "LiteralF64":
        nodes.append(OrcNode(N_LIT_F64, 0, 0, 1, 0, float(p["value"])))
    elif k == "LiteralBool":
        nodes.append(OrcNode(N_LIT_BOOL, 0, 0, 2, int(bool(p["value"])), 0.0))
    elif k == "Compare":
        _lower_pred(p["lhs"], frame, names, nodes)
        _lower_pred(p["rhs"], frame, names, nodes)

        def is_f(s):
            return s["kind"] == "LiteralF64" or (s["kind"] == "Column" and frame[s["name"]][0] in ("F64", "F32"))

        dt = p.get("dtype")
        f = (dt in ("F64", "F32")) if dt else (is_f(p["lhs"]) or is_f(p["rhs"]))
        nodes.append(OrcNode(N_COMPARE, CMP[p["op"]], 0, TAGS["F64" if f else "I64"], 0, 0.0))
    elif k in ("And", "Or"):
        _lower_pred(p["lhs"], frame, names, nodes)
        _lower_pred(p["rhs"], frame, names, nodes)
        nodes.append(OrcNode(N_AND if k == "And" else N_OR, 0, 0, 2, 0, 0.0))
    else:
        raise NotImplementedError(k)


def eval_predicate(pred: dict, frame: dict):
    """-> (data, valid) bool arrays, one entry per row."""
    n = frame_rows(frame)
    names: list = []
    nodes: list = []
    _lower_pred(pred, frame, names, nodes)
    keep = _Keep()
    cols = (OrcCol * max(1, len(names)))(*[keep.col(*frame[nm]) for nm in names])
    narr = (OrcNode * len(nodes))(*nodes)
    d = np.zeros(n, np.uint8)
    v = np.zeros(n, np.uint8)
    rc = lib().orc_eval_predicate(narr, len(nodes), cols, n, d.ctypes.data, v.ctypes.data)
    if rc != 0:
        raise ValueError("oracle: bad predicate program")
    return d.astype(bool), v.astype(bool)


def compact(keep_mask: np.ndarray, tag: str, values: np.ndarray, valid):
    """Stable compaction of one column -> (values, valid)."""
    n = len(values)
    km = np.ascontiguousarray(keep_mask, dtype=np.uint8)
    k = _Keep()
    c = k.col(tag, values, valid)
    if tag == "Bool":
        src = np.ascontiguousarray(values, dtype=np.uint8)
        esz = 1
    else:
        src = np.ascontiguousarray(values, dtype=NP_DTYPE[tag])
        esz = src.dtype.itemsize
    out = np.zeros(max(1, n), dtype=src.dtype)
    ov = np.zeros(max(1, n), np.uint8)
    m = lib().orc_compact(km.ctypes.data, n, esz, src.ctypes.data if n else None, c.valid, out.ctypes.data,
                          ov.ctypes.data)
    vals = out[:m]
    if tag == "Bool":
        vals = vals.astype(bool)
    return vals, ov[:m].astype(bool)


def exact_sum(x: np.ndarray) -> float:
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().orc_exact_sum(x.ctypes.data if x.size else None, x.size)


def encode_keys(frame: dict, keys: list):
    n = frame_rows(frame)
    k = _Keep()
    cols = (OrcCol * len(keys))(*[k.col(*frame[nm]) for nm in keys])
    lo = np.zeros(max(1, n), np.uint64)
    hi = np.zeros(max(1, n), np.uint64)
    bits = lib().orc_encode_keys(len(keys), cols, n, lo.ctypes.data, hi.ctypes.data)
    if bits == -2:
        raise ValueError("composite key too wide")
    return lo[:n], hi[:n], bits


# ---------------------------------------------------------------------------
# expression aggregates
# ---------------------------------------------------------------------------
# The reference evaluates an Expression agg's AggExpr tree
# (crates/polars-metal-core/src/plan/mod.rs:133-142) through Polars before
# the group-by call and folds the result as a Simple agg
# (python/polars_metal/_udf.py:1164-1209).  Restated here with Polars'
# arithmetic rules: Add/Sub/Mul in the numeric supertype (integers wrap,
# floats round once per operation), Div true division (Float64 for two
# integers), literals typed by the other operand, nulls propagated from every
# column read.  Shapes Polars types outside these rules raise
# NotImplementedError, as the backend does.
_INT_BITS = {"I8": 8, "I16": 16, "I32": 32, "I64": 64, "U8": 8, "U16": 16, "U32": 32, "U64": 64}
_FLOATS = ("F32", "F64")


def _supertype(a: str, b: str) -> str:
    if a == b:
        return a
    if a in _FLOATS and b in _FLOATS:
        return "F64"
    if a in _FLOATS or b in _FLOATS:
        f, i = (a, b) if a in _FLOATS else (b, a)
        return "F64" if f == "F64" or _INT_BITS[i] > 16 else "F32"
    sa, sb = a.startswith("I"), b.startswith("I")
    if sa == sb:
        return a if _INT_BITS[a] >= _INT_BITS[b] else b
    si, ui = (a, b) if sa else (b, a)
    if _INT_BITS[si] > _INT_BITS[ui]:
        return si
    if ui == "U64":
        raise NotImplementedError("signed integer with UInt64")
    return "I" + str(2 * _INT_BITS[ui])


def _lit_value(lit, tag: str):
    """An open literal ('int'|'float', value) as a numpy scalar of tag."""
    kind, v = lit
    if tag in _FLOATS:
        return NP_DTYPE[tag](float(v) if kind == "float" else int(v))
    if kind == "float":
        raise NotImplementedError("float literal as an integer")
    info = np.iinfo(NP_DTYPE[tag])
    if not info.min <= int(v) <= info.max:
        raise NotImplementedError(f"literal {v} does not fit {tag}")
    return NP_DTYPE[tag](int(v))


def _expr(e: dict, frame: dict, used: list):
    """-> ('col', tag, values) or ('lit', (kind, value))."""
    k = e["kind"]
    if k == "Column":
        tag, values, _ = frame[e["name"]]
        if tag not in _INT_BITS and tag not in _FLOATS:
            raise NotImplementedError(f"{tag} in arithmetic")
        if e["name"] not in used:
            used.append(e["name"])
        return ("col", tag, np.asarray(values, dtype=NP_DTYPE[tag]))
    if k == "LiteralI64":
        return ("lit", ("int", int(e["value"])))
    if k == "LiteralF64":
        return ("lit", ("float", float(e["value"])))
    if k != "Binary":
        raise ValueError(k)
    op = e["op"]
    lhs, rhs = _expr(e["lhs"], frame, used), _expr(e["rhs"], frame, used)
    if lhs[0] == "lit" and rhs[0] == "lit":
        (ka, a), (kb, b) = lhs[1], rhs[1]
        if ka == kb == "int" and op != "Div":
            r = {"Add": a + b, "Sub": a - b, "Mul": a * b}[op]
            r = (r + 2 ** 63) % 2 ** 64 - 2 ** 63
            return ("lit", ("int", r))
        a, b = float(a), float(b)
        with np.errstate(all="ignore"):
            r = {"Add": np.float64(a) + b, "Sub": np.float64(a) - b, "Mul": np.float64(a) * b,
                 "Div": np.float64(a) / b}[op]
        return ("lit", ("float", float(r)))

    def side(me, other):
        if me[0] == "col":
            return me[1]
        return other[1] if me[1][0] == "int" else ("F32" if other[1] == "F32" else "F64")

    tl, tr = side(lhs, rhs), side(rhs, lhs)
    if op == "Div" and tl in _INT_BITS and tr in _INT_BITS:
        ct = "F64"
    else:
        ct = _supertype(tl, tr)

    def operand(x, t):
        if x[0] == "lit":
            _lit_value(x[1], t)  # must hold its operand type
            return _lit_value(x[1], ct)
        return x[2].astype(NP_DTYPE[ct])

    a, b = operand(lhs, tl), operand(rhs, tr)
    with np.errstate(all="ignore"):
        if op == "Add":
            r = a + b
        elif op == "Sub":
            r = a - b
        elif op == "Mul":
            r = a * b
        else:
            r = a / b
    n = frame_rows(frame)
    return ("col", ct, np.broadcast_to(np.asarray(r, dtype=NP_DTYPE[ct]), (n,)).copy())


def eval_expr(expr: dict, frame: dict) -> tuple:
    """AggExpr dict -> (tag, values, valid) for every row of frame."""
    used: list = []
    r = _expr(expr, frame, used)
    if r[0] == "lit":
        raise NotImplementedError("a tree of literals only")
    tag, values = r[1], r[2]
    valid = np.ones(len(values), bool)
    for nm in used:
        v = frame[nm][2]
        if v is not None:
            valid &= np.asarray(v, bool)
    values = np.where(valid, values, np.zeros(1, values.dtype))
    return tag, values, valid


def materialize_expressions(plan: dict, frame: dict) -> tuple:
    """The reference's _materialize_expression_aggs (_udf.py:1173-1209): each
    Expression agg becomes a Simple agg over a column holding its values."""
    if not any(a.get("kind") == "Expression" for a in plan["aggs"]):
        return plan, frame
    import json
    frame = dict(frame)
    aggs = []
    names: dict = {}  # one column per distinct tree
    for a in plan["aggs"]:
        if a.get("kind") != "Expression" or a.get("op") == "Len":
            aggs.append(a)
            continue
        text = json.dumps(a["expr"], sort_keys=True)
        if text not in names:
            names[text] = f"__pm_expr_agg_{len(names)}__"
            frame[names[text]] = eval_expr(a["expr"], frame)
        aggs.append({"kind": "Simple", "input_col": names[text], "op": a["op"], "output_alias": a["output_alias"]})
    return dict(plan, aggs=aggs), frame


# ---------------------------------------------------------------------------
# group-by
# ---------------------------------------------------------------------------
def groupby(plan: dict, frame: dict, sum_mode: int = 0, threads: int | None = None) -> dict:
    """Run plan (keys, aggs, optional predicate) on frame.  Returns an ordered
    dict ``name -> (tag, values, valid)`` with groups in first-occurrence
    order, plus the keys '__n_groups__' and '__ops__' (output alias -> agg
    op, which the comparison helpers read).  ``threads`` runs the all-core
    form (orc_groupby_par): the same bits, groups in partition order."""
    plan, frame = materialize_expressions(plan, frame)
    n = frame_rows(frame)
    keep_mask = None
    if plan.get("predicate"):
        d, v = eval_predicate(plan["predicate"], frame)
        keep_mask = np.ascontiguousarray(d & v, dtype=np.uint8)
    keys = [k[0] for k in plan["keys"]]
    val_names: list = []
    for a in plan["aggs"]:
        if a.get("op") != "Len" and a.get("kind") != "Length" and a["input_col"] not in val_names:
            val_names.append(a["input_col"])
    k = _Keep()

    def key_col(nm):
        tag, values, valid = frame[nm]
        if tag == "Utf8":
            _, codes, ok = utf8_dict(values, valid)
            return k.col("U32", codes, ok)
        return k.col(tag, values, valid)

    kcols = (OrcCol * max(1, len(keys)))(*[key_col(nm) for nm in keys])
    vcols = (OrcCol * max(1, len(val_names)))(*[k.col(*frame[nm]) for nm in val_names])
    L = lib()
    km = keep_mask.ctypes.data if keep_mask is not None else None
    if threads is None:
        g = L.orc_groupby(n, km, len(keys), kcols, len(val_names), vcols, sum_mode)
    else:
        g = L.orc_groupby_par(n, km, len(keys), kcols, len(val_names), vcols, sum_mode, threads)
    if not g:
        raise ValueError("oracle: group-by rejected (key too wide?)")
    try:
        ng = L.orc_gb_n_groups(g)
        lo = np.zeros(max(1, ng), np.uint64)
        hi = np.zeros(max(1, ng), np.uint64)
        first = np.zeros(max(1, ng), np.uint32)
        ln = np.zeros(max(1, ng), np.uint32)
        L.orc_gb_lanes(g, lo.ctypes.data, hi.ctypes.data, first.ctypes.data, ln.ctypes.data)
        first, ln = first[:ng], ln[:ng]
        vals = {}
        for j, nm in enumerate(val_names):
            cnt = np.zeros(max(1, ng), np.uint32)
            si = np.zeros(max(1, ng), np.int64)
            mn = np.zeros(max(1, ng), np.int64)
            mx = np.zeros(max(1, ng), np.int64)
            sf = np.zeros(max(1, ng), np.float64)
            mnf = np.zeros(max(1, ng), np.float64)
            mxf = np.zeros(max(1, ng), np.float64)
            L.orc_gb_val(g, j, cnt.ctypes.data, si.ctypes.data, mn.ctypes.data, mx.ctypes.data, sf.ctypes.data,
                         mnf.ctypes.data, mxf.ctypes.data)
            vals[nm] = dict(cnt=cnt[:ng], si=si[:ng], mn=mn[:ng], mx=mx[:ng], sf=sf[:ng], mnf=mnf[:ng],
                            mxf=mxf[:ng])
    finally:
        L.orc_gb_free(g)
    out: dict = {}
    for nm in keys:
        tag, values, valid = frame[nm]
        kv = np.asarray(values)[first] if ng else np.asarray(values)[:0]
        vv = (np.ones(ng, bool) if valid is None else np.asarray(valid)[first]) if ng else np.zeros(0, bool)
        kv = np.where(vv, kv, "" if tag == "Utf8" else np.zeros_like(kv))
        if tag == "Utf8":
            kv = kv.astype(object)
        out[nm] = (tag, kv, vv)
    with np.errstate(invalid="ignore", divide="ignore"):
        for a in plan["aggs"]:
            op = a.get("op")
            alias = a["output_alias"]
            if op == "Len" or a.get("kind") == "Length":
                out[alias] = ("U32", ln.copy(), np.ones(ng, bool))
                continue
            nm = a["input_col"]
            tag = frame[nm][0]
            s = vals[nm]
            cnt = s["cnt"]
            has = cnt > 0
            fl = tag in ("F64", "F32")
            if op == "Count":
                out[alias] = ("U32", cnt.copy(), np.ones(ng, bool))
            elif op == "Sum":
                if fl:
                    v = s["sf"].astype(np.float32) if tag == "F32" else s["sf"].copy()
                elif tag == "I32":
                    v = s["si"].astype(np.int32)  # wrapping
                else:
                    v = s["si"].copy()
                out[alias] = (tag, v, np.ones(ng, bool))
            elif op == "Mean":
                num = s["sf"] if fl else s["si"].astype(np.float64)
                m = np.where(has, num / np.where(has, cnt, 1).astype(np.float64), 0.0)
                if tag == "F32":
                    out[alias] = ("F32", m.astype(np.float32), has)
                else:
                    out[alias] = ("F64", m, has)
            elif op in ("Min", "Max"):
                key = "mnf" if op == "Min" else "mxf"
                ikey = "mn" if op == "Min" else "mx"
                if fl:
                    v = np.where(has, s[key], 0.0)
                    v = v.astype(np.float32) if tag == "F32" else v
                else:
                    v = np.where(has, s[ikey], 0).astype(NP_DTYPE[tag])
                out[alias] = (tag, v, has)
            else:
                raise ValueError(op)
    out["__n_groups__"] = ng
    out["__ops__"] = {a["output_alias"]: ("Len" if a.get("kind") == "Length" else a.get("op")) for a in plan["aggs"]}
    return out


def from_wire(outputs: list) -> dict:
    """The reference's output tuples -> ``name -> (tag, values, valid)``."""
    res: dict = {}
    n = None
    for name, tag, data, valid in outputs:
        if tag == "Bool":
            continue
        if tag == "Utf8":
            n = int(np.frombuffer(data, np.uint32, 1)[0])
            break
        arr = np.frombuffer(data, dtype=NP_DTYPE[tag])
        n = len(arr)
        break
    if n is None:
        n = 0
    for name, tag, data, valid in outputs:
        if tag == "Bool":
            values = unpack_bits(data, n)
        elif tag == "Utf8":
            values = utf8_unpacked(data)
        else:
            values = np.frombuffer(data, dtype=NP_DTYPE[tag]).copy()
        res[name] = (tag, values, unpack_bits(valid, n))
    res["__n_groups__"] = n
    return res


def canonical(result: dict, keys: list) -> dict:
    """Sort a result by (key validity, key value) columns, nulls first."""
    n = result["__n_groups__"]
    if n == 0 or not keys:
        return result
    sort_cols = []
    for nm in reversed(keys):
        tag, v, ok = result[nm]
        vv = np.asarray(v)
        if vv.dtype.kind == "f":
            vv = vv.view(np.int64 if vv.dtype.itemsize == 8 else np.int32)
        sort_cols.append(np.where(ok, vv, "" if vv.dtype == object else 0))
        sort_cols.append(ok.astype(np.int8))
    order = np.lexsort(sort_cols)
    out = {}
    for nm, val in result.items():
        if nm.startswith("__"):
            out[nm] = val
            continue
        tag, v, ok = val
        out[nm] = (tag, np.asarray(v)[order], np.asarray(ok)[order])
    return out
