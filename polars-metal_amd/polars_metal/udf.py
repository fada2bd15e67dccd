"""Group-by dispatch of the engine hook over Arrow tables.

``dispatch_groupby(table, wire_plan)`` does for a ``pyarrow.Table`` what the
reference's ``_udf._dispatch_groupby`` (python/polars_metal/_udf.py:955-1127)
does for a Polars frame: resolve the GroupBy's upstream, hand the key and
aggregate columns to the backend, and reassemble the result.  Three things
differ, each on purpose:

  * The upstream ``Filter`` is not compacted on the host.  The reference
    compacts through Polars CPU (``_filter_via_polars``, _udf.py:550-597)
    and then groups the survivors; here the walker's PredicateAst goes into
    the plan and one fused device call filters, groups and aggregates.
    Nested Filters are conjoined, a ``Project`` between them passes
    through, and a predicate Polars pushed into the scan
    (``Scan["selection"]``, which the reference's walker refuses at
    _walker.py:195-196) is lowered the same way.
  * Columns are not copied into ``bytes`` (_udf.py:1068-1084).  Each Arrow
    buffer is passed by address; the backend uploads it through its pinned
    staging ring (or, ``cache=True``, serves it from its device column
    cache, so a repeated query copies nothing).  A column without a
    validity buffer is passed with none, which the backend reads as all
    valid (_udf.py:1574-1582 materialises all-ones bytes instead).
  * A column of several chunks is passed chunk by chunk and assembled on
    the device (kept resident with ``cache=True``, dropped after the query
    otherwise), where the reference's ``_materialize_arrow`` concatenates
    on the host (_udf.py:1441-1456).

  * Expression aggregates (``sum(a * (1 - b))``) are not evaluated through
    Polars into temporary columns (_udf.py:1164-1209); the AggExpr tree goes
    into the plan and the device evaluates it inside the same call.

Shapes this module does not lower (HStack, other upstream nodes) raise
NotImplementedError; ``polars_metal.install`` then hands them to the
reference's own dispatch.
"""
from __future__ import annotations

from typing import Any

from . import _lib, _native, colcache

__all__ = ["dispatch_groupby", "dispatch_groupby_pydf", "resolve_upstream", "arrow_tag"]


def arrow_tag(pa_type) -> str | None:
    """Wire tag for a pyarrow type (the reference's ``_DTYPE_TO_TAG``,
    _udf.py:1415-1425, widened to every fixed-width type the kernels take;
    Date32 travels as I32, plan/mod.rs:78-101)."""
    import pyarrow as pa
    table = {pa.int64(): "I64", pa.float64(): "F64", pa.bool_(): "Bool", pa.int32(): "I32",
             pa.float32(): "F32", pa.int8(): "I8", pa.int16(): "I16", pa.uint8(): "U8", pa.uint16(): "U16",
             pa.uint32(): "U32", pa.uint64(): "U64", pa.string(): "Utf8", pa.large_string(): "Utf8",
             pa.date32(): "I32"}
    return table.get(pa_type)


def _conjoin(a: dict | None, b: dict | None) -> dict | None:
    if a is None:
        return b
    if b is None:
        return a
    return {"kind": "And", "lhs": a, "rhs": b}


def resolve_upstream(node: dict) -> tuple[dict | None, list | None]:
    """The GroupBy's input subtree -> (predicate or None, projected column
    names or None).  Accepts Scan (with an optional pushed-down
    ``selection``), Filter and Project over any of those."""
    kind = node.get("kind")
    if kind == "Scan":
        return node.get("selection"), node.get("projection")
    if kind == "Filter":
        pred, proj = resolve_upstream(node["input"])
        return _conjoin(pred, node["predicate"]), proj
    if kind == "Project":
        pred, _ = resolve_upstream(node["input"])
        return pred, list(node["columns"])
    raise NotImplementedError(f"polars_metal: GroupBy over {kind!r} is not lowered here")


def _needed_columns(plan: dict, predicate: dict | None) -> list:
    needed: list = []

    def add(name):
        if name and name not in needed:
            needed.append(name)

    for k in plan["keys"]:
        add(k[0])
    def walk(p):
        if p is None:
            return
        if p.get("kind") == "Column":
            add(p["name"])
        walk(p.get("lhs"))
        walk(p.get("rhs"))

    for a in plan["aggs"]:
        if a.get("kind") == "Length" or a.get("op") == "Len":
            continue
        if a.get("kind") == "Expression":
            walk(a.get("expr"))  # evaluated on the device inside the call
        else:
            add(a.get("input_col"))
    walk(predicate)
    return needed


class _Columns:
    """Arrow columns as pmh_column descriptors over their own buffers."""

    def __init__(self):
        self.cols: list = []
        self.keep: list = []
        self.dicts: dict = {}

    def add_fixed(self, name: str, tag: str, arr) -> None:
        """A single-chunk fixed-width (or Bool) array, by buffer address."""
        import pyarrow as pa
        width = _lib.TAG_SIZE[tag]
        off = arr.offset
        if (off & 7) and (arr.null_count or tag == "Bool"):
            arr = pa.concat_arrays([arr])  # bit offset inside a byte: re-base once
            off = 0
        vbuf, dbuf = arr.buffers()[0], arr.buffers()[1]
        n = len(arr)
        self.keep.append(arr)
        if tag == "Bool":
            dp, dl = dbuf.address + off // 8, (n + 7) // 8
        else:
            dp, dl = dbuf.address + off * width, n * width
        vp, vl = (vbuf.address + off // 8, (n + 7) // 8) if (vbuf is not None and arr.null_count) else (0, 0)
        self.cols.append((name, tag, dp, dl, vp, vl))

    def add_utf8(self, name: str, arr) -> None:
        """Utf8 keys group on first-seen dictionary codes built on the
        device (the reference's build_dict_nullable,
        crates/polars-metal-buffer/src/dict.rs:40-66, runs on the host):
        null rows get code 0 and stay null; "" is a value of its own.  The
        Arrow buffers are read in place."""
        d = _native.Utf8Dict.from_arrow(arr)
        self.keep.append(d)
        self.dicts[name] = d
        self.cols.append(d.column(name))


def _single(col):
    import pyarrow as pa
    if isinstance(col, pa.ChunkedArray):
        return col.chunk(0) if col.num_chunks == 1 else (col.combine_chunks() if col.num_chunks else
                                                        pa.array([], type=col.type))
    return col


def _out_array(tag: str, data: bytes, valid: bytes, n: int):
    """One output column as a pyarrow array (the reference's
    ``_assemble_series``, _udf.py:1585-1669, minus the Polars wrap); a Utf8
    key arrives in the packed wire form and is wrapped without a copy."""
    import pyarrow as pa
    nb = (n + 7) // 8
    vb = pa.py_buffer(valid[:nb]) if n else None
    if tag == "Utf8":
        buf = pa.py_buffer(data)
        return pa.Array.from_buffers(pa.string(), n, [vb, buf.slice(4, 4 * (n + 1)), buf.slice(4 + 4 * (n + 1))],
                                     null_count=-1)
    pa_type = {"I64": pa.int64(), "F64": pa.float64(), "Bool": pa.bool_(), "I32": pa.int32(),
               "F32": pa.float32(), "I8": pa.int8(), "I16": pa.int16(), "U8": pa.uint8(), "U16": pa.uint16(),
               "U32": pa.uint32(), "U64": pa.uint64()}[tag]
    if not n:
        return pa.array([], type=pa_type)
    width = _lib.TAG_SIZE[tag]
    db = pa.py_buffer(data[:nb] if tag == "Bool" else data[:n * width])
    return pa.Array.from_buffers(pa_type, n, [vb, db], null_count=-1)


def dispatch_groupby(table, wire_plan: dict, cache: bool = False):
    """Run a GroupBy wire plan (``{"kind": "GroupBy", "input": ..., "keys",
    "aggs"}``) over a ``pyarrow.Table``; returns a ``pyarrow.Table`` with the
    keys, then the aggregates (group order unspecified, as the reference's).
    ``cache=True`` serves the input columns from the device column cache."""
    import pyarrow as pa
    predicate, _ = resolve_upstream(wire_plan["input"]) if "input" in wire_plan else (None, None)
    plan = {"keys": [list(k) for k in wire_plan["keys"]], "aggs": list(wire_plan["aggs"])}
    if "options" in wire_plan:
        plan["options"] = wire_plan["options"]
    if predicate is not None:
        plan["predicate"] = predicate
    needed = _needed_columns(plan, predicate)
    n_rows = table.num_rows
    cols = _Columns()
    handles: list = []
    try:
        for name in needed:
            if name not in table.column_names:
                raise RuntimeError(f"polars_metal: column {name!r} required by GroupBy not in upstream df")
            col = table.column(name)
            tag = arrow_tag(col.type)
            if tag is None:
                raise RuntimeError(f"polars_metal: unsupported dtype {col.type} on column {name!r}")
            if tag == "Utf8":
                cols.add_utf8(name, _single(col))
            elif cache or col.num_chunks > 1:
                c, h = colcache.device_column(name, tag, col, transient=not cache)
                cols.cols.append(c)
                handles.append(h)
            else:
                cols.add_fixed(name, tag, _single(col))
        for k in plan["keys"]:
            if k[1] == "String":
                k[1] = "Utf8"
            elif k[1] == "Date":
                k[1] = "I32"
        out = _native.execute_filter_groupby_ptrs(plan, n_rows, cols.cols, _lib.PMH_HOST, cols.dicts)
    finally:
        colcache.release(handles)
    n = 0
    for _, tag, data, _ in out:
        if tag == "Utf8":
            n = int.from_bytes(data[:4], "little")
            break
        if tag != "Bool":
            n = len(data) // _lib.TAG_SIZE[tag]
            break
    arrays = [_out_array(tag, data, valid, n) for name, tag, data, valid in out]
    return pa.table(arrays, names=[c[0] for c in out])


def dispatch_groupby_pydf(df_pydf: Any, wire_plan: dict):
    """The reference's ``_dispatch_groupby(df_pydf, wire_plan)`` signature
    (_udf.py:955): a PyDataFrame in, a Polars DataFrame out."""
    predicate = resolve_upstream(wire_plan["input"])  # unsupported shapes raise before Polars is touched
    import polars as pl  # only the engine hook needs Polars
    df = pl.DataFrame._from_pydf(df_pydf)
    return pl.from_arrow(dispatch_groupby(df.to_arrow(), wire_plan))
