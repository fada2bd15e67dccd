"""Python face of the MI355X backend, signature-compatible with the
reference's PyO3 module ``polars_metal._native``
(crates/polars-metal-core/src/lib.rs:57-83).

Every function marshals Python ``bytes`` into ``pmh_column`` descriptors,
calls the C ABI (include/pmh.h) and returns the reference's tuple layouts.
Errors surface as the reference's exception classes (ValueError, KeyError,
NotImplementedError, RuntimeError, polars ComputeError).

Two additions have no reference counterpart:
  * ``execute_filter_groupby`` -- the fused filter + group-by entry that
    replaces the Polars-CPU compaction on the GroupBy(Filter(Scan)) shape
    (python/polars_metal/_udf.py:550-597);
  * ``execute_filter_groupby_device`` -- the same call on device-resident
    columns (raw device pointers), leaving outputs in device memory.
"""
from __future__ import annotations

import ctypes
import json
import math
import struct
from dataclasses import dataclass
from typing import Any, Iterable, Sequence

from . import _lib
from ._lib import PMH_DEVICE, PMH_HOST, PMH_OUT_DEVICE, PMH_OUT_HOST, TAGS, TAG_ALIASES, TAG_NAMES, check

__all__ = [
    "version_string", "device_name", "execute_groupby", "execute_filter_groupby",
    "execute_filter_groupby_device", "execute_filter_groupby_ptrs", "execute_filter_groupby_noalloc",
    "execute_filter_compact", "cmp_i64_col_scalar",
    "cmp_i64_col_col", "cmp_f64_col_scalar", "cmp_f64_col_col", "bool_and_dispatch",
    "bool_or_dispatch", "compute_lifting_plan", "warmup_common_fused_signatures",
    "set_profiling", "kernel_times", "last_stats",
]


def _tag(t: str) -> int:
    t = TAG_ALIASES.get(t, t)
    if t not in TAGS:
        raise ValueError(f"polars_metal: unknown MetalDtype tag {t!r}")
    return TAGS[t]


def _jsonable(obj: Any) -> Any:
    if isinstance(obj, float) and not math.isfinite(obj):
        return obj  # json.dumps writes NaN / Infinity, which the C reader accepts
    if isinstance(obj, tuple):
        return [_jsonable(x) for x in obj]
    if isinstance(obj, list):
        return [_jsonable(x) for x in obj]
    if isinstance(obj, dict):
        return {str(k): _jsonable(v) for k, v in obj.items()}
    return obj


def _plan_json(plan: dict) -> bytes:
    if not isinstance(plan, dict):
        raise ValueError("polars_metal: plan parse error: plan must be a dict")
    try:
        # json writes tuples as lists and stringifies int/float/bool keys
        # itself; the normalising copy is only needed for other key types
        return json.dumps(plan, allow_nan=True).encode("utf-8")
    except TypeError:
        return json.dumps(_jsonable(plan), allow_nan=True).encode("utf-8")


# ---------------------------------------------------------------------------
# Utf8 keys: the reference builds a dictionary on the host
# (crates/polars-metal-core/src/udf/groupby/mod.rs:818-936) and groups on u32
# codes; here the library builds it on the device (pmh_utf8_encode) and the
# key column of the result decodes there too (pmh_utf8_decode).
# ---------------------------------------------------------------------------
class Utf8Dict:
    """A Utf8 column's device dictionary: ``column(name)`` is its codes as a
    PMH_DEVICE key column (tag Utf8), ``decode`` turns result codes back
    into the packed wire form ``[n u32][(n+1) x i32 offsets][bytes]``."""

    def __init__(self, col: "_lib.PmhUtf8Column", keep=None):
        self._keep = keep
        h = ctypes.c_void_p()
        check(_lib.load().pmh_utf8_encode(ctypes.byref(col), ctypes.byref(h)))
        self._h = h
        codes, valid = ctypes.c_void_p(), ctypes.c_void_p()
        nr, nd = ctypes.c_uint64(), ctypes.c_uint64()
        check(_lib.load().pmh_utf8_dict_codes(h, ctypes.byref(codes), ctypes.byref(valid), ctypes.byref(nr),
                                              ctypes.byref(nd)))
        self.codes, self.valid = codes.value or 0, valid.value or 0
        self.n_rows, self.n_distinct = nr.value, nd.value
        self._keep = None  # the library holds its own device copy from here on

    @classmethod
    def from_packed(cls, data: bytes, valid: bytes | None, n_rows: int) -> "Utf8Dict":
        """The reference's packed wire form (python/polars_metal/_udf.py:1497-1548),
        checked as the reference checks it (mod.rs:818-875)."""
        if len(data) < 4:
            raise ValueError(f"polars_metal: Utf8 column wire payload too short ({len(data)} B; need >= 4 for header)")
        (n,) = struct.unpack_from("<I", data, 0)
        if n != n_rows:
            raise ValueError(f"polars_metal: Utf8 column wire n_rows={n} disagrees with column n_rows={n_rows}")
        if len(data) < 4 + 4 * (n + 1):
            raise ValueError("polars_metal: Utf8 column wire payload truncated in offsets")
        if valid is not None and len(valid) < (n + 7) // 8:
            raise ValueError(f"polars_metal: Utf8 column validity buffer is {len(valid)} B, need at least "
                             f"{(n + 7) // 8} B for {n} rows")
        base = ctypes.cast(ctypes.c_char_p(data), ctypes.c_void_p).value
        col = _lib.PmhUtf8Column()
        col.n_rows, col.location, col.large = n, PMH_HOST, 0
        col.offsets, col.bytes = base + 4, base + 4 + 4 * (n + 1)
        col.bytes_len = len(data) - 4 - 4 * (n + 1)
        col.valid = ctypes.cast(ctypes.c_char_p(valid), ctypes.c_void_p).value if valid else None
        return cls(col, (data, valid))

    @classmethod
    def from_arrow(cls, arr) -> "Utf8Dict":
        """A pyarrow string or large_string Array, read in place (its slice
        offset folded into the offsets pointer and a validity bit offset)."""
        import pyarrow as pa
        large = pa.types.is_large_string(arr.type)
        ow = 8 if large else 4
        vbuf, obuf, dbuf = arr.buffers()
        off = arr.offset
        col = _lib.PmhUtf8Column()
        col.n_rows, col.location, col.large = len(arr), PMH_HOST, int(large)
        col.offsets = obuf.address + ow * off
        col.bytes = dbuf.address if dbuf is not None else None
        col.bytes_len = dbuf.size if dbuf is not None else 0
        if vbuf is not None and arr.null_count:
            col.valid, col.valid_bit_offset = vbuf.address + off // 8, off % 8
        return cls(col, arr)

    def column(self, name: str) -> tuple:
        return (name, "Utf8", self.codes, self.n_rows * 4, self.valid, (self.n_rows + 7) // 8 if self.valid else 0,
                PMH_DEVICE)

    def decode(self, codes: bytes, valid: bytes | None, n: int) -> bytes:
        """Host codes (and validity) of n result rows -> packed wire form."""
        out, ln = ctypes.c_void_p(), ctypes.c_uint64()
        cp = ctypes.cast(ctypes.c_char_p(codes), ctypes.c_void_p).value if n else None
        vp = ctypes.cast(ctypes.c_char_p(valid), ctypes.c_void_p).value if valid else None
        check(_lib.load().pmh_utf8_decode(self._h, cp, vp, n, PMH_HOST, ctypes.byref(out), ctypes.byref(ln)))
        try:
            return ctypes.string_at(out.value, ln.value)
        finally:
            _lib.load().pmh_string_free(out)

    def free(self) -> None:
        if self._h:
            _lib.load().pmh_utf8_dict_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclass
class _Prepared:
    cols: ctypes.Array
    keep: list
    dicts: dict


def _prepare_columns(columns: Sequence[tuple], n_rows: int) -> _Prepared:
    arr = (_lib.PmhColumn * max(1, len(columns)))()
    keep: list = []
    dicts: dict = {}
    for i, col in enumerate(columns):
        if len(col) != 4:
            raise ValueError("polars_metal: columns must be (name, dtype_tag, data, valid) tuples")
        name, tag, data, valid = col
        tag = TAG_ALIASES.get(tag, tag)
        t = _tag(tag)
        if tag == "Utf8":
            d = Utf8Dict.from_packed(bytes(data), bytes(valid) if valid else None, n_rows)
            dicts[name] = d
            nb = name.encode("utf-8")
            keep.append(nb)
            arr[i].name, arr[i].tag, arr[i].location = nb, t, PMH_DEVICE
            arr[i].data, arr[i].data_len = d.codes or None, n_rows * 4
            arr[i].valid, arr[i].valid_len = (d.valid or None), ((n_rows + 7) // 8 if d.valid else 0)
            continue
        data = bytes(data) if not isinstance(data, bytes) else data
        valid = bytes(valid) if valid is not None and not isinstance(valid, bytes) else valid
        nb = name.encode("utf-8")
        keep += [nb, data, valid]
        arr[i].name = nb
        arr[i].tag = t
        arr[i].location = PMH_HOST
        arr[i].data = ctypes.cast(ctypes.c_char_p(data), ctypes.c_void_p) if data else None
        arr[i].data_len = len(data)
        if valid:
            arr[i].valid = ctypes.cast(ctypes.c_char_p(valid), ctypes.c_void_p)
            arr[i].valid_len = len(valid)
        else:
            arr[i].valid = None
            arr[i].valid_len = 0
    return _Prepared(arr, keep, dicts)


_LAST_STATS: dict = {}


def last_stats() -> dict:
    """Execution statistics of the last group-by call (method, passes,
    capacities, group and selected-row counts, device milliseconds)."""
    return dict(_LAST_STATS)


def _stats_dict(st: _lib.PmhExecStats) -> dict:
    return {
        "method": "sort" if st.method == 1 else "hash",
        "n_passes": st.n_passes,
        "initial_capacity": st.initial_capacity,
        "final_capacity": st.final_capacity,
        "n_groups": st.n_groups,
        "n_selected": st.n_selected,
        "device_ms": st.device_ms,
        "sample_rows": st.sample_rows,
        "sample_distinct": st.sample_distinct,
        "sample_f1": st.sample_f1,
        "sample_f2": st.sample_f2,
        "hot_keys": st.hot_keys,
        "partitions": st.partitions,
        "tail_rows": st.tail_rows,
        "distinct_estimate": st.distinct_estimate,
        "scan_kernel": "specialised" if st.scan_kernel == 1 else "generic",
        "sizer_choice": {0: None, 1: "hash", 2: "sort"}.get(st.sizer_choice),
    }


def _collect_host_result(res_p, dicts: dict) -> list:
    res = res_p.contents
    out = []
    n = res.n_rows
    for i in range(res.n_columns):
        c = res.columns[i]
        name = c.name.decode("utf-8")
        tag = TAG_NAMES[c.tag]
        data = ctypes.string_at(c.data, c.data_len) if c.data_len else b""
        valid = ctypes.string_at(c.valid, c.valid_len) if c.valid_len else b""
        if tag == "Utf8":
            data = dicts[name].decode(data, valid, n)
        out.append((name, tag, data, valid))
    return out


def _run(entry: str, plan: dict, n_rows: int, columns: Sequence[tuple], flags: int | None) -> list:
    lib = _lib.load()
    prep = _prepare_columns(columns, n_rows)
    res_p = ctypes.POINTER(_lib.PmhResult)()
    pj = _plan_json(plan)
    if entry == "pmh_execute_groupby":
        st = lib.pmh_execute_groupby(pj, n_rows, prep.cols, len(columns), ctypes.byref(res_p))
    else:
        st = lib.pmh_execute_filter_groupby(pj, n_rows, prep.cols, len(columns), flags or 0, ctypes.byref(res_p))
    check(st)
    try:
        _LAST_STATS.clear()
        _LAST_STATS.update(_stats_dict(res_p.contents.stats))
        return _collect_host_result(res_p, prep.dicts)
    finally:
        lib.pmh_result_free(res_p)


def version_string() -> str:
    """lib.rs:31-35."""
    return _lib.load().pmh_version_string().decode()


def device_name() -> str:
    """lib.rs:36-40."""
    buf = ctypes.create_string_buffer(256)
    check(_lib.load().pmh_device_name(buf, 256))
    return buf.value.decode()


def execute_groupby(plan_dict: dict, n_rows: int, columns: Sequence[tuple]) -> list:
    """crates/polars-metal-core/src/udf/groupby/mod.rs:747-1144."""
    return _run("pmh_execute_groupby", plan_dict, n_rows, columns, None)


def execute_filter_groupby(plan_dict: dict, n_rows: int, columns: Sequence[tuple]) -> list:
    """Fused predicate + group-by: ``plan_dict`` is an execute_groupby plan
    with an optional ``"predicate"`` in PredicateAst dict form
    (crates/polars-metal-core/src/udf/predicate.rs:92-200)."""
    return _run("pmh_execute_filter_groupby", plan_dict, n_rows, columns, PMH_OUT_HOST)


class DeviceResult:
    """Outputs of execute_filter_groupby_device (or of one batch query).
    Columns are pointers owned by the library, valid until the next call or
    free(); ``columns`` and ``stats`` are read from the result on first use."""

    __slots__ = ("_p", "_columns", "_stats")

    def __init__(self, res_p):
        self._p = res_p
        self._columns = None
        self._stats = None

    @property
    def n_rows(self) -> int:
        return self._p.contents.n_rows

    @property
    def stats(self) -> dict:
        if self._stats is None:
            self._stats = _stats_dict(self._p.contents.stats)
        return self._stats

    @property
    def columns(self) -> list:
        if self._columns is None:
            r = self._p.contents
            cols = []
            for i in range(r.n_columns):
                c = r.columns[i]
                cols.append({
                    "name": c.name.decode(), "tag": TAG_NAMES[c.tag], "data": c.data or 0,
                    "data_len": c.data_len, "valid": c.valid or 0, "valid_len": c.valid_len,
                    "location": c.location,
                })
            self._columns = cols
        return self._columns

    def fetch(self, name: str):
        """Copy one output column to host: ``(tag, data bytes, valid bytes
        or None)``."""
        for c in self.columns:
            if c["name"] == name:
                if c["location"] == PMH_HOST:  # a batch query re-run alone
                    data = ctypes.string_at(c["data"], c["data_len"]) if c["data_len"] else b""
                    valid = ctypes.string_at(c["valid"], c["valid_len"]) if c["valid"] else None
                    return c["tag"], data, valid
                data = _device_to_host(c["data"], c["data_len"])
                valid = _device_to_host(c["valid"], c["valid_len"]) if c["valid"] else None
                return c["tag"], data, valid
        raise KeyError(name)

    def free(self) -> None:
        if self._p:
            _lib.load().pmh_result_free(self._p)
            self._p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _device_to_host(ptr: int, nbytes: int) -> bytes:
    """Copy device memory into a new bytes object."""
    if nbytes == 0:
        return b""
    buf = ctypes.create_string_buffer(nbytes)
    check(_lib.load().pmh_buffer_to_host(ctypes.c_void_p(ptr), buf, nbytes))
    return buf.raw


def _device_column_array(columns: list, location: int = PMH_DEVICE):
    """PmhColumn array for ``(name, tag, data_ptr, data_len, valid_ptr,
    valid_len)`` columns (device pointers, or host addresses with
    ``location=PMH_HOST``; a seventh element gives one column's own
    location), and the objects that must outlive it."""
    arr = (_lib.PmhColumn * max(1, len(columns)))()
    keep = []
    for i, c in enumerate(columns):
        name, tag, dp, dl, vp, vl = c[:6]
        nb = name.encode()
        keep.append(nb)
        arr[i].name = nb
        arr[i].tag = _tag(tag)
        arr[i].location = c[6] if len(c) > 6 else location
        arr[i].data = dp
        arr[i].data_len = dl
        arr[i].valid = vp or None
        arr[i].valid_len = vl if vp else 0
    return arr, keep


def execute_filter_groupby_device(plan_dict: dict, n_rows: int, columns: Iterable[tuple]) -> DeviceResult:
    """``columns``: ``(name, tag, data_ptr, data_len, valid_ptr, valid_len)``
    with device pointers (``valid_ptr`` 0 for no validity)."""
    lib = _lib.load()
    columns = list(columns)
    arr, keep = _device_column_array(columns)
    res_p = ctypes.POINTER(_lib.PmhResult)()
    check(lib.pmh_execute_filter_groupby(_plan_json(plan_dict), n_rows, arr, len(columns), PMH_OUT_DEVICE,
                                         ctypes.byref(res_p)))
    return DeviceResult(res_p)


def execute_filter_groupby_ptrs(plan_dict: dict, n_rows: int, columns: Iterable[tuple],
                                location: int = PMH_HOST, dicts: dict | None = None) -> list:
    """The fused call over columns given by address, ``(name, tag, data_ptr,
    data_len, valid_ptr, valid_len)`` (host memory by default, device
    memory with ``location=PMH_DEVICE``); outputs come back as the
    reference's host tuples, Utf8 keys decoded through ``dicts`` (name ->
    Utf8Dict) into the packed wire form.  Nothing is copied on the Python
    side."""
    lib = _lib.load()
    columns = list(columns)
    arr, keep = _device_column_array(columns, location)
    res_p = ctypes.POINTER(_lib.PmhResult)()
    check(lib.pmh_execute_filter_groupby(_plan_json(plan_dict), n_rows, arr, len(columns), PMH_OUT_HOST,
                                         ctypes.byref(res_p)))
    try:
        _LAST_STATS.clear()
        _LAST_STATS.update(_stats_dict(res_p.contents.stats))
        return _collect_host_result(res_p, dicts or {})
    finally:
        lib.pmh_result_free(res_p)


def execute_filter_groupby_noalloc(plan_dict: dict, n_rows: int, columns: Iterable[tuple], outputs: Sequence[tuple]):
    """Allocation-free fused call (pmh_execute_filter_groupby_noalloc).
    ``columns`` as for execute_filter_groupby_device; ``outputs`` one
    ``(data_ptr, data_len, valid_ptr, valid_len)`` caller-owned device
    buffer per output column (keys, then aggregates; valid_ptr 0 for none).
    Returns ``(n_groups, [(tag, data_len, valid_len), ...])``."""
    lib = _lib.load()
    columns = list(columns)
    arr, keep = _device_column_array(columns)
    outs = (_lib.PmhOutColumn * max(1, len(outputs)))()
    for i, (dp, dl, vp, vl) in enumerate(outputs):
        outs[i].data, outs[i].data_len = dp or None, dl
        outs[i].valid, outs[i].valid_len = vp or None, vl if vp else 0
    ng = ctypes.c_uint64(0)
    st = lib.pmh_execute_filter_groupby_noalloc(_plan_json(plan_dict), n_rows, arr, len(columns), outs, len(outputs),
                                                ctypes.byref(ng))
    if st != _lib.PMH_OK:
        try:
            check(st)
        except ValueError as e:
            e.n_groups = ng.value
            raise
    return ng.value, [(TAG_NAMES[outs[i].tag], outs[i].data_len, outs[i].valid_len) for i in range(len(outputs))]


def execute_filter_compact(pred_data: bytes, pred_valid: bytes, n_rows: int, columns: Sequence[tuple]) -> list:
    """crates/polars-metal-core/src/udf/compact.rs:46-147.  ``columns`` are
    ``(name, tag, data, valid)``; returns ``[(data, valid, n_out), ...]``
    with empty bytes when nothing survives (compact.rs:188-190)."""
    lib = _lib.load()
    prep = _prepare_columns(columns, n_rows)
    res_p = ctypes.POINTER(_lib.PmhResult)()
    pd = bytes(pred_data)
    pv = bytes(pred_valid)
    check(lib.pmh_execute_filter_compact(pd, pv, n_rows, prep.cols, len(columns), ctypes.byref(res_p)))
    try:
        r = res_p.contents
        n_out = r.n_rows
        out = []
        for i in range(r.n_columns):
            c = r.columns[i]
            if n_out == 0:
                out.append((b"", b"", 0))
                continue
            out.append((ctypes.string_at(c.data, c.data_len), ctypes.string_at(c.valid, c.valid_len), n_out))
        return out
    finally:
        lib.pmh_result_free(res_p)


def _bitmap_len(n: int) -> int:
    return max(4, (((n + 7) // 8) + 3) & ~3)


def _cmp(fn: str, lhs: bytes, lv: bytes | None, rhs, rv: bytes | None, op: str, n_rows: int, col_col: bool):
    if op not in _lib.CMP_OPS:
        raise ValueError(f"polars_metal: unknown compare op {op!r}")
    lib = _lib.load()
    if len(lhs) < 8 * n_rows:
        raise ValueError("polars_metal: lhs buffer too short")
    if col_col and len(rhs) < 8 * n_rows:
        raise ValueError("polars_metal: rhs buffer too short")
    nb = _bitmap_len(n_rows)
    od = ctypes.create_string_buffer(nb)
    ov = ctypes.create_string_buffer(nb)
    lvb = bytes(lv) if lv else None
    if lvb is not None and len(lvb) < (n_rows + 7) // 8:
        raise ValueError("polars_metal: lhs validity buffer too short")
    if col_col:
        rvb = bytes(rv) if rv else None
        st = getattr(lib, fn)(bytes(lhs), lvb, bytes(rhs), rvb, _lib.CMP_OPS[op], n_rows, od, ov)
    else:
        st = getattr(lib, fn)(bytes(lhs), lvb, rhs, _lib.CMP_OPS[op], n_rows, od, ov)
    check(st)
    return od.raw, ov.raw


def cmp_i64_col_scalar(lhs_data: bytes, lhs_valid: bytes, rhs: int, op: str, n_rows: int):
    """crates/polars-metal-core/src/udf/compare.rs:23-145."""
    return _cmp("pmh_cmp_i64_col_scalar", lhs_data, lhs_valid, int(rhs), None, op, n_rows, False)


def cmp_f64_col_scalar(lhs_data: bytes, lhs_valid: bytes, rhs: float, op: str, n_rows: int):
    return _cmp("pmh_cmp_f64_col_scalar", lhs_data, lhs_valid, float(rhs), None, op, n_rows, False)


def cmp_i64_col_col(lhs_data: bytes, lhs_valid: bytes, rhs_data: bytes, rhs_valid: bytes, op: str, n_rows: int):
    return _cmp("pmh_cmp_i64_col_col", lhs_data, lhs_valid, rhs_data, rhs_valid, op, n_rows, True)


def cmp_f64_col_col(lhs_data: bytes, lhs_valid: bytes, rhs_data: bytes, rhs_valid: bytes, op: str, n_rows: int):
    return _cmp("pmh_cmp_f64_col_col", lhs_data, lhs_valid, rhs_data, rhs_valid, op, n_rows, True)


def _bool(fn: str, ld: bytes, lv: bytes, rd: bytes, rv: bytes, n_rows: int):
    lib = _lib.load()
    need = (n_rows + 7) // 8
    for b in (ld, rd):
        if len(b) < need:
            raise ValueError("polars_metal: bool data buffer too short")
    nb = _bitmap_len(n_rows)
    od = ctypes.create_string_buffer(nb)
    ov = ctypes.create_string_buffer(nb)
    check(getattr(lib, fn)(bytes(ld), bytes(lv) if lv else None, bytes(rd), bytes(rv) if rv else None,
                           n_rows, od, ov))
    return od.raw, ov.raw


def bool_and_dispatch(ld: bytes, lv: bytes, rd: bytes, rv: bytes, n_rows: int):
    """crates/polars-metal-core/src/udf/logical.rs:26."""
    return _bool("pmh_bool_and", ld, lv, rd, rv, n_rows)


def bool_or_dispatch(ld: bytes, lv: bytes, rd: bytes, rv: bytes, n_rows: int):
    """crates/polars-metal-core/src/udf/logical.rs:44."""
    return _bool("pmh_bool_or", ld, lv, rd, rv, n_rows)


def compute_lifting_plan(plan_dict: dict) -> dict:
    """crates/polars-metal-core/src/router_udf.rs:22-40 (no device needed)."""
    lib = _lib.load()
    out = ctypes.c_void_p()
    check(lib.pmh_compute_lifting_plan(_plan_json(plan_dict), ctypes.byref(out)))
    return json.loads(_lib.take_string(out))


def warmup_common_fused_signatures() -> int:
    """crates/polars-metal-core/src/udf/groupby/mod.rs:187-201."""
    n = _lib.load().pmh_warmup_common_fused_signatures()
    if n < 0:
        check(-n)
    return n


def set_profiling(enabled: bool) -> None:
    check(_lib.load().pmh_set_profiling(1 if enabled else 0))


def kernel_times() -> dict:
    """``{kernel: (launches, total_ms)}`` measured with HIP events on the
    library's stream while profiling is enabled."""
    out = ctypes.c_void_p()
    check(_lib.load().pmh_kernel_times(ctypes.byref(out)))
    return {k: tuple(v) for k, v in json.loads(_lib.take_string(out)).items()}
