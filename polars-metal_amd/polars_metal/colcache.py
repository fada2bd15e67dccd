"""Device-resident Arrow columns for repeated queries.

The reference reads page-aligned Arrow buffers in place through its
zero-copy bridge (crates/polars-metal-buffer/src/bridge.rs:56-64) and
stages the rest through a reusable buffer (staging.rs:24-58).  On a discrete
MI355X the equivalent of "in place" is a device copy that outlives the
query: ``device_column`` uploads an Arrow column once (``pmh_column_resident``,
through the library's pinned staging ring) and finds it again on the next
query over the same buffers, copying nothing.

Identity.  An entry is keyed by the column's chunk buffer addresses, row
counts and bit offsets, plus a stamp.  Arrow buffers are immutable, and this
module keeps every cached chunk alive while its record exists, so an address
cannot be reused under a live record; each record gets a fresh stamp, so a
device entry left behind by a dropped record can never answer for new
buffers at a recycled address.

Chunked columns are assembled on the device (values DMA into place, bitmaps
re-based by a kernel), replacing the host concatenation of the reference's
``_materialize_arrow`` (python/polars_metal/_udf.py:1441-1456).

Use::

    col, entry = colcache.device_column("v", "F64", table.column("v"))
    try:
        ... execute with col (a PMH_DEVICE column tuple) ...
    finally:
        colcache.release([entry])
"""
from __future__ import annotations

import ctypes
import secrets
from collections import OrderedDict

from . import _lib
from ._lib import check

__all__ = ["device_column", "release", "evict_transient", "stats", "clear", "set_capacity"]


class _Record:
    __slots__ = ("stamp", "entry", "keep", "nbytes", "held")

    def __init__(self, stamp: int, keep: list, nbytes: int):
        self.stamp = stamp
        self.entry = 0
        self.keep = keep
        self.nbytes = nbytes
        self.held = 0


_RECORDS: "OrderedDict[tuple, _Record]" = OrderedDict()
_BY_ENTRY: dict = {}
_STAMP = [secrets.randbits(62) << 1]


def _next_stamp() -> int:
    _STAMP[0] = (_STAMP[0] + 2) & ((1 << 64) - 1)
    return _STAMP[0]


def _chunks(tag: str, col):
    """``(key, PmhChunk array, n_chunks, keep, bytes)`` of an Arrow Array or
    ChunkedArray: each chunk by address, with its offset folded into the
    pointers (and, for bitmaps, a bit offset 0..7)."""
    import pyarrow as pa
    arrays = col.chunks if isinstance(col, pa.ChunkedArray) else [col]
    arrays = [a for a in arrays if len(a)]
    width = _lib.TAG_SIZE[tag]
    parts, keep, nbytes = [], [], 0
    arr = (_lib.PmhChunk * max(1, len(arrays)))()
    for i, a in enumerate(arrays):
        off, n = a.offset, len(a)
        vbuf, dbuf = a.buffers()[0], a.buffers()[1]
        if tag == "Bool":
            dp, dbit = dbuf.address + off // 8, off % 8
            nbytes += (n + 7) // 8
        else:
            dp, dbit = dbuf.address + off * width, 0
            nbytes += n * width
        if vbuf is not None and a.null_count:
            vp, vbit = vbuf.address + off // 8, off % 8
            nbytes += (n + 7) // 8
        else:
            vp, vbit = 0, 0
        arr[i].n_rows, arr[i].data, arr[i].valid = n, dp, vp or None
        arr[i].data_bit_offset, arr[i].valid_bit_offset = dbit, vbit
        parts.append((n, dp, dbit, vp, vbit))
        keep.append(a)
    return (tag, tuple(parts)), arr, len(arrays), keep, nbytes


def _resident(tag: str, arr, n: int, stamp: int):
    out = _lib.PmhResidentColumn()
    check(_lib.load().pmh_column_resident(_native_tag(tag), arr, n, stamp, ctypes.byref(out)))
    return out


def _native_tag(tag: str) -> int:
    from ._native import _tag
    return _tag(tag)


def device_column(name: str, tag: str, col, transient: bool = False) -> tuple:
    """The column ``(name, tag, data, data_len, valid, valid_len, PMH_DEVICE)``
    for the fused entry, and the entry handle to ``release`` after the query.
    ``transient``: upload for this query only (a fresh stamp; the entry is
    dropped by ``evict_transient``), for chunked columns that should be
    assembled on the device but not kept."""
    key, arr, n, keep, nbytes = _chunks(tag, col)
    if transient:
        rec = _Record(_next_stamp(), keep, nbytes)
        out = _resident(tag, arr, n, rec.stamp)
    else:
        rec = _RECORDS.get(key)
        if rec is None:
            rec = _Record(_next_stamp(), keep, nbytes)
            _RECORDS[key] = rec
        _RECORDS.move_to_end(key)
        out = _resident(tag, arr, n, rec.stamp)
    if rec.entry and rec.entry != out.entry:
        _BY_ENTRY.pop(rec.entry, None)  # the library evicted the old copy and uploaded again
    rec.entry = out.entry
    rec.held += 1
    _BY_ENTRY[out.entry] = rec
    col_t = (name, tag, out.data or 0, out.data_len, out.valid or 0, out.valid_len if out.valid else 0,
             _lib.PMH_DEVICE)
    if not transient:
        _trim()
    return col_t, (out.entry, transient, bool(out.hit))


def release(handles) -> None:
    """Ends the hold of each ``device_column`` handle (transient entries are
    freed)."""
    lib = _lib.load()
    for entry, transient, _ in handles:
        check(lib.pmh_column_release(entry))
        rec = _BY_ENTRY.get(entry)
        if rec is not None:
            rec.held -= 1
        if transient:
            check(lib.pmh_column_evict(entry))
            _BY_ENTRY.pop(entry, None)


def evict_transient(handles) -> None:
    release([h for h in handles if h[1]])


def _trim() -> None:
    """Drops the least recently used unheld records beyond the library's
    capacity (host buffers are released with them)."""
    cap = stats()["capacity"]
    total = sum(r.nbytes for r in _RECORDS.values())
    lib = _lib.load()
    for key in list(_RECORDS):
        if total <= cap:
            break
        rec = _RECORDS[key]
        if rec.held:
            continue
        check(lib.pmh_column_evict(rec.entry))
        _BY_ENTRY.pop(rec.entry, None)
        del _RECORDS[key]
        total -= rec.nbytes


def stats() -> dict:
    info = _lib.PmhColcacheInfo()
    check(_lib.load().pmh_colcache_stats(ctypes.byref(info)))
    return {name: getattr(info, name) for name, _ in _lib.PmhColcacheInfo._fields_}


def set_capacity(nbytes: int) -> None:
    check(_lib.load().pmh_colcache_set_capacity(nbytes))
    _trim()


def clear() -> None:
    """Drops every unheld record and device entry."""
    lib = _lib.load()
    for key in list(_RECORDS):
        rec = _RECORDS[key]
        if rec.held:
            continue
        check(lib.pmh_column_evict(rec.entry))
        _BY_ENTRY.pop(rec.entry, None)
        del _RECORDS[key]
    check(lib.pmh_colcache_clear())
