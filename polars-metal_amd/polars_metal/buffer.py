"""hipMalloc-backed Arrow buffers (the MI355X rework of the reference's
polars-metal-buffer crate, crates/polars-metal-buffer/src/device.rs:55-153).

``DeviceBuffer`` owns one device allocation made by libpmh.so's runtime, so
device columns can be built without any other GPU library in the process.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check


class DeviceBuffer:
    def __init__(self, ptr: int, nbytes: int):
        self.ptr = ptr
        self.nbytes = nbytes

    @classmethod
    def alloc(cls, nbytes: int, zeroed: bool = False) -> "DeviceBuffer":
        p = ctypes.c_void_p()
        fn = _lib.load().pmh_buffer_zeroed if zeroed else _lib.load().pmh_buffer_alloc
        check(fn(nbytes, ctypes.byref(p)))
        return cls(p.value or 0, nbytes)

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        check(_lib.load().pmh_buffer_from_host(a.ctypes.data if a.nbytes else None, a.nbytes, ctypes.byref(p)))
        return cls(p.value or 0, a.nbytes)

    def to_numpy(self, dtype, count: int | None = None) -> np.ndarray:
        dtype = np.dtype(dtype)
        n = self.nbytes // dtype.itemsize if count is None else count
        out = np.empty(n, dtype)
        if out.nbytes:
            check(_lib.load().pmh_buffer_to_host(ctypes.c_void_p(self.ptr), out.ctypes.data, out.nbytes))
        return out

    def copy_from(self, src: "DeviceBuffer", nbytes: int | None = None) -> None:
        """Device-to-device copy of ``nbytes`` (default: all of src)."""
        n = src.nbytes if nbytes is None else nbytes
        if n > min(self.nbytes, src.nbytes):
            raise ValueError("copy larger than a buffer")
        check(_lib.load().pmh_buffer_copy(ctypes.c_void_p(self.ptr), ctypes.c_void_p(src.ptr), n))

    def free(self) -> None:
        if self.ptr:
            check(_lib.load().pmh_buffer_free(ctypes.c_void_p(self.ptr)))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_columns(frame: dict) -> tuple[list, list]:
    """Upload a frame ``name -> (tag, values, valid)`` (Bool values as numpy
    bool) and return ``(columns, buffers)``: the column tuples for
    ``execute_filter_groupby_device`` and the buffers that keep them alive."""
    cols, bufs = [], []
    for name, (tag, values, valid) in frame.items():
        if tag == "Bool":
            data = np.packbits(np.asarray(values, bool), bitorder="little")
        else:
            data = np.ascontiguousarray(values)
        d = DeviceBuffer.from_numpy(data)
        bufs.append(d)
        vp, vl = 0, 0
        if valid is not None:
            vb = np.packbits(np.asarray(valid, bool), bitorder="little")
            v = DeviceBuffer.from_numpy(vb)
            bufs.append(v)
            vp, vl = v.ptr, v.nbytes
        cols.append((name, tag, d.ptr, d.nbytes, vp, vl))
    return cols, bufs
