"""The staging ring and the device column cache (pmh_ingest.hip,
polars_metal.colcache): chunk descriptors on the host; on the GPU, device
assembly of chunked and sliced Arrow columns against host concatenation,
a repeated query that copies no input bytes, stamps, holds and eviction."""
import ctypes

import numpy as np
import pyarrow as pa
import pytest

import oracle as orc
from helpers import assert_results_equal
from polars_metal import _lib, colcache, udf


def test_chunk_descriptors_fold_offsets():
    """Slices keep their parent's buffers with an offset: values by byte,
    bitmaps (and Bool values) by byte plus a bit 0..7."""
    a = pa.array(np.arange(100, dtype=np.int64), mask=np.arange(100) % 7 == 0)
    b = pa.array([True, False, None] * 10)
    col = pa.chunked_array([a.slice(13, 50), a.slice(0, 0), a.slice(70)])
    key, arr, n, keep, nbytes = colcache._chunks("I64", col)
    assert n == 2  # the empty chunk is dropped
    (r0, d0, db0, v0, vb0), (r1, d1, db1, v1, vb1) = key[1]
    assert (r0, r1) == (50, 30)
    assert d0 == a.buffers()[1].address + 13 * 8 and db0 == 0
    assert v0 == a.buffers()[0].address + 1 and vb0 == 5
    assert d1 == a.buffers()[1].address + 70 * 8 and (v1, vb1) == (a.buffers()[0].address + 8, 6)
    assert arr[0].n_rows == 50 and arr[0].valid_bit_offset == 5 and arr[1].data == d1
    assert nbytes == 80 * 8 + (50 + 7) // 8 + (30 + 7) // 8
    key, arr, n, _, _ = colcache._chunks("Bool", b.slice(3, 20))
    ((rows, dp, dbit, vp, vbit),) = key[1]
    assert rows == 20 and dp == b.buffers()[1].address and dbit == 3 and vbit == 3
    # no nulls: no bitmap is passed, even where the array carries one
    c = pa.array([1.0, 2.0, None]).slice(0, 2)
    key, _, _, _, _ = colcache._chunks("F64", c)
    assert key[1][0][3] == 0


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------
def _device_bytes(ptr, n):
    if not n:
        return b""
    out = np.empty(n, np.uint8)
    _lib.check(_lib.load().pmh_buffer_to_host(ctypes.c_void_p(ptr), out.ctypes.data, n))
    return out.tobytes()


def _bits(buf: bytes, n: int) -> np.ndarray:
    return np.unpackbits(np.frombuffer(buf, np.uint8), bitorder="little")[:n].astype(bool)


@pytest.mark.gpu
@pytest.mark.parametrize("tag", ["I64", "I32", "F64", "U8", "Bool"])
def test_device_assembly_matches_host_concatenation(native, tag):
    rng = np.random.default_rng(3)
    n = 70_003
    if tag == "Bool":
        vals = rng.random(n) < 0.4
        base = pa.array(vals, mask=rng.random(n) < 0.1)
    else:
        dt = {"I64": np.int64, "I32": np.int32, "F64": np.float64, "U8": np.uint8}[tag]
        vals = (rng.standard_normal(n) * 1000).astype(dt)
        base = pa.array(vals, mask=rng.random(n) < 0.1)
    # chunk boundaries and slice offsets inside bytes; one chunk without nulls
    cuts = [(0, 5), (5, 13), (21, 40_000), (40_000, 40_001), (40_009, n)]
    chunks = [base.slice(a, b - a) for a, b in cuts]
    chunks.append(pa.array(vals[:1003], type=base.type))  # no validity buffer at all
    col = pa.chunked_array(chunks)
    want = col.combine_chunks()
    m = len(want)
    c, h = colcache.device_column("x", tag, col, transient=True)
    try:
        _, _, dp, dl, vp, vl, loc = c
        assert loc == _lib.PMH_DEVICE and vp != 0 and vl == (m + 7) // 8
        got_v = _bits(_device_bytes(vp, vl + 16), vl * 8 + 128)
        np.testing.assert_array_equal(got_v[:m], ~np.asarray(want.is_null()))
        assert not got_v[m:].any()  # bits past the last row, and the padding, are zero
        if tag == "Bool":
            got = _bits(_device_bytes(dp, dl), m)
            ok = ~np.asarray(want.is_null())
            np.testing.assert_array_equal(got[ok], np.asarray(want.fill_null(False))[ok])
        else:
            got = np.frombuffer(_device_bytes(dp, dl), want.type.to_pandas_dtype())
            ok = ~np.asarray(want.is_null())
            np.testing.assert_array_equal(got[ok], want.fill_null(0).to_numpy()[ok])
    finally:
        colcache.release([h])


def _frame_table(rng, n):
    k = rng.integers(0, 5000, n)
    v = rng.standard_normal(n)
    kv = rng.random(n) < 0.95
    vv = rng.random(n) < 0.9
    frame = {"k": ("I64", k, kv), "v": ("F64", v, vv)}
    table = pa.table({"k": pa.array(k, mask=~kv), "v": pa.array(v, mask=~vv)})
    return frame, table


AGGS = [{"kind": "Simple", "input_col": "v", "op": op, "output_alias": f"v_{op.lower()}"}
        for op in ("Sum", "Count", "Min", "Max")] + [{"kind": "Length", "op": "Len", "output_alias": "n"}]


def _from_table(t):
    out = {"__n_groups__": t.num_rows}
    for name in t.column_names:
        a = t.column(name).combine_chunks()
        tag = udf.arrow_tag(a.type)
        ok = ~np.asarray(a.is_null().to_numpy(zero_copy_only=False))
        out[name] = (tag, a.fill_null(0).to_numpy(zero_copy_only=False).astype(orc.NP_DTYPE[tag]), ok)
    return out


@pytest.mark.gpu
def test_second_query_copies_no_input_bytes(native):
    rng = np.random.default_rng(11)
    n = 2_000_000
    frame, table = _frame_table(rng, n)
    table = pa.concat_tables([table.slice(0, 700_001), table.slice(700_001)])  # two chunks per column
    plan = {"kind": "GroupBy", "input": {"kind": "Scan"}, "keys": [["k", "I64"]], "aggs": AGGS}
    want = orc.groupby({"keys": [["k", "I64"]], "aggs": AGGS}, frame)
    colcache.clear()
    s0 = colcache.stats()
    first = udf.dispatch_groupby(table, plan, cache=True)
    s1 = colcache.stats()
    second = udf.dispatch_groupby(table, plan, cache=True)
    s2 = colcache.stats()
    print(s0, s1, s2, flush=True)
    in_bytes = n * 16 + 2 * ((700_001 + 7) // 8 + (n - 700_001 + 7) // 8)
    up = s1["bytes_uploaded"] - s0["bytes_uploaded"]  # the columns, and the bitmap assembly's chunk table
    assert s1["misses"] - s0["misses"] == 2 and in_bytes <= up <= in_bytes + 1024
    assert s2["h2d_bytes"] == s1["h2d_bytes"]  # nothing crossed PCIe the second time
    assert s2["hits"] - s1["hits"] == 2 and s2["misses"] == s1["misses"]
    for t in (first, second):
        assert_results_equal(_from_table(t), want, ["k"], label="cached")
    # without the cache every query uploads (single chunks by address, through the ring)
    one = pa.table({"k": table.column("k").combine_chunks(), "v": table.column("v").combine_chunks()})
    h0 = colcache.stats()["h2d_bytes"]
    udf.dispatch_groupby(one, plan)
    assert colcache.stats()["h2d_bytes"] - h0 >= n * 16
    colcache.clear()
    assert colcache.stats()["entries"] == 0


@pytest.mark.gpu
def test_stamps_holds_and_eviction(native):
    """A new stamp re-uploads the same address; held entries survive a
    capacity cut, unheld ones go least recently used first."""
    lib = _lib.load()
    rng = np.random.default_rng(2)
    xs = [np.ascontiguousarray(rng.integers(0, 1 << 40, 100_000)) for _ in range(3)]

    def resident(x, stamp):
        ch = (_lib.PmhChunk * 1)()
        ch[0].n_rows, ch[0].data = len(x), x.ctypes.data
        out = _lib.PmhResidentColumn()
        _lib.check(lib.pmh_column_resident(0, ch, 1, stamp, ctypes.byref(out)))  # tag 0: I64
        return out

    colcache.clear()
    try:
        a = resident(xs[0], 7)
        assert a.hit == 0
        b = resident(xs[0], 7)
        assert b.hit == 1 and b.entry == a.entry and b.data == a.data
        c = resident(xs[0], 8)  # same bytes, new stamp: a new entry
        assert c.hit == 0 and c.entry != a.entry
        for e in (a.entry, a.entry, c.entry):
            _lib.check(lib.pmh_column_release(e))
        with pytest.raises(ValueError):
            _lib.check(lib.pmh_column_release(a.entry))  # not held any more
        d = resident(xs[1], 7)  # held
        e = resident(xs[2], 7)
        _lib.check(lib.pmh_column_release(e.entry))
        entry_bytes = 100_000 * 8 + 16
        _lib.check(lib.pmh_colcache_set_capacity(entry_bytes))  # room for one: only the held one stays
        st = colcache.stats()
        assert st["entries"] == 1 and st["bytes_resident"] == entry_bytes
        assert np.array_equal(np.frombuffer(_device_bytes(d.data, d.data_len), np.int64), xs[1])
        f = resident(xs[0], 7)  # evicted earlier: uploaded again
        assert f.hit == 0
        _lib.check(lib.pmh_column_release(f.entry))
        _lib.check(lib.pmh_column_release(d.entry))
        with pytest.raises(ValueError):
            _lib.check(lib.pmh_column_resident(0, None, 1, 0, ctypes.byref(_lib.PmhResidentColumn())))
    finally:
        _lib.check(lib.pmh_colcache_set_capacity(1 << 36))
        colcache.clear()
