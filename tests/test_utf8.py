"""Utf8 keys on the device (pmh_utf8_encode / pmh_utf8_decode): codes equal
the oracle's restatement of build_dict_nullable (first seen wins, null rows
code 0, "" a value of its own), decoding returns the reference's packed wire
form, rows with bad offsets or bytes that are not UTF-8 are refused by row,
and Utf8-keyed group-bys through the bytes API and the Arrow dispatch equal
the oracle."""
import ctypes

import numpy as np
import pyarrow as pa
import pytest

import oracle as orc
from helpers import assert_results_equal, run_gpu


def _strings(rng, n, pool=None):
    words = pool if pool is not None else np.array(
        ["", "a", "ab", "abc", "Zürich", "日本語", "x" * 8, "x" * 9, "y" * 63, "long " * 40, "ab ", "été",
         "emoji \U0001F600"], object)
    return words[rng.integers(0, len(words), n)]


def _codes_of(d):
    from polars_metal import _lib
    out = np.empty(max(1, d.n_rows), np.uint32)
    if d.n_rows:
        _lib.check(_lib.load().pmh_buffer_to_host(ctypes.c_void_p(d.codes), out.ctypes.data, d.n_rows * 4))
    return out[:d.n_rows]


def _valid_of(d):
    from polars_metal import _lib
    if not d.valid:
        return np.ones(d.n_rows, bool)
    nb = (d.n_rows + 7) // 8
    out = np.empty(max(1, nb), np.uint8)
    _lib.check(_lib.load().pmh_buffer_to_host(ctypes.c_void_p(d.valid), out.ctypes.data, nb))
    return np.unpackbits(out, bitorder="little")[:d.n_rows].astype(bool)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["mixed", "high_card", "sliced", "large_string", "all_null", "empty", "one"])
def test_device_codes_equal_first_seen_dictionary(native, case):
    from polars_metal._native import Utf8Dict
    rng = np.random.default_rng(["mixed", "high_card", "sliced", "large_string", "all_null", "empty", "one"].index(case))
    n = {"mixed": 50_000, "high_card": 300_000, "sliced": 40_000, "large_string": 20_000, "all_null": 1000,
         "empty": 0, "one": 1}[case]
    if case == "high_card":
        vals = np.array([f"k{v}" for v in rng.integers(0, 200_000, n)], object)
    else:
        vals = _strings(rng, n)
    ok = rng.random(n) < 0.9 if case != "all_null" else np.zeros(n, bool)
    arr = pa.array(list(vals), pa.large_string() if case == "large_string" else pa.string(), mask=~ok)
    if case == "sliced":  # a slice: nonzero first offset, validity starting inside a byte
        arr, vals, ok = arr.slice(13, n - 20), vals[13:n - 7], ok[13:n - 7]
    d = Utf8Dict.from_arrow(arr)
    try:
        want_dict, want_codes, want_ok = orc.utf8_dict(vals, ok)
        assert d.n_rows == len(vals) and d.n_distinct == len(want_dict)
        np.testing.assert_array_equal(_valid_of(d), want_ok)
        np.testing.assert_array_equal(_codes_of(d), want_codes)
        # every row decoded back: the packed form of the input (nulls empty)
        codes = want_codes.tobytes()
        vb = orc.pack_bits(want_ok).tobytes()
        assert d.decode(codes, vb, len(vals)) == orc.utf8_packed(vals, want_ok)
        # the distinct strings, in code order
        m = len(want_dict)
        got = orc.utf8_unpacked(d.decode(np.arange(m, dtype=np.uint32).tobytes(), None, m))
        assert got.tolist() == want_dict
    finally:
        d.free()


@pytest.mark.gpu
def test_bad_rows_are_refused_by_row(native):
    from polars_metal._native import Utf8Dict
    good = ["ok", "fine", "", "fine"]
    # bytes that are not well-formed UTF-8: a lone continuation byte, an
    # overlong form, a surrogate, a cut-off sequence, beyond U+10FFFF
    for bad in (b"\x80", b"\xc0\xaf", b"\xed\xa0\x80", b"ab\xe6\x97", b"\xf4\x90\x80\x80"):
        blobs = [s.encode() for s in good] + [bad, b"ok"]
        offs = np.zeros(len(blobs) + 1, np.int32)
        np.cumsum([len(b) for b in blobs], out=offs[1:])
        wire = np.uint32(len(blobs)).tobytes() + offs.tobytes() + b"".join(blobs)
        with pytest.raises(ValueError, match="row 4 is not valid UTF-8"):
            Utf8Dict.from_packed(wire, None, len(blobs))
        # the same bytes in a null row are not read
        vb = orc.pack_bits(np.array([1, 1, 1, 1, 0, 1], bool)).tobytes()
        Utf8Dict.from_packed(wire, vb, len(blobs)).free()
    # offsets that decrease inside the buffer
    offs = np.array([0, 2, 1, 3], np.int32)
    wire = np.uint32(3).tobytes() + offs.tobytes() + b"abc"
    with pytest.raises(ValueError, match="row 1 has invalid offsets"):
        Utf8Dict.from_packed(wire, None, 3)
    # the last offset past the buffer
    wire = np.uint32(1).tobytes() + np.array([0, 9], np.int32).tobytes() + b"abc"
    with pytest.raises(ValueError, match="do not lie in"):
        Utf8Dict.from_packed(wire, None, 1)
    # header disagreeing with the row count (the reference's wire check)
    with pytest.raises(ValueError, match="disagrees"):
        Utf8Dict.from_packed(wire, None, 2)


AGGS = [{"kind": "Simple", "input_col": "v", "op": op, "output_alias": f"v_{op.lower()}"}
        for op in ("Sum", "Count", "Min", "Max", "Mean")] + [{"kind": "Length", "op": "Len", "output_alias": "n"}]


def _frame(rng, n):
    s = _strings(rng, n)
    sok = rng.random(n) < 0.92
    v = rng.standard_normal(n)
    return {"s": ("Utf8", s, sok), "k": ("I64", rng.integers(0, 4, n), None), "v": ("F64", v, rng.random(n) < 0.95)}


@pytest.mark.gpu
@pytest.mark.parametrize("keys", [["s"], ["s", "k"]])
def test_utf8_groupby_bytes_api_matches_oracle(native, keys):
    rng = np.random.default_rng(len(keys))
    frame = _frame(rng, 200_000)
    plan = {"keys": [[k, frame[k][0]] for k in keys], "aggs": AGGS}
    want = orc.groupby(plan, frame)
    for entry in ("fused", "groupby"):
        got = run_gpu(native, plan, frame, entry=entry)
        assert_results_equal(got, want, keys, label=f"utf8-{entry}")


@pytest.mark.gpu
def test_utf8_groupby_arrow_dispatch_matches_oracle(native):
    from polars_metal import udf
    rng = np.random.default_rng(9)
    n = 150_000
    frame = _frame(rng, n)
    s, sok = frame["s"][1], frame["s"][2]
    v, vok = frame["v"][1], frame["v"][2]
    table = pa.table({"s": pa.array(list(s), pa.string(), mask=~sok), "k": pa.array(frame["k"][1]),
                      "v": pa.array(v, mask=~vok)})
    plan = {"kind": "GroupBy", "input": {"kind": "Scan"}, "keys": [["s", "Utf8"], ["k", "I64"]], "aggs": AGGS}
    t = udf.dispatch_groupby(table, plan)
    got = {"__n_groups__": t.num_rows}
    for name in t.column_names:
        a = t.column(name).combine_chunks()
        ok = ~np.asarray(a.is_null())
        if pa.types.is_string(a.type):
            got[name] = ("Utf8", np.array([x if x is not None else "" for x in a.to_pylist()], object), ok)
        else:
            tag = udf.arrow_tag(a.type) if a.type != pa.uint32() else "U32"
            got[name] = (tag, a.fill_null(0).to_numpy(zero_copy_only=False).astype(orc.NP_DTYPE[tag]), ok)
    want = orc.groupby({"keys": [["s", "Utf8"], ["k", "I64"]], "aggs": AGGS}, frame)
    assert_results_equal(got, want, ["s", "k"], label="utf8-arrow")
