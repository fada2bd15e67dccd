"""The allocation-free fused group-by (pmh_execute_filter_groupby_noalloc)
against the allocating entry point it mirrors (pmh_execute_filter_groupby
with device outputs) and the oracle, bit for bit, on two calls in a row,
the second with a smaller input; then its refusals.  Runs alone in a fresh
process and after every other test."""
import numpy as np
import pytest

import oracle as orc
from helpers import assert_results_equal
from polars_metal import synth

pytestmark = pytest.mark.gpu

MAX_ROWS = 262144


def _out_buffers(plan, frame, capacity):
    """Caller-owned device outputs for every plan output (keys, then aggs)."""
    from polars_metal.buffer import DeviceBuffer
    tags = [k[1] for k in plan["keys"]]
    for a in plan["aggs"]:
        if a.get("op") in ("Len", "Count") or a.get("kind") == "Length":
            tags.append("U32")
        else:
            t = frame[a["input_col"]][0]
            tags.append(("F32" if t == "F32" else "F64") if a["op"] == "Mean" else t)
    vbytes = max(4, (((capacity + 7) // 8) + 3) & ~3)
    bufs, outs = [], []
    for t in tags:
        nbytes = vbytes if t == "Bool" else capacity * {"I64": 8, "F64": 8, "U64": 8, "I32": 4, "F32": 4, "U32": 4,
                                                          "I16": 2, "U16": 2, "I8": 1, "U8": 1}[t]
        d = DeviceBuffer.alloc(max(nbytes, 1))
        v = DeviceBuffer.alloc(vbytes)
        bufs += [d, v]
        outs.append((d, v))
    return outs, bufs


def _fetch(names, outs, meta, ng):
    res = {"__n_groups__": ng}
    for name, (d, v), (tag, dlen, vlen) in zip(names, outs, meta):
        assert vlen == max(4, (((ng + 7) // 8) + 3) & ~3)
        raw = d.to_numpy(np.uint8, dlen) if dlen else np.zeros(0, np.uint8)
        vals = orc.unpack_bits(raw, ng) if tag == "Bool" else raw.view(orc.NP_DTYPE[tag])[:ng]
        vb = v.to_numpy(np.uint8, vlen)
        # padding past the last group is zero, as in the allocating entry
        assert not orc.unpack_bits(vb, vlen * 8)[ng:].any()
        res[name] = (tag, vals, orc.unpack_bits(vb, ng))
    return res


def _fetch_device(res):
    out = {"__n_groups__": res.n_rows}
    for c in res.columns:
        tag, data, valid = res.fetch(c["name"])
        vals = orc.unpack_bits(data, res.n_rows) if tag == "Bool" else np.frombuffer(data, orc.NP_DTYPE[tag])[:res.n_rows]
        ok = orc.unpack_bits(valid, res.n_rows) if valid is not None else np.ones(res.n_rows, bool)
        out[c["name"]] = (tag, vals, ok)
    return out


def _names(plan):
    return [k[0] for k in plan["keys"]] + [a["output_alias"] for a in plan["aggs"]]


def _check_pair(native, plan, frame, rows, label):
    from polars_metal.buffer import device_columns
    cols, keep = device_columns(frame)
    outs, bufs = _out_buffers(plan, frame, rows)
    print(f"noalloc {label}: {rows} rows, {plan.get('options')}", flush=True)
    ng, meta = native.execute_filter_groupby_noalloc(
        plan, rows, cols, [(d.ptr, d.nbytes, v.ptr, v.nbytes) for d, v in outs])
    got = _fetch(_names(plan), outs, meta, ng)
    ref = native.execute_filter_groupby_device(plan, rows, cols)
    want_dev = _fetch_device(ref)
    ref.free()
    want = orc.groupby(plan, frame)
    assert ng == want["__n_groups__"]
    want_dev["__ops__"] = want["__ops__"]
    keys = [k[0] for k in plan["keys"]]
    assert_results_equal(got, want_dev, keys, label=f"{label}/vs-entry")
    assert_results_equal(got, want, keys, label=f"{label}/vs-oracle")
    del keep, bufs


@pytest.mark.parametrize("method", ["auto", "hash_global", "sort"])
def test_noalloc_matches_allocating_entry_twice(native, method):
    # the largest input the scratch array takes, then a smaller one: the
    # second call must not see the first's leftovers
    frame, plan = synth.zipf(MAX_ROWS, seed=41)
    plan = dict(plan, options={"method": method})
    _check_pair(native, plan, frame, MAX_ROWS, f"{method}/first")
    frame2, plan2 = synth.zipf(70_001, seed=42)
    plan2 = dict(plan2, options={"method": method})
    _check_pair(native, plan2, frame2, 70_001, f"{method}/second")


def test_noalloc_tail_partitions_and_bool_keys(native):
    # every group through the tail partitions and the fallback table, and a
    # Bool key whose data comes back as bits
    rng = np.random.default_rng(43)
    n = 150_000
    frame = {"k": ("I32", rng.integers(0, 40_000, n).astype(np.int32), rng.random(n) < 0.97),
             "f": ("Bool", rng.random(n) < 0.5, rng.random(n) < 0.9),
             "v": ("F64", rng.standard_normal(n), rng.random(n) < 0.9)}
    aggs = [{"kind": "Simple", "input_col": "v", "op": op, "output_alias": op} for op in ("Sum", "Min", "Max", "Mean")]
    plan = {"keys": [["k", "I32"], ["f", "Bool"]], "aggs": aggs,
            "options": {"method": "hash", "lds_slots": 0, "partitions": 4, "tail_slots": 256}}
    _check_pair(native, plan, frame, n, "tail")
    plan = {"keys": [["k", "I32"], ["f", "Bool"]], "aggs": aggs}
    small = {k: (t, v[:10_000], None if ok is None else ok[:10_000]) for k, (t, v, ok) in frame.items()}
    _check_pair(native, plan, small, 10_000, "tail-small")


def test_noalloc_refusals(native):
    from polars_metal.buffer import device_columns
    frame, plan = synth.zipf(20_000, seed=44)
    cols, keep = device_columns(frame)
    want = orc.groupby(plan, frame)["__n_groups__"]
    # outputs too small: refused with the group count, nothing written
    outs, bufs = _out_buffers(plan, frame, want - 1)
    with pytest.raises(ValueError, match="holds fewer than") as ei:
        native.execute_filter_groupby_noalloc(plan, 20_000, cols, [(d.ptr, d.nbytes, v.ptr, v.nbytes) for d, v in outs])
    assert ei.value.n_groups == want
    outs, bufs = _out_buffers(plan, frame, want)
    args = [(d.ptr, d.nbytes, v.ptr, v.nbytes) for d, v in outs]
    # more rows than the scratch array is sized for, before any device work
    with pytest.raises(ValueError, match="at most 262144 rows"):
        native.execute_filter_groupby_noalloc(plan, MAX_ROWS + 1, cols, args)
    # host columns
    host = [(c[0], c[1], 0x1000, c[3], 0, 0) for c in cols]
    with pytest.raises(ValueError):
        from polars_metal import _lib
        import ctypes
        arr = (_lib.PmhColumn * len(host))()
        for i, (name, tag, dp, dl, vp, vl) in enumerate(host):
            arr[i].name, arr[i].tag, arr[i].location = name.encode(), _lib.TAGS[tag], _lib.PMH_HOST
            arr[i].data, arr[i].data_len = dp, dl
        o = (_lib.PmhOutColumn * len(args))()
        ng = ctypes.c_uint64()
        _lib.check(_lib.load().pmh_execute_filter_groupby_noalloc(
            native._plan_json(plan), 20_000, arr, len(host), o, len(args), ctypes.byref(ng)))
    # a wrong number of outputs
    with pytest.raises(ValueError, match="output columns"):
        native.execute_filter_groupby_noalloc(plan, 20_000, cols, args[:-1])
    # the entry still works after the refusals
    ng, meta = native.execute_filter_groupby_noalloc(plan, 20_000, cols, args)
    assert ng == want
    del keep, bufs
