"""The 100M-row Zipf workload at full size, checked through properties that
do not need the CPU oracle to finish: selected-row and non-null counts,
distinct-key count, global min/max, the total of the exact group sums, and
mean == sum / count per group."""
import numpy as np
import pytest

import oracle as orc
from polars_metal import synth

pytestmark = pytest.mark.gpu


def test_zipf_100m_properties(native):
    from polars_metal.buffer import device_columns
    n = 100_000_000
    print("generating the 100M-row zipf workload (numpy)", flush=True)
    frame, plan = synth.zipf(n, seed=2)
    cols, bufs = device_columns(frame)
    print("running fused filter + group-by on device-resident columns", flush=True)
    res = native.execute_filter_groupby_device(plan, n, cols)
    key = frame["key"][1]
    v, v_ok = frame["v"][1], frame["v"][2]
    w = frame["w"][1]
    sel = (w >= 200) & (w < 500)
    vsel = sel & v_ok
    ng = res.n_rows
    assert res.stats["n_selected"] == int(sel.sum())
    assert ng == len(np.unique(key[sel]))
    out = {}
    for c in res.columns:
        tag, data, valid = res.fetch(c["name"])
        vals = np.frombuffer(data, orc.NP_DTYPE[tag])
        ok = orc.unpack_bits(valid, ng) if valid is not None else np.ones(ng, bool)
        out[c["name"]] = (vals, ok)
    res.free()
    assert len(np.unique(out["key"][0])) == ng
    assert int(out["v_count"][0].sum(dtype=np.int64)) == int(vsel.sum())
    assert out["w_min"][0].min() == w[sel].min() and out["w_max"][0].max() == w[sel].max()
    exact_total = orc.exact_sum(v[vsel])
    group_total = orc.exact_sum(out["v_sum"][0])
    # each group sum is correctly rounded: the total of the rounded sums is
    # within ng half-ulps of the exact total
    assert abs(group_total - exact_total) <= 1e-12 * float(np.abs(v[vsel]).sum())
    cnt = out["v_count"][0]
    has = cnt > 0
    np.testing.assert_array_equal(out["v_mean"][1], has)
    m = out["v_sum"][0][has] / cnt[has].astype(np.float64)
    assert (m.view(np.uint64) == out["v_mean"][0][has].view(np.uint64)).all()
    # spot-check 200 groups against exact per-group sums
    rng = np.random.default_rng(0)
    ks = out["key"][0]
    for i in rng.choice(ng, 200, replace=False):
        rows = vsel & (key == ks[i])
        assert out["v_sum"][0][i] == orc.exact_sum(v[rows]), ks[i]
        assert out["v_count"][0][i] == rows.sum()


def _splitmix(x):
    x = x + np.uint64(0x9E3779B97F4A7C15)
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def _fetch_all(res):
    ng = res.n_rows
    out = {}
    for c in res.columns:
        tag, data, valid = res.fetch(c["name"])
        vals = np.frombuffer(data, orc.NP_DTYPE[tag])[:ng]
        ok = orc.unpack_bits(valid, ng) if valid is not None else np.ones(ng, bool)
        out[c["name"]] = (vals, ok)
    return out


def _group_checksum(out):
    """Order-independent checksum over groups: the wrapping sum of a mix of
    each group's key and aggregates (equal multisets give equal sums)."""
    with np.errstate(over="ignore"):
        a = out["a"][0].astype(np.int64).view(np.uint64)
        b = out["b"][0].view(np.uint64)
        s = out["v_sum"][0].view(np.uint64)
        c = out["v_count"][0].astype(np.uint64)
        x = _splitmix(a * np.uint64(0xD6E8FEB86659FD93) ^ _splitmix(b) ^ (s * np.uint64(0x9E3779B97F4A7C15))
                      ^ (c << np.uint64(40)))
        return int(x.sum(dtype=np.uint64))


@pytest.mark.timeout(600)
def test_highcard_200m_hash_and_sort_agree(native):
    """The 200M-row composite-key workload (~49M groups) through both the
    partitioned hash engine and the sort engine: same group count, same
    multiset of (key, sum, count) rows, totals conserved, and the sort
    engine's output in lane order (b, then a)."""
    from polars_metal.buffer import device_columns
    n = 200_000_000
    print("generating the 200M-row high-cardinality workload (numpy)", flush=True)
    frame, plan = synth.highcard(n, seed=3)
    cols, bufs = device_columns(frame)
    v, v_ok = frame["v"][1], frame["v"][2]
    with np.errstate(over="ignore"):
        total = int(v[v_ok].sum(dtype=np.int64))
    n_valid = int(v_ok.sum())
    del frame
    outs = {}
    for method in ("hash", "sort"):
        print(f"running {method}", flush=True)
        res = native.execute_filter_groupby_device(dict(plan, options={"method": method}), n, cols)
        print(f"{method}: {res.n_rows} groups, {res.stats['device_ms']:.2f} ms", flush=True)
        assert res.stats["method"] == method
        outs[method] = _fetch_all(res)
        res.free()
    h, s = outs["hash"], outs["sort"]
    assert len(h["a"][0]) == len(s["a"][0]) > 45_000_000
    for o in (h, s):
        assert int(o["v_count"][0].sum(dtype=np.int64)) == n_valid
        with np.errstate(over="ignore"):
            assert int(o["v_sum"][0].sum(dtype=np.int64)) == total
    assert _group_checksum(h) == _group_checksum(s)
    # sort output: strictly increasing in (b, a) as unsigned bit patterns
    bs = s["b"][0].view(np.uint64)
    as_ = s["a"][0].view(np.uint32)
    step_b = bs[1:] > bs[:-1]
    same_b = bs[1:] == bs[:-1]
    assert (step_b | (same_b & (as_[1:] > as_[:-1]))).all()
