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
