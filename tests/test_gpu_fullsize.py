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
    import torch
    n = 100_000_000
    print("generating 100M-row zipf workload on the device", flush=True)
    d = synth.zipf_device(n, seed=2)
    plan = synth.zipf_plan()
    cols = [("key", "I64", d["key"].data_ptr(), n * 8, 0, 0),
            ("v", "F64", d["v"].data_ptr(), n * 8, d["v_valid"].data_ptr(), d["v_valid"].numel()),
            ("w", "I32", d["w"].data_ptr(), n * 4, 0, 0)]
    torch.cuda.synchronize()
    print("running fused filter + group-by", flush=True)
    res = native.execute_filter_groupby_device(plan, n, cols)
    ng = res.n_rows
    sel = (d["w"] >= 200) & (d["w"] < 500)
    vsel = sel & d["valid_bool"]
    assert res.stats["n_selected"] == int(sel.sum())
    assert ng == int(torch.unique(d["key"][sel]).numel())
    out = {}
    for c in res.columns:
        tag, data, valid = res.fetch(c["name"])
        vals = np.frombuffer(data, orc.NP_DTYPE[tag])
        ok = orc.unpack_bits(valid, ng) if valid is not None else np.ones(ng, bool)
        out[c["name"]] = (vals, ok)
    res.free()
    assert len(out["key"][0]) == ng and len(np.unique(out["key"][0])) == ng
    assert int(out["v_count"][0].sum(dtype=np.int64)) == int(vsel.sum())
    w_sel = d["w"][sel]
    assert out["w_min"][0].min() == int(w_sel.min()) and out["w_max"][0].max() == int(w_sel.max())
    total = float(d["v"][vsel].sum())
    exact_groups = out["v_sum"][0]
    assert abs(orc.exact_sum(exact_groups) - total) <= 1e-9 * float(d["v"][vsel].abs().sum())
    cnt = out["v_count"][0]
    has = cnt > 0
    np.testing.assert_array_equal(out["v_mean"][1], has)
    m = exact_groups[has] / cnt[has].astype(np.float64)
    assert (m.view(np.uint64) == out["v_mean"][0][has].view(np.uint64)).all()
