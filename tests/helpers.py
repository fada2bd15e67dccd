"""Comparison helpers shared by the parity tests."""
import numpy as np

import oracle as orc


def run_gpu(native, plan, frame, entry="fused"):
    wire = orc.to_wire(frame)
    n = orc.frame_rows(frame)
    if entry == "fused":
        out = native.execute_filter_groupby(plan, n, wire)
    else:
        out = native.execute_groupby(plan, n, wire)
    return orc.from_wire(out)


def assert_results_equal(got, want, keys, float_exact=True, rtol=0.0, label="", zero_sign_free=()):
    """Bit-exact comparison after sorting by key.  Floats compare by bits
    when float_exact (the path computes correctly rounded exact sums), else
    within rtol relative to max(|want|, 1e-300).  Outputs named in
    zero_sign_free compare -0.0 and +0.0 as equal: an f64 min/max whose
    extreme is zero carries, in the reference, the sign of the group's first
    zero in row order (crates/polars-metal-kernels/src/groupby.rs:1471-1520),
    while the device tables keep -0.0 below +0.0."""
    assert got["__n_groups__"] == want["__n_groups__"], f"{label}: n_groups {got['__n_groups__']} != {want['__n_groups__']}"
    g = orc.canonical(got, keys)
    w = orc.canonical(want, keys)
    for name, val in w.items():
        if name == "__n_groups__":
            continue
        assert name in g, f"{label}: missing output {name}"
        wt, wv, wok = val
        gt, gv, gok = g[name]
        assert gt == wt, f"{label}: {name} tag {gt} != {wt}"
        np.testing.assert_array_equal(gok, wok, err_msg=f"{label}: {name} validity")
        wv = np.asarray(wv)[wok]
        gv = np.asarray(gv)[gok]
        if name in zero_sign_free and wv.dtype.kind == "f":
            wv = np.where(wv == 0, 0.0, wv).astype(wv.dtype)
            gv = np.where(gv == 0, 0.0, gv).astype(gv.dtype)
        if wv.dtype.kind == "f":
            if float_exact:
                same = (wv.view(np.uint64 if wv.dtype.itemsize == 8 else np.uint32)
                        == gv.view(np.uint64 if gv.dtype.itemsize == 8 else np.uint32)) | (np.isnan(wv) & np.isnan(gv))
                bad = np.flatnonzero(~same)
                assert bad.size == 0, f"{label}: {name} differs at {bad[:5]}: got {gv[bad[:5]]} want {wv[bad[:5]]}"
            else:
                np.testing.assert_allclose(gv, wv, rtol=rtol, atol=0, err_msg=f"{label}: {name}")
        else:
            np.testing.assert_array_equal(gv, wv, err_msg=f"{label}: {name}")
