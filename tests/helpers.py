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


def _bits(a):
    return a.view(np.uint64 if a.dtype.itemsize == 8 else np.uint32)


def assert_results_equal(got, want, keys, float_exact=True, rtol=0.0, label=""):
    """Comparison after sorting by key.  Floats compare by bits when
    float_exact (the path computes correctly rounded exact sums), else
    within rtol relative.  Min/Max outputs always compare by bits, NaN
    included: a group holding a NaN yields the canonical quiet NaN and -0.0
    orders below +0.0.  A NaN Sum or Mean matches any NaN, since IEEE 754
    leaves the payload of a computed NaN (inf + -inf) to the hardware.
    ``want['__ops__']`` (alias -> op) says which output is which."""
    assert got["__n_groups__"] == want["__n_groups__"], f"{label}: n_groups {got['__n_groups__']} != {want['__n_groups__']}"
    ops = want.get("__ops__", {})
    g = orc.canonical(got, keys)
    w = orc.canonical(want, keys)
    for name, val in w.items():
        if name.startswith("__"):
            continue
        assert name in g, f"{label}: missing output {name}"
        wt, wv, wok = val
        gt, gv, gok = g[name]
        assert gt == wt, f"{label}: {name} tag {gt} != {wt}"
        np.testing.assert_array_equal(gok, wok, err_msg=f"{label}: {name} validity")
        wv = np.asarray(wv)[wok]
        gv = np.asarray(gv)[gok]
        if wv.dtype.kind == "f":
            by_bits = ops.get(name) in ("Min", "Max")
            if float_exact or by_bits:
                same = _bits(wv) == _bits(gv)
                if not by_bits:
                    same |= np.isnan(wv) & np.isnan(gv)
                bad = np.flatnonzero(~same)
                assert bad.size == 0, (f"{label}: {name} differs at {bad[:5]}: got {gv[bad[:5]]} "
                                       f"({[hex(x) for x in _bits(gv[bad[:5]])]}) want {wv[bad[:5]]} "
                                       f"({[hex(x) for x in _bits(wv[bad[:5]])]})")
            else:
                np.testing.assert_allclose(gv, wv, rtol=rtol, atol=0, err_msg=f"{label}: {name}")
        else:
            np.testing.assert_array_equal(gv, wv, err_msg=f"{label}: {name}")
