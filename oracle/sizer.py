"""CPU restatement (float64) of the learned sizer's inputs -- test
infrastructure only: the tests compare what the partitioned engine's sketch
block computes on the device (polars-metal_amd/csrc/pmh_table.h, the
estimators and features, evaluated in pmh_hashagg.hip's hs_sample_count)
with it, and the product never imports it.

No counterpart in the reference, which sizes nothing from data.  The
statistics are textbook distinct-value estimators over a row sample of s
selected rows out of an estimated n_sel, with f_c the sampled keys seen c
times and d the sampled distinct keys:

    GEE       sqrt(n_sel / s) f1 + (d - f1)               (Charikar et al., 2000)
    Shlosser  d + f1 sum_c (1-q)^c f_c / sum_c c q (1-q)^(c-1) f_c,  q = s / n_sel
                                                          (Haas et al., 1995)
    Chao1     d + f1 (f1 - 1) / (2 (f2 + 1))             (bias-corrected, Chao 1984)

and the plain estimate the partitions are sized from combines them: linear
extrapolation n_sel d / s for a sample of singletons (or one whose
fingerprint table dropped most of it), d when the sample is the whole
selection, else max(d, sqrt(GEE * Shlosser), min(n_sel, Chao1)).  Rows the
sample table dropped count as singletons.

Parity pins (tests/test_oracle.py): each estimator against a hand-worked
value on a small histogram, Shlosser against d at full coverage (q = 1),
the plain estimate's three regimes, and the features' closed forms
(lane-width one-hot, log scale)."""
from __future__ import annotations

import math

import numpy as np

N_FEATURES = 16


def sketch(hist, s_sel: int, n_sample: int, n_rows: int, d: int, f1: int, f2: int, dropped: int) -> dict:
    """The sketch fields from the sample counts; hist[c] = keys seen c times."""
    k = {"s": s_sel, "sample_size": n_sample, "n_rows": n_rows, "d": d + dropped, "f1": f1 + dropped,
         "f2": f2, "dropped": dropped}
    n_sel = n_rows * s_sel / n_sample if n_sample else 0.0
    k["n_sel"] = n_sel
    q = min(1.0, s_sel / n_sel) if n_sel > 0 else 1.0
    F2 = num = den = 0.0
    for c in range(1, min(len(hist), 1024)):
        h = float(hist[c]) + (dropped if c == 1 else 0.0)
        if h == 0:
            continue
        F2 += c * c * h
        num += (1.0 - q) ** c * h
        den += c * q * (1.0 - q) ** (c - 1) * h
    k.update(F2=F2, shl_num=num, shl_den=den)
    return k


def gee(k: dict) -> float:
    if k["s"] == 0:
        return 0.0
    return math.sqrt(max(k["n_sel"], k["s"]) / k["s"]) * k["f1"] + (k["d"] - k["f1"])


def shlosser(k: dict) -> float:
    if k["s"] == 0:
        return 0.0
    return k["d"] + k["f1"] * k["shl_num"] / k["shl_den"] if k["shl_den"] > 0 else gee(k)


def chao1(k: dict) -> float:
    f1 = k["f1"]
    return min(k["n_sel"], k["d"] + f1 * (f1 - 1.0 if f1 > 0 else 0.0) / (2.0 * (k["f2"] + 1.0)))


def plain_estimate(k: dict) -> float:
    s, d = k["s"], k["d"]
    if s == 0:
        return 0.0
    if 2.0 * k["dropped"] > s or (k["f2"] == 0 and k["f1"] >= 0.95 * s):
        return max(d, k["n_sel"] * d / s)
    if k["n_sel"] <= s * 1.0000001:
        return float(d)
    return max(d, math.sqrt(max(gee(k), 1.0) * max(shlosser(k), 1.0)), chao1(k))


def _lg(x: float) -> float:
    return math.log2(max(0.0, x) + 1.0) / 32.0


def features(k: dict, null_keys: int, lane_bits: int) -> np.ndarray:
    s = max(1.0, float(k["s"]))
    f = np.zeros(N_FEATURES, np.float64)
    f[0] = _lg(k["n_rows"])
    f[1] = k["s"] / k["sample_size"] if k["sample_size"] else 0.0
    f[2] = _lg(k["d"])
    f[3] = k["f1"] / s
    f[4] = k["f2"] / s
    f[5] = k["d"] / s
    f[6] = _lg(gee(k))
    f[7] = _lg(shlosser(k))
    f[8] = _lg(plain_estimate(k))
    f[9] = k["F2"] / (s * s)
    f[10] = null_keys / s
    f[11] = k["dropped"] / s
    f[12] = 1.0 if lane_bits <= 33 else 0.0
    f[13] = 1.0 if 33 < lane_bits <= 65 else 0.0
    f[14] = 1.0 if lane_bits > 65 else 0.0
    f[15] = _lg(k["n_sel"])
    return f
