"""Seeded synthetic workloads for the filtered group-by hot path.

Each generator returns a frame ``name -> (tag, values, valid)`` (numpy;
``valid`` None for no nulls) plus the query plan in the reference's wire
form (group-by plan with an optional PredicateAst ``predicate``).

  small     1M rows; key i64 U{0..999}; v f64 U[0,1); w i32 U{0..99};
            filter v > 0.5; sum(v), count(v), min(w), max(w), mean(v)
  zipf      100M rows; key i64 bounded Zipf(s=1.1) over ranks 1..1e6
            (inverse-CDF sampling); v f64 N(0,1) with 5% nulls; w i32
            U{0..999}; filter (w >= 200) & (w < 500); same five aggregates
  highcard  200M rows; composite key (a i32 U[0,2^20), b i64 U[0,2^40))
            drawn from a pool of 50M distinct pairs; v i64 U[-1e9,1e9] with
            10% nulls; group_by([a, b]).agg(sum(v), count(v))
  batch     4096 queries, rows log-uniform in [1e3, 1e6]; key i64
            U{1..rows/10}; v f64 U[0,1); 20% of queries filter v > 0.5

``rows`` scales any workload down for parity tests; the distributions stay
the same.  ``*_device`` variants draw the same distributions with torch on
the GPU (used where host generation plus upload would dominate).
"""
from __future__ import annotations

import numpy as np

FIVE_AGGS = [
    {"kind": "Simple", "input_col": "v", "op": "Sum", "output_alias": "v_sum"},
    {"kind": "Simple", "input_col": "v", "op": "Count", "output_alias": "v_count"},
    {"kind": "Simple", "input_col": "w", "op": "Min", "output_alias": "w_min"},
    {"kind": "Simple", "input_col": "w", "op": "Max", "output_alias": "w_max"},
    {"kind": "Simple", "input_col": "v", "op": "Mean", "output_alias": "v_mean"},
]


def _col(name: str) -> dict:
    return {"kind": "Column", "name": name, "dtype": "I64"}


def pred_gt_f64(col: str, x: float) -> dict:
    return {"kind": "Compare", "op": "Gt", "dtype": "F64",
            "lhs": {"kind": "Column", "name": col, "dtype": "F64"},
            "rhs": {"kind": "LiteralF64", "value": float(x)}}


def pred_range_i64(col: str, lo: int, hi: int) -> dict:
    """(col >= lo) & (col < hi): the spelling the reference's walker accepts
    for a half-open range (python/polars_metal/_walker.py:1508-1620)."""
    return {"kind": "And",
            "lhs": {"kind": "Compare", "op": "Ge", "dtype": "I64", "lhs": _col(col),
                    "rhs": {"kind": "LiteralI64", "value": int(lo)}},
            "rhs": {"kind": "Compare", "op": "Lt", "dtype": "I64", "lhs": _col(col),
                    "rhs": {"kind": "LiteralI64", "value": int(hi)}}}


def small_plan() -> dict:
    return {"keys": [["key", "I64"]], "aggs": FIVE_AGGS, "predicate": pred_gt_f64("v", 0.5)}


def zipf_plan() -> dict:
    return {"keys": [["key", "I64"]], "aggs": FIVE_AGGS, "predicate": pred_range_i64("w", 200, 500)}


def highcard_plan() -> dict:
    return {"keys": [["a", "I32"], ["b", "I64"]],
            "aggs": [{"kind": "Simple", "input_col": "v", "op": "Sum", "output_alias": "v_sum"},
                     {"kind": "Simple", "input_col": "v", "op": "Count", "output_alias": "v_count"}]}


def small(rows: int = 1_000_000, seed: int = 1) -> tuple[dict, dict]:
    rng = np.random.default_rng(seed)
    frame = {
        "key": ("I64", rng.integers(0, 1000, rows, dtype=np.int64), None),
        "v": ("F64", rng.random(rows), None),
        "w": ("I32", rng.integers(0, 100, rows, dtype=np.int32), None),
    }
    return frame, small_plan()


def zipf_cdf(n_ranks: int = 1_000_000, s: float = 1.1) -> np.ndarray:
    w = np.arange(1, n_ranks + 1, dtype=np.float64) ** (-s)
    c = np.cumsum(w)
    return c / c[-1]


def zipf(rows: int = 100_000_000, seed: int = 2, n_ranks: int = 1_000_000, s: float = 1.1) -> tuple[dict, dict]:
    rng = np.random.default_rng(seed)
    cdf = zipf_cdf(n_ranks, s)
    key = np.searchsorted(cdf, rng.random(rows), side="right").astype(np.int64) + 1
    key = np.minimum(key, n_ranks)
    v = rng.standard_normal(rows)
    v_valid = rng.random(rows) >= 0.05
    w = rng.integers(0, 1000, rows, dtype=np.int32)
    frame = {"key": ("I64", key, None), "v": ("F64", v, v_valid), "w": ("I32", w, None)}
    return frame, zipf_plan()


def highcard(rows: int = 200_000_000, seed: int = 3, pool: int = 50_000_000) -> tuple[dict, dict]:
    rng = np.random.default_rng(seed)
    pool = min(pool, max(1, rows // 4))
    pa = rng.integers(0, 1 << 20, pool, dtype=np.int32)
    pb = rng.integers(0, 1 << 40, pool, dtype=np.int64)
    idx = rng.integers(0, pool, rows)
    v = rng.integers(-1_000_000_000, 1_000_000_001, rows, dtype=np.int64)
    frame = {"a": ("I32", pa[idx], None), "b": ("I64", pb[idx], None),
             "v": ("I64", v, rng.random(rows) >= 0.10)}
    return frame, highcard_plan()


def batch_specs(n_queries: int = 4096, seed: int = 5, lo: float = 1e3, hi: float = 1e6) -> list[dict]:
    rng = np.random.default_rng(seed)
    rows = np.exp(rng.uniform(np.log(lo), np.log(hi), n_queries)).astype(np.int64)
    filt = rng.random(n_queries) < 0.2
    return [{"rows": int(r), "distinct": max(1, int(r) // 10), "filter": bool(f), "seed": int(seed * 100003 + i)}
            for i, (r, f) in enumerate(zip(rows, filt))]


_BATCH_AGGS = [{"kind": "Simple", "input_col": "v", "op": op, "output_alias": f"v_{op.lower()}"}
               for op in ("Sum", "Count", "Min", "Max", "Mean")]
_BATCH_PRED = pred_gt_f64("v", 0.5)  # one object shared by the filtered queries


def batch_query(spec: dict) -> tuple[dict, dict]:
    rng = np.random.default_rng(spec["seed"])
    r = spec["rows"]
    frame = {"key": ("I64", rng.integers(1, spec["distinct"] + 1, r, dtype=np.int64), None),
             "v": ("F64", rng.random(r), None)}
    plan = {"keys": [["key", "I64"]], "aggs": _BATCH_AGGS}
    if spec["filter"]:
        plan["predicate"] = _BATCH_PRED
    return frame, plan


# ---------------------------------------------------------------------------
# device-side generation (torch on the GPU): same distributions
# ---------------------------------------------------------------------------
def zipf_device(rows: int, seed: int = 2, n_ranks: int = 1_000_000, s: float = 1.1, device: str = "cuda"):
    """Returns dict of torch tensors: key i64, v f64, v_valid packed u8
    bitmap, w i32."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    cdf = torch.from_numpy(zipf_cdf(n_ranks, s)).to(device)
    u = torch.rand(rows, generator=g, device=device, dtype=torch.float64)
    key = torch.searchsorted(cdf, u, right=True).to(torch.int64) + 1
    key.clamp_(max=n_ranks)
    v = torch.randn(rows, generator=g, device=device, dtype=torch.float64)
    valid = torch.rand(rows, generator=g, device=device, dtype=torch.float32) >= 0.05
    w = torch.randint(0, 1000, (rows,), generator=g, device=device, dtype=torch.int32)
    return {"key": key, "v": v, "v_valid": pack_bits_device(valid), "w": w, "valid_bool": valid}


def pack_bits_device(b):
    """bool tensor (n,) -> LSB-first packed uint8 tensor of ceil(n/8) bytes."""
    import torch
    n = b.numel()
    pad = (-n) % 8
    bb = torch.nn.functional.pad(b.to(torch.uint8), (0, pad)).view(-1, 8)
    weights = torch.tensor([1, 2, 4, 8, 16, 32, 64, 128], dtype=torch.uint8, device=b.device)
    return (bb * weights).sum(dim=1, dtype=torch.uint8)


# ---------------------------------------------------------------------------
# the learned sizer's job stream: rows and distinct keys log-uniform, key
# Zipf exponent in [0, 1.5], key dtype i32 / i64 / u128 (an (i64, i32)
# composite: a 98-bit lane), null fraction in [0, 0.2]
# ---------------------------------------------------------------------------
def sizer_job_specs(n_jobs: int, seed: int, rows_lo: float = 1e3, rows_hi: float = 1e7) -> list[dict]:
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n_jobs):
        rows = int(np.exp(rng.uniform(np.log(rows_lo), np.log(rows_hi))))
        distinct = max(1, int(np.exp(rng.uniform(0.0, np.log(rows)))))
        specs.append({"rows": rows, "distinct": distinct, "s": float(rng.uniform(0.0, 1.5)),
                      "dtype": ["I32", "I64", "U128"][int(rng.integers(0, 3))],
                      "null_frac": float(rng.uniform(0.0, 0.2)), "seed": int(seed * 1_000_003 + i)})
    return specs


def _power_law_ranks(rng, n: int, d: int, s: float) -> np.ndarray:
    """n ranks in [0, d) with P(rank k) ~ (k+1)^-s: the inverse CDF of the
    continuous bounded power law on [1, d + 1), floored."""
    u = rng.random(n)
    if abs(s - 1.0) < 1e-9:
        x = np.power(float(d + 1), u)
    else:
        a = 1.0 - s
        x = np.power(1.0 + u * (np.power(float(d + 1), a) - 1.0), 1.0 / a)
    return np.minimum(np.floor(x).astype(np.int64) - 1, d - 1)


def sizer_job(spec: dict) -> tuple[dict, dict]:
    rng = np.random.default_rng(spec["seed"])
    n, d = spec["rows"], spec["distinct"]
    r = _power_law_ranks(rng, n, d, spec["s"]).astype(np.uint64)
    with np.errstate(over="ignore"):
        # odd multipliers: bijections of the 64- and 32-bit rings, so d ranks
        # give d distinct keys
        k64 = (r * np.uint64(0x9E3779B97F4A7C15) + np.uint64(0x632BE59BD9B4E019)).view(np.int64)
        k32 = ((r * np.uint64(2654435761) + np.uint64(12345)) & np.uint64(0xFFFFFFFF)).astype(np.uint32).view(np.int32)
    kv = rng.random(n) >= spec["null_frac"]
    frame = {"v": ("F64", rng.random(n), None)}
    if spec["dtype"] == "I32":
        frame["k"] = ("I32", k32, kv)
        keys = [["k", "I32"]]
    elif spec["dtype"] == "I64":
        frame["k"] = ("I64", k64, kv)
        keys = [["k", "I64"]]
    else:
        frame["k"] = ("I64", k64, kv)
        frame["k2"] = ("I32", k32, None)
        keys = [["k", "I64"], ["k2", "I32"]]
    aggs = [{"kind": "Simple", "input_col": "v", "op": "Sum", "output_alias": "v_sum"},
            {"kind": "Simple", "input_col": "v", "op": "Count", "output_alias": "v_count"},
            {"kind": "Length", "op": "Len", "output_alias": "n"}]
    return frame, {"keys": keys, "aggs": aggs}
