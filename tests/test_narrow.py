"""Key-lane narrowing in the partitioned engine: a composite key wider than
64 bits whose columns' value ranges fit 64 bits runs as a narrow lane (each
column's data as its value minus the column's smallest, in the bits its
range needs; a null bit only for a column with nulls).  Every case is checked
bit for bit against the oracle and against the same engine without
narrowing; the cases cover signed and unsigned, 8- to 64-bit and float keys,
nulls, a constant column, and a narrowed lane of exactly 64 bits whose
all-ones value is a real key.  Ranges come from a 65536-row sample that
the specialised scan checks every selected key against (a key outside them
re-runs the query with ranges from every row), or from every row."""
import numpy as np
import pytest

import oracle as orc
from helpers import assert_results_equal, run_gpu

N = 1_200_000  # above the narrowing threshold (2^20 rows)

AGGS = [{"kind": "Simple", "input_col": "v", "op": op, "output_alias": f"v_{op.lower()}"}
        for op in ("Sum", "Count", "Min", "Max")] + [{"kind": "Length", "op": "Len", "output_alias": "n"}]


def _case(name, rng):
    v = rng.integers(-10**9, 10**9, N)
    vv = rng.random(N) < 0.9
    if name == "signed_ranges":  # i32 in [-5000, 5000), i64 in [-2^40, -2^40 + 2^30)
        f = {"a": ("I32", rng.integers(-5000, 5000, N).astype(np.int32), None),
             "b": ("I64", rng.integers(-(1 << 40), -(1 << 40) + (1 << 30), N), rng.random(N) < 0.97)}
    elif name == "small_ints":  # i8 / u16 / u64 keys with nulls
        f = {"a": ("I8", rng.integers(-128, 128, N).astype(np.int8), rng.random(N) < 0.95),
             "b": ("U16", rng.integers(0, 1 << 16, N).astype(np.uint16), None),
             "c": ("U64", (np.uint64(1) << np.uint64(63)) + rng.integers(0, 1 << 20, N).astype(np.uint64), None)}
    elif name == "constant_and_float":  # one constant column, one f32 key of few values
        f = {"a": ("I32", np.full(N, 7, np.int32), None),
             "b": ("F32", rng.choice(np.array([-0.0, 0.0, 1.5, -2.25, np.inf], np.float32), N), None),
             "c": ("I16", rng.integers(-150, 150, N).astype(np.int16), rng.random(N) < 0.9)}
    else:  # "full_64": two full-range i32 keys -> exactly 64 bits; the all-ones lane occurs
        a = rng.integers(-(1 << 31), 1 << 31, N).astype(np.int32)
        b = rng.integers(-(1 << 31), 1 << 31, N).astype(np.int32)
        a[:3] = [np.iinfo(np.int32).min, np.iinfo(np.int32).max, np.iinfo(np.int32).max]
        b[:3] = [np.iinfo(np.int32).min, np.iinfo(np.int32).max, np.iinfo(np.int32).max]
        a[::1000], b[::1000] = np.iinfo(np.int32).max, np.iinfo(np.int32).max  # the all-ones lane
        f = {"a": ("I32", a, None), "b": ("I32", b, None)}
    keys = list(f)
    f["v"] = ("I64", v, vv)
    return f, {"keys": [[k, f[k][0]] for k in keys], "aggs": AGGS}, keys


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["signed_ranges", "small_ints", "constant_and_float", "full_64"])
def test_narrowed_lane_matches_oracle(native, name):
    rng = np.random.default_rng(["signed_ranges", "small_ints", "constant_and_float", "full_64"].index(name))
    frame, plan, keys = _case(name, rng)
    want = orc.groupby(plan, frame)
    for narrow in (1, 2, 0):  # ranges from a checked sample, from every row, no narrowing
        native.set_profiling(True)
        try:
            got = run_gpu(native, dict(plan, options={"method": "hash", "narrow_keys": narrow}), frame)
            kt = native.kernel_times()
        finally:
            native.set_profiling(False)
        st = native.last_stats()
        print(name, narrow, st["method"], st["n_groups"], st["partitions"], st["scan_kernel"], sorted(kt), flush=True)
        assert st["method"] == "hash"
        assert ("hs_key_range" in kt) == bool(narrow)
        # narrowed, the query takes the narrow lane's specialised scan
        assert st["scan_kernel"] == ("specialised" if narrow else "generic")
        assert_results_equal(got, want, keys, label=f"{name}-narrow{narrow}")


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["high", "low", "null"])
def test_sampled_ranges_miss(native, case):
    """A key outside the sampled ranges in a row off the sample grid (above
    them, below them, or a null where the sample saw none) sends the query
    through ranges from every row; the result matches the oracle and the
    exact ranges."""
    rng = np.random.default_rng(["high", "low", "null"].index(case) + 30)
    n = (1 << 21) + 3
    a = rng.integers(0, 1000, n).astype(np.int32)
    b = rng.integers(1 << 40, (1 << 40) + (1 << 20), n)
    av = None
    if case == "high":
        b[5] = 1 << 50
    elif case == "low":
        a[7] = -5
    else:
        av = np.ones(n, bool)
        av[9] = False
    frame = {"a": ("I32", a, av), "b": ("I64", b, None), "v": ("I64", rng.integers(-100, 100, n), None)}
    plan = {"keys": [["a", "I32"], ["b", "I64"]], "aggs": AGGS}
    want = orc.groupby(plan, frame)
    for narrow in (1, 2):
        print(f"sampled ranges {case} narrow={narrow}", flush=True)
        got = run_gpu(native, dict(plan, options={"method": "hash", "narrow_keys": narrow}), frame)
        assert native.last_stats()["method"] == "hash"
        assert_results_equal(got, want, ["a", "b"], label=f"{case}-narrow{narrow}")
