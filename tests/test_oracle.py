"""Pins the CPU oracle before it is trusted as the checker (no GPU needed).

Each check is chosen so a plausible oracle mistake fails it: golden cases
computed independently in pure Python; exhaustive Kleene tables; the
reference's TotalOrd examples; math.fsum for exact sums; numpy.unique for
group identity; brute-force compaction; key-encoding width KATs."""
import itertools
import math

import numpy as np
import pytest

import oracle as orc
from golden_cases import CASES, expected_of, frame_of
from helpers import assert_results_equal


@pytest.fixture(scope="module", autouse=True)
def _build():
    orc.build()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_matches_golden(case):
    frame = frame_of(case)
    got = orc.groupby(case["plan"], frame)
    want = expected_of(case, frame)
    keys = [k[0] for k in case["plan"]["keys"]]
    assert_results_equal(got, want, keys, label=case["name"])


def _tri(x):
    return {None: (False, False), True: (True, True), False: (False, True)}[x]


@pytest.mark.parametrize("kind", ["And", "Or"])
def test_kleene_truth_table_exhaustive(kind):
    # shaders/logical_bool.metal:1-31 and test_logical_bool.rs:131-270:
    # AND: false dominates, OR: true dominates, otherwise null is null.
    table = {
        "And": {(True, True): True, (True, False): False, (True, None): None, (False, None): False,
                (False, False): False, (None, None): None},
        "Or": {(True, True): True, (True, False): True, (True, None): True, (False, None): None,
               (False, False): False, (None, None): None},
    }[kind]
    combos = list(itertools.product([True, False, None], repeat=2))
    a = [c[0] for c in combos]
    b = [c[1] for c in combos]
    frame = {"a": ("Bool", np.array([bool(x) for x in a]), np.array([x is not None for x in a])),
             "b": ("Bool", np.array([bool(x) for x in b]), np.array([x is not None for x in b]))}
    pred = {"kind": kind, "lhs": {"kind": "Column", "name": "a", "dtype": "Bool"},
            "rhs": {"kind": "Column", "name": "b", "dtype": "Bool"}}
    d, v = orc.eval_predicate(pred, frame)
    for i, (x, y) in enumerate(combos):
        want = table.get((x, y), table.get((y, x)))
        assert (bool(d[i]) if v[i] else None) == want, (kind, x, y)


def test_totalord_examples():
    # docs/open-questions.md:57-64: pl.Series([1.0, NaN, 3.0]) > 0 -> all true;
    # NaN == NaN -> true; NaN > NaN -> false; -0.0 == 0.0.
    x = np.array([1.0, np.nan, 3.0, -0.0, np.nan])
    y = np.array([0.0, 0.0, 0.0, 0.0, np.nan])
    frame = {"x": ("F64", x, None), "y": ("F64", y, None)}
    col = lambda n: {"kind": "Column", "name": n, "dtype": "F64"}  # noqa: E731
    gt = orc.eval_predicate({"kind": "Compare", "op": "Gt", "dtype": "F64", "lhs": col("x"), "rhs": col("y")}, frame)[0]
    eq = orc.eval_predicate({"kind": "Compare", "op": "Eq", "dtype": "F64", "lhs": col("x"), "rhs": col("y")}, frame)[0]
    assert list(gt) == [True, True, True, False, False]
    assert list(eq) == [False, False, False, True, True]


def test_compaction_brute_force():
    rng = np.random.default_rng(0)
    for n in [0, 1, 7, 8, 9, 63, 64, 65, 1000, 4099]:
        keep = rng.random(n) < 0.37
        vals = rng.integers(-1 << 40, 1 << 40, n)
        valid = rng.random(n) < 0.8
        got, gv = orc.compact(keep, "I64", vals, valid)
        np.testing.assert_array_equal(got, vals[keep])
        np.testing.assert_array_equal(gv, valid[keep])
        bits = rng.random(n) < 0.5
        gb, _ = orc.compact(keep, "Bool", bits, None)
        np.testing.assert_array_equal(gb, bits[keep])


def test_scan_kat():
    # crates/polars-metal-mlx-sys/tests/test_cumsum.rs:12-53: [1,0,1,1,0,1] -> [1,1,2,3,3,4];
    # compaction places kept row i at prefix[i] - 1.
    keep = np.array([1, 0, 1, 1, 0, 1], bool)
    got, _ = orc.compact(keep, "I64", np.arange(6), None)
    assert list(got) == [0, 2, 3, 5]
    assert list(np.cumsum(keep)) == [1, 1, 2, 3, 3, 4]


@pytest.mark.parametrize("seed", range(6))
def test_exact_sum_vs_fsum(seed):
    rng = np.random.default_rng(seed)
    parts = [rng.standard_normal(1000), rng.standard_normal(50) * 1e16, -rng.standard_normal(50) * 1e16,
             np.ldexp(rng.random(20), rng.integers(-1070, -1000, 20)), rng.random(30) * 1e-30]
    x = np.concatenate(parts)
    rng.shuffle(x)
    x = np.concatenate([x, -x[:400]])
    assert orc.exact_sum(x) == math.fsum(x.tolist())
    assert orc.exact_sum(np.array([1e16, 1.0, -1e16])) == 1.0
    assert orc.exact_sum(np.array([0.1] * 10)) == 1.0
    assert math.isnan(orc.exact_sum(np.array([np.inf, -np.inf])))
    assert orc.exact_sum(np.array([np.inf, 1.0])) == np.inf


def test_group_identity_vs_numpy_unique():
    rng = np.random.default_rng(7)
    n = 20000
    k = rng.integers(0, 500, n)
    kv = rng.random(n) < 0.9
    v = rng.standard_normal(n)
    frame = {"k": ("I64", k, kv), "v": ("F64", v, None)}
    plan = {"keys": [["k", "I64"]], "aggs": [{"kind": "Simple", "input_col": "v", "op": "Sum", "output_alias": "s"},
                                             {"kind": "Length", "op": "Len", "output_alias": "n"}]}
    got = orc.canonical(orc.groupby(plan, frame), ["k"])
    keys_valid = np.unique(k[kv])
    assert got["__n_groups__"] == len(keys_valid) + 1  # + the null key
    tag, gk, gok = got["k"]
    assert not gok[0] and gok[1:].all()
    np.testing.assert_array_equal(gk[1:], keys_valid)
    lens = got["n"][1]
    assert lens[0] == (~kv).sum()
    np.testing.assert_array_equal(lens[1:], [np.sum(k[kv] == x) for x in keys_valid])
    sums = got["s"][1]
    for i, x in enumerate(keys_valid[:50]):
        assert sums[i + 1] == math.fsum(v[kv & (k == x)].tolist())


def test_signed_zero_sum_and_min_max():
    # The reference folds f64 sums from +0.0 (groupby.rs:1303-1321): an
    # all -0.0 group sums to +0.0 in any fold order.  Min/max order -0.0
    # below +0.0, so a group holding both zeros gives min -0.0 and max +0.0
    # in every row order (the reference's Rayon fold, groupby.rs:1471-1520,
    # fixes no order).
    v = np.array([-0.0, -0.0, 0.0, -0.0, -0.0, 1.0, -1.0])
    k = np.array([0, 0, 1, 1, 2, 2, 3])
    frame = {"k": ("I64", k, None), "v": ("F64", v, None)}
    plan = {"keys": [["k", "I64"]],
            "aggs": [{"kind": "Simple", "input_col": "v", "op": op, "output_alias": op} for op in ("Sum", "Min", "Max")]}
    for mode in (0, 1):
        got = orc.canonical(orc.groupby(plan, frame, sum_mode=mode), ["k"])
        s, mn, mx = got["Sum"][1], got["Min"][1], got["Max"][1]
        assert [math.copysign(1.0, x) for x in s[:2]] == [1.0, 1.0], mode   # +0.0 sums
        assert math.copysign(1.0, mn[0]) == -1.0 and math.copysign(1.0, mx[0]) == -1.0
        assert math.copysign(1.0, mn[1]) == -1.0 and math.copysign(1.0, mx[1]) == 1.0  # both zeros: -0.0 / +0.0
        assert mn[2] == -0.0 and math.copysign(1.0, mn[2]) == -1.0 and mx[2] == 1.0 and mn[3] == -1.0


def _bits64(x):
    return np.asarray(x, np.float64).view(np.uint64)


def test_nan_poisoning_kats():
    # crates/polars-metal-kernels/tests/test_groupby_aggregate.rs:469-497:
    # min of [1, NaN, 3] and max of [10, NaN, 5] are NaN.  The NaN written is
    # Rust's f64::NAN (groupby.rs:1494), the quiet NaN 0x7FF8000000000000,
    # whatever NaN the input held.
    odd_nan = np.array([0xFFF8000000000001], np.uint64).view(np.float64)[0]
    for vals, op in (([1.0, np.nan, 3.0], "Min"), ([10.0, np.nan, 5.0], "Max"), ([odd_nan, 2.0], "Min"),
                     ([4.0, odd_nan], "Max")):
        frame = {"k": ("I64", np.zeros(len(vals), np.int64), None), "v": ("F64", np.array(vals), None)}
        plan = {"keys": [["k", "I64"]], "aggs": [{"kind": "Simple", "input_col": "v", "op": op, "output_alias": "m"}]}
        got = orc.groupby(plan, frame)["m"]
        assert got[2][0] and _bits64(got[1])[0] == 0x7FF8000000000000, (vals, op)
    # a null NaN does not poison; a NaN under a filter does not either
    frame = {"k": ("I64", np.zeros(3, np.int64), None), "v": ("F64", np.array([np.nan, 2.0, -3.0]),
                                                               np.array([False, True, True]))}
    plan = {"keys": [["k", "I64"]], "aggs": [{"kind": "Simple", "input_col": "v", "op": op, "output_alias": op}
                                             for op in ("Min", "Max")]}
    got = orc.groupby(plan, frame)
    assert got["Min"][1][0] == -3.0 and got["Max"][1][0] == 2.0


def test_float_min_max_independent_of_row_order():
    # every permutation of a group with both zeros, a NaN-free mix and
    # +-inf gives the same bits
    import itertools
    base = [0.0, -0.0, 3.5, -0.0, 0.0]
    plan = {"keys": [["k", "I64"]], "aggs": [{"kind": "Simple", "input_col": "v", "op": op, "output_alias": op}
                                             for op in ("Min", "Max")]}
    seen = set()
    for perm in itertools.permutations(base):
        frame = {"k": ("I64", np.zeros(5, np.int64), None), "v": ("F64", np.array(perm), None)}
        got = orc.groupby(plan, frame)
        seen.add((int(_bits64(got["Min"][1])[0]), int(_bits64(got["Max"][1])[0])))
    assert seen == {(0x8000000000000000, int(_bits64(3.5)))}
    for perm in itertools.permutations([0.0, -0.0, -0.0]):
        frame = {"k": ("I64", np.zeros(3, np.int64), None), "v": ("F64", np.array(perm), None)}
        got = orc.groupby(plan, frame)
        assert (_bits64(got["Min"][1])[0], _bits64(got["Max"][1])[0]) == (0x8000000000000000, 0)


def test_f32_sum_rounds_once_from_exact():
    # F32 values widen exactly to f64; the group sum is the correctly rounded
    # f64 sum, then rounded to f32.  [1, 2^-24, 2^-24]: exact 1 + 2^-23, which
    # f32 holds, where an f32 running sum rounds each step back to 1.
    # [1e8, 1, -1e8]: exact 1, where an f32 running sum gives 0.
    t = np.float32(2.0 ** -24)
    cases = [([1.0, t, t], np.float32(1.0 + 2.0 ** -23)), ([1e8, 1.0, -1e8], np.float32(1.0)),
             ([0.1, 0.2, 0.3], np.float32(float(np.float32(0.1)) + float(np.float32(0.2)) + float(np.float32(0.3))))]
    for vals, want in cases:
        frame = {"k": ("I32", np.zeros(len(vals), np.int32), None), "v": ("F32", np.array(vals, np.float32), None)}
        plan = {"keys": [["k", "I32"]], "aggs": [{"kind": "Simple", "input_col": "v", "op": op, "output_alias": op}
                                                 for op in ("Sum", "Mean", "Min", "Max")]}
        got = orc.groupby(plan, frame)
        assert got["Sum"][0] == "F32" and got["Mean"][0] == "F32"
        assert got["Sum"][1][0].view(np.uint32) == want.view(np.uint32), vals
        exact = math.fsum(float(x) for x in np.array(vals, np.float32))
        assert got["Mean"][1][0] == np.float32(exact / len(vals))
        assert got["Min"][1][0] == min(np.array(vals, np.float32)) and got["Max"][1][0] == max(np.array(vals, np.float32))
    # an F32 NaN min is the f32 quiet NaN 0x7FC00000
    frame = {"k": ("I32", np.zeros(2, np.int32), None), "v": ("F32", np.array([np.nan, 1.0], np.float32), None)}
    plan = {"keys": [["k", "I32"]], "aggs": [{"kind": "Simple", "input_col": "v", "op": "Min", "output_alias": "m"}]}
    assert orc.groupby(plan, frame)["m"][1].view(np.uint32)[0] == 0x7FC00000


def test_key_encoding_widths():
    # crates/polars-metal-kernels/tests/test_key_encoding.rs: i64 = 65 bits,
    # i64 + i32 = 98 bits, two i64 + bool... exceed 128 -> TooWide.
    n = 4
    f = {"a": ("I64", np.arange(n), None), "b": ("I32", np.arange(n, dtype=np.int32), None),
         "c": ("Bool", np.array([1, 0, 1, 0], bool), None), "d": ("I64", np.arange(n), None)}
    assert orc.encode_keys(f, ["a"])[2] == 65
    assert orc.encode_keys(f, ["a", "b"])[2] == 98
    assert orc.encode_keys(f, ["c", "b"])[2] == 2 + 33
    with pytest.raises(ValueError):
        orc.encode_keys(f, ["a", "d"])
    # null sets only the null bit: all nulls collapse into one lane
    g = {"a": ("I64", np.array([5, 6]), np.array([False, False]))}
    lo, hi, _ = orc.encode_keys(g, ["a"])
    assert lo[0] == lo[1] == 1 and hi[0] == hi[1] == 0
    # data bits start right after the null bit, LSB first
    h = {"a": ("I64", np.array([-1]), None)}
    lo, hi, _ = orc.encode_keys(h, ["a"])
    assert int(lo[0]) == (2 ** 64 - 1) << 1 & (2 ** 64 - 1) and int(hi[0]) == 1


def test_sequential_sum_mode_differs_only_by_rounding():
    frame = {"k": ("I64", np.zeros(1000, np.int64), None), "v": ("F64", np.full(1000, 0.1), None)}
    plan = {"keys": [["k", "I64"]], "aggs": [{"kind": "Simple", "input_col": "v", "op": "Sum", "output_alias": "s"}]}
    exact = orc.groupby(plan, frame, sum_mode=0)["s"][1][0]
    seq = orc.groupby(plan, frame, sum_mode=1)["s"][1][0]
    assert exact == 100.0 and seq != exact and abs(seq - exact) < 1e-10


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_all_core_groupby_is_bit_identical(threads):
    """The all-core oracle (the CPU baseline) against the one-thread oracle,
    in both sum modes: same groups, same bits (its partitions keep every
    group's rows in row order, so even row-order f64 sums agree)."""
    from polars_metal import synth
    from helpers import assert_results_equal
    frame, plan = synth.zipf(300_000, seed=3)
    frame["v"][1][::101] = np.nan
    frame["v"][1][::103] = -0.0
    for mode in (0, 1):
        a = orc.groupby(plan, frame, sum_mode=mode)
        b = orc.groupby(plan, frame, sum_mode=mode, threads=threads)
        assert_results_equal(b, a, ["key"], label=f"par{threads}/mode{mode}")
    hf, hp = synth.highcard(200_000, seed=4, pool=60_000)
    a = orc.groupby(hp, hf, sum_mode=1)
    b = orc.groupby(hp, hf, sum_mode=1, threads=threads)
    assert a["__n_groups__"] > 40_000
    assert_results_equal(b, a, ["a", "b"], label=f"par{threads}/highcard")
    # one hot key: a partition far larger than the rest, few distinct keys
    n = 100_000
    f = {"k": ("I64", np.where(np.arange(n) % 3 == 0, 7, np.arange(n)), None),
         "v": ("F64", np.random.default_rng(1).standard_normal(n), None)}
    p = {"keys": [["k", "I64"]], "aggs": [{"kind": "Simple", "input_col": "v", "op": op, "output_alias": op}
                                          for op in ("Sum", "Min", "Max", "Count")]}
    assert_results_equal(orc.groupby(p, f, sum_mode=1, threads=threads), orc.groupby(p, f, sum_mode=1), ["k"])
    # empty input and a predicate that keeps nothing
    e = {k: (t, v[:0], None) for k, (t, v, _) in frame.items()}
    assert orc.groupby(plan, e, threads=threads)["__n_groups__"] == 0


# ---------------------------------------------------------------------------
# Utf8 keys: the dictionary restatement against the reference's own KATs
# (crates/polars-metal-buffer/tests/test_dict.rs) and numpy
# ---------------------------------------------------------------------------
def test_utf8_dict_reference_kats():
    d, codes, ok = orc.utf8_dict(np.array(["b", "a", "b", "c", "a"], object), None)
    assert d == ["b", "a", "c"] and codes.tolist() == [0, 1, 0, 2, 1]  # test_dict.rs:20-28
    d, codes, ok = orc.utf8_dict(np.array(["a", "x", "", "a", "y"], object),
                                 np.array([True, False, True, True, False]))
    assert len(d) == 2 and ok.tolist() == [True, False, True, True, False]  # test_dict.rs:39-49
    assert codes[0] == 0 and codes[2] == 1 and codes[3] == 0
    d, codes, ok = orc.utf8_dict(np.array(["p", "q", "r"], object), np.zeros(3, bool))
    assert d == [] and codes.tolist() == [0, 0, 0] and not ok.any()  # test_dict.rs:52-58
    d, codes, _ = orc.utf8_dict(np.array([], object), None)
    assert d == [] and codes.size == 0  # test_dict.rs:31-36


def test_utf8_dict_first_seen_order_vs_numpy():
    rng = np.random.default_rng(4)
    words = np.array(["", "a", "ab", "Zürich", "x" * 30, "ab ", "日本"], object)
    vals = words[rng.integers(0, len(words), 5000)]
    ok = rng.random(5000) < 0.9
    d, codes, _ = orc.utf8_dict(vals, ok)
    uniq, first, inv = np.unique(vals[ok].astype(str), return_index=True, return_inverse=True)
    order = np.argsort(first)  # distinct strings by first valid row
    rank = np.empty(len(uniq), np.int64)
    rank[order] = np.arange(len(uniq))
    assert d == [str(x) for x in uniq[order]]
    np.testing.assert_array_equal(codes[ok], rank[inv])
    assert (codes[~ok] == 0).all()


def test_utf8_packed_form_round_trip():
    vals = np.array(["ab", "", "日本", "zz"], object)
    ok = np.array([True, True, True, False])
    p = orc.utf8_packed(vals, ok)
    assert p[:4] == (4).to_bytes(4, "little")
    assert np.frombuffer(p, np.int32, 5, 4).tolist() == [0, 2, 2, 8, 8]  # the null row is empty
    assert orc.utf8_unpacked(p).tolist() == ["ab", "", "日本", ""]


def test_utf8_key_groupby_vs_python():
    rng = np.random.default_rng(6)
    words = np.array(["", "a", "b", "ab"], object)
    s = words[rng.integers(0, 4, 400)]
    ok = rng.random(400) < 0.85
    k2 = rng.integers(0, 3, 400)
    v = rng.integers(-9, 9, 400)
    frame = {"s": ("Utf8", s, ok), "k2": ("I64", k2, None), "v": ("I64", v, None)}
    plan = {"keys": [["s", "Utf8"], ["k2", "I64"]],
            "aggs": [{"kind": "Simple", "input_col": "v", "op": "Sum", "output_alias": "t"}]}
    got = orc.groupby(plan, frame)
    want: dict = {}
    for i in range(400):
        key = (s[i] if ok[i] else None, int(k2[i]))
        want[key] = want.get(key, 0) + int(v[i])
    gs, gk, gt = got["s"], got["k2"], got["t"]
    seen = {((gs[1][j] if gs[2][j] else None), int(gk[1][j])): int(gt[1][j]) for j in range(got["__n_groups__"])}
    assert seen == want


# Polars' numeric supertypes for arithmetic, stated per pair (polars-core
# get_supertype): the expression restatement must type each pair this way.
SUPERTYPE_KATS = [
    ("I32", "I64", "I64"), ("I8", "I16", "I16"), ("U8", "U32", "U32"), ("I8", "U8", "I16"),
    ("I16", "U16", "I32"), ("I32", "U32", "I64"), ("I8", "U32", "I64"), ("I16", "U8", "I16"),
    ("I64", "U32", "I64"), ("I8", "F32", "F32"), ("U16", "F32", "F32"), ("I32", "F32", "F64"),
    ("U64", "F32", "F64"), ("I64", "F64", "F64"), ("F32", "F64", "F64"), ("F32", "F32", "F32"),
]


def _two_cols(a, b):
    return {"x": (a, np.array([3], orc.NP_DTYPE[a]), None), "y": (b, np.array([2], orc.NP_DTYPE[b]), None)}


@pytest.mark.parametrize("a,b,want", SUPERTYPE_KATS)
def test_expression_supertypes(a, b, want):
    add = {"kind": "Binary", "op": "Add", "lhs": {"kind": "Column", "name": "x"}, "rhs": {"kind": "Column", "name": "y"}}
    for e in (add, dict(add, lhs=add["rhs"], rhs=add["lhs"])):
        tag, vals, ok = orc.eval_expr(e, _two_cols(a, b))
        assert tag == want and vals[0] == 5 and ok.all()


def test_expression_division_and_literal_typing():
    def ev(expr, frame):
        return orc.eval_expr(expr, frame)

    col = lambda n: {"kind": "Column", "name": n}  # noqa: E731
    li = lambda v: {"kind": "LiteralI64", "value": v}  # noqa: E731
    lf = lambda v: {"kind": "LiteralF64", "value": v}  # noqa: E731
    b = lambda op, x, y: {"kind": "Binary", "op": op, "lhs": x, "rhs": y}  # noqa: E731
    fr = {"i": ("I32", np.array([7, -7, 0], np.int32), None), "f": ("F32", np.array([1.5, 0.1, -2], np.float32), None),
          "h": ("I16", np.array([3, 0, 1], np.int16), None), "u": ("U8", np.array([0, 1, 255], np.uint8), None)}
    # true division of integers is Float64, with IEEE infinities
    tag, v, _ = ev(b("Div", col("i"), col("h")), fr)
    assert tag == "F64" and v[0] == 7 / 3 and v[1] == -math.inf and v[2] == 0.0
    tag, v, _ = ev(b("Div", col("i"), li(2)), fr)
    assert tag == "F64" and v[0] == 3.5
    # F32 with I16 or with a literal stays F32, rounded once per operation
    tag, v, _ = ev(b("Div", col("f"), col("h")), fr)
    assert tag == "F32" and v[0] == np.float32(0.5)
    tag, v, _ = ev(b("Mul", col("f"), lf(0.1)), fr)
    assert tag == "F32" and v[1] == np.float32(np.float32(0.1) * np.float32(0.1))
    # an integer literal takes the column's integer type and wraps in it
    tag, v, _ = ev(b("Sub", col("u"), li(1)), fr)
    assert tag == "U8" and list(v) == [255, 0, 254]
    tag, v, _ = ev(b("Mul", col("i"), lf(0.5)), fr)
    assert tag == "F64" and v[0] == 3.5
    tag, v, _ = ev(b("Add", col("i"), b("Mul", li(2), li(3))), fr)  # literal subtree folded
    assert tag == "I32" and v[0] == 13
    with pytest.raises(NotImplementedError):
        ev(b("Add", col("u"), li(300)), fr)
    with pytest.raises(NotImplementedError):
        ev(b("Add", col("i"), {"kind": "Column", "name": "q"}), dict(fr, q=("U64", np.zeros(3, np.uint64), None)))
    # nulls of every column read propagate
    fr2 = dict(fr, h=("I16", np.array([3, 0, 1], np.int16), np.array([True, False, True])))
    _, _, ok = ev(b("Add", col("i"), b("Mul", col("h"), li(0))), fr2)
    assert list(ok) == [True, False, True]
