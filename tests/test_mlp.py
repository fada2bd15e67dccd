"""The learned sizer's network: the CPU restatement (oracle/mlp.py) pinned by
finite differences and Adam's closed form, then the device kernels
(pmh_mlp_*) checked against it, trained, and used for routing."""
import math

import numpy as np
import pytest

import mlp as mo
import oracle as orc
import sizer as osz
from helpers import assert_results_equal, run_gpu


def _batch(seed, n):
    rng = np.random.default_rng(seed)
    x = rng.random((n, mo.F)) * 2 - 0.5
    y = np.stack([rng.random(n), rng.choice([-1.0, 0.0, 1.0], n)], axis=1)
    return x, y


# ---------------------------------------------------------------------------
# CPU: pin the oracle
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("lam", [0.0, 1.0, 2.5])
def test_oracle_grad_matches_finite_differences(lam):
    p = mo.init_params(3)
    x, y = _batch(4, 17)
    g = mo.grad(p, x, y, lam)
    rng = np.random.default_rng(5)
    # every bias and a sample of weights in each layer
    idx = np.concatenate([rng.choice(mo.N_PARAMS, 300, replace=False), np.arange(mo.N_PARAMS - 2, mo.N_PARAMS)])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (mo.loss(p + e, x, y, lam) - mo.loss(p - e, x, y, lam)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_oracle_adam_first_step_is_lr_sign():
    p = np.zeros(5)
    g = np.array([3.0, -2.0, 1e-3, -1e-3, 0.0])
    p1, m, v = mo.adam_step(p, g, np.zeros(5), np.zeros(5), 1, 0.01)
    np.testing.assert_allclose(p1[:4], -0.01 * np.sign(g[:4]), rtol=1e-4)
    assert p1[4] == 0.0


def test_oracle_layout_matches_library():
    from polars_metal import learned
    assert mo.N_PARAMS == learned.MLP_PARAMS
    assert mo.F == learned.MLP_IN


# ---------------------------------------------------------------------------
# GPU: the kernels against the oracle
# ---------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 5, 64, 333])
def test_device_forward_matches_oracle(native, n):
    from polars_metal import learned
    p = learned.init_params(7)
    x, _ = _batch(8, n)
    m = learned.DeviceMLP(p)
    got = m.forward(x.astype(np.float32))
    want, _ = mo.forward(p.astype(np.float64), x.astype(np.float32).astype(np.float64))
    np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-5)
    np.testing.assert_array_equal(m.params(), p)


@pytest.mark.gpu
@pytest.mark.parametrize("n,lam", [(1, 1.0), (40, 1.0), (513, 0.5), (96, 0.0)])
def test_device_gradient_matches_oracle(native, n, lam):
    from polars_metal import learned
    p = learned.init_params(9)
    x, y = _batch(10, n)
    x32, y32 = x.astype(np.float32), y.astype(np.float32)
    m = learned.DeviceMLP(p)
    g, loss = m.grad(x32, y32, lam)
    gw = mo.grad(p.astype(np.float64), x32.astype(np.float64), y32.astype(np.float64), lam)
    lw = mo.loss(p.astype(np.float64), x32.astype(np.float64), y32.astype(np.float64), lam)
    assert abs(loss - lw) <= 1e-4 * max(1.0, abs(lw))
    # fp32 accumulation against float64: scale the tolerance by the largest
    # gradient component
    scale = float(np.abs(gw).max())
    np.testing.assert_allclose(g, gw, rtol=0, atol=2e-4 * scale)
    # and the direction is the same
    cos = float(g @ gw / (np.linalg.norm(g) * np.linalg.norm(gw)))
    assert cos > 0.99999


@pytest.mark.gpu
def test_device_training_fits_a_known_function(native):
    """Adam on the device fits y0 = mean(x) and y1 = [x0 > x1]; its first
    step equals the oracle's Adam step from the same gradient."""
    from polars_metal import learned
    rng = np.random.default_rng(12)
    x = rng.random((2048, mo.F)).astype(np.float32)
    y = np.stack([x.mean(1), (x[:, 0] > x[:, 1]).astype(np.float32)], axis=1).astype(np.float32)
    p0 = learned.init_params(13)
    m = learned.DeviceMLP(p0)
    g, l0 = m.grad(x[:64], y[:64])
    # one step on one batch of 64 (the whole "data set")
    m1 = learned.DeviceMLP(p0)
    m1.fit(x[:64], y[:64], epochs=1, batch=64, lr=1e-3)
    p1, _, _ = mo.adam_step(p0.astype(np.float64), g.astype(np.float64), 0.0, 0.0, 1, 1e-3)
    np.testing.assert_allclose(m1.params(), p1, rtol=0, atol=1e-5)
    loss = m.fit(x, y, epochs=60, batch=64, lr=3e-3)
    assert loss < 0.05 * l0
    pred = m.forward(x)
    assert np.mean((pred[:, 1] > 0) == (y[:, 1] > 0.5)) > 0.95
    assert np.median(np.abs(pred[:, 0] - y[:, 0])) < 0.02


# ---------------------------------------------------------------------------
# CPU: pin the restatement of the sketch estimators and features
# ---------------------------------------------------------------------------
def _hist(**f):
    h = np.zeros(1024, np.int64)
    for c, v in f.items():
        h[int(c[1:])] = v
    return h


def test_sketch_estimators_hand_worked():
    """f1 = 4, f2 = 2, f3 = 1 (d = 7, s = 11 rows) sampled from 22 rows
    (q = 1/2): every estimator worked out by hand."""
    k = osz.sketch(_hist(c1=4, c2=2, c3=1), s_sel=11, n_sample=11, n_rows=22, d=7, f1=4, f2=2, dropped=0)
    assert k["n_sel"] == 22.0
    assert k["F2"] == 1 * 4 + 4 * 2 + 9 * 1
    # sum (1-q)^c f_c = 4/2 + 2/4 + 1/8; sum c q (1-q)^(c-1) f_c = 4/2 + 2*2/4 + 3/8
    assert math.isclose(k["shl_num"], 2.625) and math.isclose(k["shl_den"], 3.375)
    assert math.isclose(osz.shlosser(k), 7 + 4 * 2.625 / 3.375)
    assert math.isclose(osz.gee(k), math.sqrt(2.0) * 4 + 3)
    assert math.isclose(osz.chao1(k), 7 + 4 * 3 / (2 * 3))
    want = max(7.0, math.sqrt((math.sqrt(2.0) * 4 + 3) * (7 + 4 * 2.625 / 3.375)), 9.0)
    assert math.isclose(osz.plain_estimate(k), want)


def test_sketch_estimator_regimes():
    # the sample is the whole selection (q = 1): Shlosser adds nothing, the estimate is d
    k = osz.sketch(_hist(c1=5, c2=3), s_sel=11, n_sample=11, n_rows=11, d=8, f1=5, f2=3, dropped=0)
    assert osz.shlosser(k) == 8 and osz.plain_estimate(k) == 8
    # GEE scales the singletons by sqrt(n_sel / s): sqrt(100) * 40 + 20
    k = osz.sketch(_hist(c1=40, c2=20, c3=20), s_sel=100, n_sample=100, n_rows=10_000, d=60, f1=40, f2=20,
                   dropped=0)
    assert math.isclose(osz.gee(k), 420.0)
    # a sample of singletons extrapolates linearly: n_sel * d / s
    k = osz.sketch(_hist(c1=1000), s_sel=1000, n_sample=2000, n_rows=1_000_000, d=1000, f1=1000, f2=0, dropped=0)
    assert math.isclose(osz.plain_estimate(k), 500_000 * 1000 / 1000)
    # dropped rows count as singletons, and dropping most of the sample extrapolates linearly
    k = osz.sketch(_hist(c1=10, c2=10), s_sel=100, n_sample=100, n_rows=1000, d=20, f1=10, f2=10, dropped=70)
    assert k["d"] == 90 and k["f1"] == 80
    assert math.isclose(osz.plain_estimate(k), 1000 * 90 / 100)
    # nothing selected
    k = osz.sketch(np.zeros(1024), s_sel=0, n_sample=100, n_rows=1000, d=0, f1=0, f2=0, dropped=0)
    assert osz.plain_estimate(k) == 0.0 and osz.gee(k) == 0.0


def test_sketch_features_closed_forms():
    k = osz.sketch(_hist(c1=4, c2=2, c3=1), s_sel=11, n_sample=16, n_rows=1 << 20, d=7, f1=4, f2=2, dropped=0)
    for bits, hot in [(17, 12), (33, 12), (34, 13), (65, 13), (66, 14), (97, 14)]:
        f = osz.features(k, null_keys=3, lane_bits=bits)
        assert f[12:15].sum() == 1 and f[hot] == 1
    f = osz.features(k, null_keys=3, lane_bits=65)
    assert math.isclose(f[0] * 32, math.log2((1 << 20) + 1))
    assert f[1] == 11 / 16 and f[3] == 4 / 11 and f[4] == 2 / 11 and f[5] == 7 / 11
    assert f[9] == 21 / 121 and f[10] == 3 / 11 and f[11] == 0
    assert math.isclose(2 ** (f[8] * 32) - 1, osz.plain_estimate(k))


# ---------------------------------------------------------------------------
# GPU: features and routing
# ---------------------------------------------------------------------------
@pytest.mark.gpu
def test_sketch_features(native):
    from polars_metal import learned, synth
    from polars_metal.buffer import device_columns
    frame, plan = synth.zipf(500_000, seed=3)
    cols, bufs = device_columns(frame)
    f, st = learned.sketch_features(plan, 500_000, cols)
    assert f.shape == (16,) and np.isfinite(f).all()
    assert st["sample_rows"] > 15_000 and 0 < st["sample_distinct"] <= st["sample_rows"]
    assert abs(f[1] - 0.3) < 0.02                   # sample selectivity of (w >= 200) & (w < 500)
    assert f[13] == 1.0 and f[12] == 0.0 and f[14] == 0.0  # one i64 key: the 65-bit lane class
    assert abs(f[0] * 32 - np.log2(500_001)) < 1e-4


def _inspect_case(kind):
    from polars_metal import synth
    rng = np.random.default_rng(11)
    if kind == "zipf":
        return synth.zipf(500_000, seed=3) + (500_000,)
    if kind == "uniform_recount":  # ~60K distinct sampled keys: the fingerprint table drops, then recounts
        n = 4_000_000
        frame = {"key": ("I64", rng.integers(1, n // 10 + 1, n).astype(np.int64), None), "v": ("F64", rng.random(n), None)}
        plan = {"keys": [["key", "I64"]], "aggs": [{"kind": "Simple", "input_col": "v", "op": "Sum",
                                                     "output_alias": "s"}]}
        return frame, plan, n
    if kind == "wide_nulls":  # 98-bit lanes: a null key is a key value of the lane
        frame, plan = synth.sizer_job({"rows": 800_000, "distinct": 30_000, "s": 0.9, "dtype": "U128",
                                       "null_frac": 0.15, "seed": 5})
        return frame, plan, 800_000
    if kind == "i64_nulls":  # 65-bit lanes: null keys are side keys, counted apart
        frame, plan = synth.sizer_job({"rows": 900_000, "distinct": 200_000, "s": 0.4, "dtype": "I64",
                                       "null_frac": 0.12, "seed": 7})
        return frame, plan, 900_000
    if kind == "i32_small":  # fewer rows than the sample: the sample is the selection
        frame, plan = synth.sizer_job({"rows": 20_000, "distinct": 3_000, "s": 1.2, "dtype": "I32",
                                       "null_frac": 0.05, "seed": 6})
        return frame, plan, 20_000
    frame, plan = synth.zipf(200_000, seed=4)  # "none": a predicate no row passes
    plan = dict(plan, predicate=synth.pred_gt_f64("v", 1e9))  # v is standard normal
    return frame, plan, 200_000


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["zipf", "uniform_recount", "wide_nulls", "i64_nulls", "i32_small", "none"])
def test_sketch_block_sizer_matches_restatement(native, kind):
    """The sketch block computes the sizer's features, the plain distinct
    estimate and (model installed) the forward pass on the device; the CPU
    restatement of the same sketch gives the same numbers."""
    from polars_metal import learned
    from polars_metal.buffer import device_columns
    frame, plan, n = _inspect_case(kind)
    cols, bufs = device_columns(frame)
    params = learned.load_model()
    learned.install(params)
    try:
        got = learned.sketch_inspect(plan, n, cols)
    finally:
        learned.install(None)
    bare = learned.sketch_inspect(plan, n, cols)
    print(kind, {k: v for k, v in got.items() if k not in ("hist", "features")}, flush=True)
    for sk in (got, bare):
        h = sk["hist"].astype(np.int64)
        assert sk["sample_size"] == min(n, 1 << 16)
        if kind in ("uniform_recount", "wide_nulls", "i64_nulls"):  # more sampled keys than fingerprint slots
            assert sk["scale"] == 8 and sk["dropped"] % 8 == 0 and (h % 8 == 0).all()
        else:
            assert sk["scale"] == 1
        if kind == "wide_nulls":
            assert sk["lane_bits"] > 65
        if kind == "i64_nulls":
            assert sk["lane_bits"] == 65 and sk["null_keys"] > 0
        if kind == "none":
            assert sk["sample_rows"] == 0 and sk["distinct_estimate"] == 0
        assert h[1:].sum() == sk["distinct"] and h[1] == sk["f1"] and h[2] == sk["f2"]
        if sk["scale"] == 1:  # the counts cover the counted rows (counts clip at 1023)
            assert (np.arange(1024) * h)[1:].sum() <= sk["sample_rows"] - sk["null_keys"] - sk["dropped"]
        k = osz.sketch(h, sk["sample_rows"], sk["sample_size"], n, sk["distinct"], sk["f1"], sk["f2"],
                       sk["dropped"])
        want = osz.features(k, sk["null_keys"], sk["lane_bits"])
        np.testing.assert_allclose(sk["features"], want, rtol=2e-6, atol=1e-7)
        assert math.isclose(sk["distinct_estimate"], osz.plain_estimate(k), rel_tol=1e-12, abs_tol=1e-12)
    o, _ = mo.forward(params, got["features"][None, :])
    np.testing.assert_allclose(got["prediction"], o[0], rtol=0, atol=2e-5)
    # without a model there is no prediction (the sketch itself varies from
    # call to call only in which keys win contended fingerprint slots)
    assert np.isnan(bare["prediction"]).all()


@pytest.mark.gpu
@pytest.mark.parametrize("bias", [-20.0, 20.0])
def test_installed_sizer_routes_without_changing_results(native, bias):
    """A model whose path logit is forced negative (hash) or positive (sort):
    the route follows it, the results stay bit-identical to the oracle."""
    from polars_metal import learned, synth
    p = learned.init_params(1)
    p[-1] = bias  # b3[1]: the path logit's bias
    frame, plan = synth.zipf(300_000, seed=9)
    want = orc.groupby(plan, frame)
    learned.install(p)
    try:
        got = run_gpu(native, dict(plan, options={"method": "auto"}), frame)
        st = native.last_stats()
    finally:
        learned.install(None)
    assert st["sizer_choice"] == ("sort" if bias > 0 else "hash")
    assert st["method"] == ("sort" if bias > 0 else "hash")
    assert_results_equal(got, want, ["key"], label=f"sizer{bias}")
    # uninstalled: auto is the plain method again
    run_gpu(native, dict(plan, options={"method": "auto"}), frame)
    assert native.last_stats()["sizer_choice"] is None


# ---------------------------------------------------------------------------
# GPU: the plain sketch estimate the partitions are sized from
# ---------------------------------------------------------------------------
def _estimate_case(kind, rows, seed):
    from polars_metal import synth
    rng = np.random.default_rng(seed)
    if kind == "zipf":  # the 100M-row workload's key distribution, with its filter
        return synth.zipf(rows, seed=seed)
    keys = rng.integers(1, rows // 10 + 1, rows) if kind == "uniform" else rng.permutation(rows) + 7
    frame = {"key": ("I64", keys.astype(np.int64), None), "v": ("F64", rng.random(rows), None)}
    plan = {"keys": [["key", "I64"]], "aggs": [{"kind": "Simple", "input_col": "v", "op": "Sum",
                                                 "output_alias": "s"}]}
    return frame, plan


@pytest.mark.gpu
@pytest.mark.parametrize("kind,rows,hi", [("zipf", 3_000_000, 2.0), ("uniform", 4_000_000, 4.0),
                                         ("unique", 2_000_000, 2.0)])
def test_sketch_estimate_tracks_group_count(native, kind, rows, hi):
    """The partitioned engine's distinct estimate (64K-row sample, 16K-slot
    fingerprint table with capped probes) lands within a small factor of the
    true count.  A sample whose rare keys overflow the probe cap must not be
    read as a near-unique one (that reading put an 846K-group zipf query at
    ~10M and split it 16x too finely); a sample with far more distinct keys
    than slots is sketched from a hash subsample.  The upper bound for
    uniform keys is looser: at 1.6% sampling, GEE and Shlosser's estimator
    straddle the truth by ~1/1.1 and ~8x, so their geometric mean overshoots
    ~3x even with an exact sample profile."""
    from polars_metal.buffer import device_columns
    frame, plan = _estimate_case(kind, rows, 21)
    cols, bufs = device_columns(frame)
    r = native.execute_filter_groupby_device(dict(plan, options={"method": "hash"}), rows, cols)
    st = r.stats
    r.free()
    ratio = st["distinct_estimate"] / st["n_groups"]
    assert 0.5 <= ratio <= hi, (kind, st["distinct_estimate"], st["n_groups"])
