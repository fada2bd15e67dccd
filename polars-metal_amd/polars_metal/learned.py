"""The learned sizer: a small MLP that reads the partitioned engine's
64K-row sample sketch and predicts the query's group count and whether the
sort engine beats the hash engine.

No counterpart in the reference, which sizes nothing from data (its nearest
precedents are fixed thresholds: crates/polars-metal-core/src/cost.rs:38,
crates/polars-metal-kernels/src/groupby.rs:493).  Training and inference
run on the device through the library's own forward, backward and Adam
kernels (pmh_mlp_*); this module only moves arrays and logs around.

The model never changes a result: under ``method: "auto"`` it chooses the
engine and the cardinality the hash partitions are sized for, and the
engines' overflow fallbacks keep every output exact whatever it predicts.

Targets per logged job: ``y0 = log2(groups + 1) / 32`` and ``y1 = 1`` if the
sort engine, reached through the sketch the router runs first, was faster
than the plain hash path, ``0`` otherwise (``-1``: not timed)."""
from __future__ import annotations

import ctypes
import json
import time
from pathlib import Path

import numpy as np

from . import _lib
from ._native import _device_column_array, _stats_dict

MLP_IN = 16
MLP_HIDDEN = 64
MLP_PARAMS = MLP_IN * 64 + 64 + 64 * 64 + 64 + 64 * 2 + 2
MODEL_PATH = Path(__file__).resolve().parent / "sizer_model.json"

_F32P = ctypes.POINTER(ctypes.c_float)


def _fp(a: np.ndarray):
    return a.ctypes.data_as(_F32P)


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights and zero biases, in the library's parameter layout
    (W1[16][64], b1, W2[64][64], b2, W3[64][2], b3)."""
    rng = np.random.default_rng(seed)
    parts = []
    for rows, cols in ((MLP_IN, 64), (64, 64), (64, 2)):
        parts.append(rng.standard_normal(rows * cols) * np.sqrt(2.0 / rows))
        parts.append(np.zeros(cols))
    return np.concatenate(parts).astype(np.float32)


class DeviceMLP:
    """A model resident on the device (pmh_mlp_*)."""

    def __init__(self, params: np.ndarray):
        p = np.ascontiguousarray(params, np.float32)
        if p.size != MLP_PARAMS:
            raise ValueError(f"expected {MLP_PARAMS} parameters, got {p.size}")
        h = ctypes.c_void_p()
        _lib.check(_lib.load().pmh_mlp_new(_fp(p), ctypes.byref(h)))
        self._h = h

    def params(self) -> np.ndarray:
        out = np.empty(MLP_PARAMS, np.float32)
        _lib.check(_lib.load().pmh_mlp_params(self._h, _fp(out)))
        return out

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_IN)
        y = np.empty((x.shape[0], 2), np.float32)
        _lib.check(_lib.load().pmh_mlp_forward(self._h, _fp(x), x.shape[0], _fp(y)))
        return y

    def grad(self, x: np.ndarray, y: np.ndarray, lam: float = 1.0) -> tuple[np.ndarray, float]:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_IN)
        y = np.ascontiguousarray(y, np.float32).reshape(-1, 2)
        g = np.empty(MLP_PARAMS, np.float32)
        loss = ctypes.c_float()
        _lib.check(_lib.load().pmh_mlp_grad(self._h, _fp(x), _fp(y), x.shape[0], lam, _fp(g), ctypes.byref(loss)))
        return g, float(loss.value)

    def fit(self, x: np.ndarray, y: np.ndarray, epochs: int = 200, batch: int = 64, lr: float = 3e-3,
            lam: float = 1.0, seed: int = 0) -> float:
        """Minibatch Adam on the device; returns the last epoch's mean loss."""
        x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_IN)
        y = np.ascontiguousarray(y, np.float32).reshape(-1, 2)
        lib = _lib.load()
        _lib.check(lib.pmh_mlp_set_data(self._h, _fp(x), _fp(y), x.shape[0]))
        loss = ctypes.c_float()
        _lib.check(lib.pmh_mlp_train(self._h, epochs, batch, lr, lam, seed, ctypes.byref(loss)))
        return float(loss.value)

    def free(self) -> None:
        if self._h:
            _lib.load().pmh_mlp_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def install(params: np.ndarray | None) -> None:
    """Make ``params`` the model that ``method: "auto"`` consults (None
    uninstalls)."""
    lib = _lib.load()
    if params is None:
        _lib.check(lib.pmh_sizer_install(None))
        return
    p = np.ascontiguousarray(params, np.float32)
    if p.size != MLP_PARAMS:
        raise ValueError(f"expected {MLP_PARAMS} parameters, got {p.size}")
    _lib.check(lib.pmh_sizer_install(_fp(p)))


def install_shipped() -> dict:
    """Install the model shipped with the package (trained on the device by
    ``python -m polars_metal.train_sizer``); returns its training report."""
    d = json.loads(MODEL_PATH.read_text())
    install(np.asarray(d["params"], np.float32))
    return d.get("meta", {})


def load_model(path: Path = MODEL_PATH) -> np.ndarray:
    d = json.loads(Path(path).read_text())
    return np.asarray(d["params"], np.float32)


def save_model(params: np.ndarray, path: Path = MODEL_PATH, meta: dict | None = None) -> None:
    Path(path).write_text(json.dumps({"params": [float(v) for v in params], "meta": meta or {}}))


def sketch_features(plan: dict, n_rows: int, columns: list) -> tuple[np.ndarray, dict]:
    """The sample sketch of a query (device columns as for
    execute_filter_groupby_device): its features and sketch stats."""
    lib = _lib.load()
    arr, keep = _device_column_array(list(columns))
    feat = np.zeros(MLP_IN, np.float32)
    st = _lib.PmhExecStats()
    _lib.check(lib.pmh_sketch_features(json.dumps(plan).encode(), n_rows, arr, len(columns), _fp(feat),
                                       ctypes.byref(st)))
    del keep
    return feat, _stats_dict(st)


def sketch_inspect(plan: dict, n_rows: int, columns: list) -> dict:
    """The whole sample sketch of a query: its counts and histogram, the
    features, the plain distinct estimate, and the installed model's outputs
    as the device's sketch block computed them (NaN without a model)."""
    lib = _lib.load()
    arr, keep = _device_column_array(list(columns))
    info = _lib.PmhSketchInfo()
    _lib.check(lib.pmh_sketch_inspect(json.dumps(plan).encode(), n_rows, arr, len(columns), ctypes.byref(info)))
    del keep
    out = {name: getattr(info, name) for name, _ in _lib.PmhSketchInfo._fields_}
    out["hist"] = np.ctypeslib.as_array(info.hist).copy()
    out["features"] = np.ctypeslib.as_array(info.features).copy()
    out["prediction"] = np.ctypeslib.as_array(info.prediction).copy()
    return out


def targets(n_groups: int, t_hash_ms: float | None, t_sort_ms: float | None) -> np.ndarray:
    y1 = -1.0 if t_hash_ms is None or t_sort_ms is None else (1.0 if t_sort_ms < t_hash_ms else 0.0)
    return np.array([np.log2(n_groups + 1.0) / 32.0, y1], np.float32)


def timed(run_query, plan: dict, n: int, cols: list) -> tuple[dict, float]:
    """Stats and wall milliseconds of one synchronous query."""
    t = time.perf_counter()
    st = run_query(plan, n, cols)
    return st, (time.perf_counter() - t) * 1e3


def collect_logs(jobs, run_query, log=print) -> tuple[np.ndarray, np.ndarray, list]:
    """Logs of the plain method: for each job ``(plan, n_rows, columns)``,
    the sketch features, the true group count and both engines' wall times
    (each call returns only after its device work).  A query the sizer
    routes to sort has run the sketch first, so the sort time logged is the
    sketch's plus the sort engine's.  ``run_query(plan, n_rows, columns)``
    runs one query and returns its stats."""
    xs, ys, meta = [], [], []
    t0 = time.perf_counter()
    for i, (plan, n, cols) in enumerate(jobs):
        t = time.perf_counter()
        feat, _ = sketch_features(plan, n, cols)
        tk = (time.perf_counter() - t) * 1e3
        sh, th = timed(run_query, dict(plan, options={"method": "hash"}), n, cols)
        ss, ts = timed(run_query, dict(plan, options={"method": "sort"}), n, cols)
        if sh["n_groups"] != ss["n_groups"]:
            raise AssertionError(f"job {i}: hash and sort disagree on the group count")
        xs.append(feat)
        ys.append(targets(sh["n_groups"], th, ts + tk))
        meta.append({"rows": n, "groups": sh["n_groups"], "hash_ms": th, "sort_ms": ts + tk, "sketch_ms": tk})
        if log and (i + 1) % 100 == 0:
            log(f"logged {i + 1} jobs in {time.perf_counter() - t0:.1f} s")
    return np.stack(xs), np.stack(ys), meta
