"""Learned tile-configuration cost model (tq_costmodel_* in the C-ABI).

A multilayer perceptron maps the features of one MFMA tile class of a plan
under one tile configuration (pair-shape histograms, tile counts, padding
efficiency) to the log of that class's grouped-GEMM time. The network is
trained and evaluated on the GPU by the library's own kernels. Samples come
from the autotuner: `record` autotunes a plan and returns its
(features, log time) rows, `CostModel.fit` trains on them, and
`CostModel.choose` configures a plan from the predictions alone.
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import lib
from .tensor import default_handle

N_FEATURES = 34  # TQ_COST_FEATURES


def _fp(a):
    return a.ctypes.data_as(_lib.c_f32_p)


def n_configs(precision="native"):
    return lib().tq_gemm_config_count(_lib.PRECISIONS[precision])


def plan_features(plan, cls):
    """[n_configs, N_FEATURES] features of tile class `cls` of `plan` under
    every configuration."""
    nc = n_configs(plan.precision)
    out = np.zeros((nc, N_FEATURES), np.float32)
    for q in range(nc):
        plan.h.check(lib().tq_plan_features(plan.ptr, int(cls), q, _fp(out[q])))
    return out


def shape_features(m, n, k, config, precision="native", n_cu=256):
    """Features of block pairs (m, n, k) under one configuration (host only)."""
    m, n, k = (np.ascontiguousarray(v, np.int64) for v in (m, n, k))
    out = np.zeros(N_FEATURES, np.float32)
    p = ctypes.POINTER(ctypes.c_int64)
    _lib.check(lib().tq_cost_features(m.ctypes.data_as(p), n.ctypes.data_as(p),
                                      k.ctypes.data_as(p), m.size, _lib.PRECISIONS[precision],
                                      int(config), int(n_cu), _fp(out)))
    return out


def record(plan, a_data, b_data, out, reps=2):
    """Autotunes `plan` and returns (x [s, N_FEATURES], log-ms [s], times)
    for every measured (class, configuration); times is the autotuner's
    [class][configuration] table (ms; 0 for a class without tiles, -1 for a
    configuration the plan cannot run)."""
    best, times = plan.autotune(a_data, b_data, out, reps=reps)
    xs, ys = [], []
    for c in range(2):
        if not any(t > 0 for t in times[c]):
            continue
        feats = plan_features(plan, c)
        for q, t in enumerate(times[c]):
            if t > 0:
                xs.append(feats[q])
                ys.append(np.log(t))
    x = np.array(xs, np.float32).reshape(-1, N_FEATURES)
    return x, np.array(ys, np.float32), times


class CostModel:
    """Handle on a GPU-resident cost model (tq_costmodel_t)."""

    def __init__(self, handle=None, seed=0):
        self.h = handle or default_handle()
        ptr = ctypes.c_void_p()
        self.h.check(lib().tq_costmodel_create(self.h.ptr, int(seed), ctypes.byref(ptr)))
        self.ptr = ptr

    def __del__(self):
        try:
            if self.ptr:
                lib().tq_costmodel_destroy(self.ptr)
                self.ptr = None
        except Exception:
            pass

    @staticmethod
    def n_params():
        return int(lib().tq_costmodel_n_params())

    def params(self):
        p = np.zeros(self.n_params(), np.float32)
        self.h.check(lib().tq_costmodel_get_params(self.ptr, _fp(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params()
        self.h.check(lib().tq_costmodel_set_params(self.ptr, _fp(p)))

    def gradient(self, x, y):
        """(gradient, loss) of the mean squared error of the raw network."""
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.zeros(self.n_params(), np.float32)
        loss = ctypes.c_float()
        self.h.check(lib().tq_costmodel_gradient(self.ptr, _fp(x), _fp(y), y.size, _fp(g),
                                                 ctypes.byref(loss)))
        return g, loss.value

    def fit(self, x, y, steps=2000, lr=3e-3):
        """Trains on (x, y) on the GPU; returns the final normalised MSE."""
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        loss = ctypes.c_float()
        self.h.check(lib().tq_costmodel_fit(self.ptr, _fp(x), _fp(y), y.size, int(steps),
                                            float(lr), ctypes.byref(loss)))
        return loss.value

    def predict(self, x, raw=False):
        x = np.ascontiguousarray(np.atleast_2d(x), np.float32)
        y = np.zeros(x.shape[0], np.float32)
        self.h.check(lib().tq_costmodel_predict(self.ptr, _fp(x), x.shape[0], int(bool(raw)),
                                                _fp(y)))
        return y

    def scaling(self):
        xm = np.zeros(N_FEATURES, np.float32)
        xs = np.zeros(N_FEATURES, np.float32)
        y = np.zeros(2, np.float32)
        self.h.check(lib().tq_costmodel_get_scaling(self.ptr, _fp(xm), _fp(xs), _fp(y)))
        return xm, xs, y

    def save(self, path, **meta):
        """Parameters and normalisation (plus any metadata arrays) to .npz."""
        xm, xs, y = self.scaling()
        np.savez(path, params=self.params(), x_mean=xm, x_std=xs, y=y, **meta)

    def load(self, path):
        z = np.load(path)
        self.set_params(z["params"])
        xm, xs, y = (np.ascontiguousarray(z[k], np.float32) for k in ("x_mean", "x_std", "y"))
        self.h.check(lib().tq_costmodel_set_scaling(self.ptr, _fp(xm), _fp(xs), _fp(y)))
        return z

    def choose(self, plan):
        """Configures each tile class of `plan` from the predictions; returns
        the two chosen configurations."""
        best = (ctypes.c_int32 * 2)()
        self.h.check(lib().tq_plan_choose_config(plan.ptr, self.ptr, best))
        return [best[0], best[1]]


# ---- training set and persisted model --------------------------------------

DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                            "costmodel_gfx950.npz")


def _block_diagonal(rng, n_lo, n_hi, s_lo, s_hi):
    """Two block-diagonal matrices with n random sectors of sizes s_lo..s_hi
    (the fused-leg shape of an enlarged-corner product)."""
    from . import workloads as W
    n = int(rng.integers(n_lo, n_hi + 1))
    sizes = rng.integers(s_lo, s_hi + 1, n)
    off = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    begin = np.stack([off, off], 1)
    end = begin + np.stack([sizes - 1, sizes - 1], 1)
    dim = int(sizes.sum())
    a = W.Tensor([dim, dim], begin, end, W.random_data(begin, end, rng))
    b = W.Tensor([dim, dim], begin.copy(), end.copy(), W.random_data(begin, end, rng))
    return a, b, [[1, 0]]


def training_contractions(n, seed):
    """Yields n contractions for the autotuner to time: U(1) rank-4 x rank-4 pairs
    with leg dimensions spanning 8..160, and (every fourth) block-diagonal
    fused-leg products with 2..12 sectors of 16..4096."""
    from . import workloads as W
    rng = np.random.default_rng(seed)
    ranges = [(8, 40), (16, 64), (24, 80), (40, 96), (60, 120), (12, 160)]
    for s in range(n):  # generated one at a time: the large ones hold 100s of MB
        if s % 4 == 3:
            lo = int(rng.choice([16, 64, 256, 512, 1024, 2048]))
            yield _block_diagonal(rng, 2, 12, lo, 2 * lo)
        else:
            lo, hi = ranges[s % len(ranges)]
            yield W.batched(1, seed=int(rng.integers(1 << 30)), dim_lo=lo, dim_hi=hi)[0]


def collect(contractions, device, reps=3, progress=None):
    """Autotuner records (x, log-ms) of FP64 plans of the given contractions
    (any iterable); progress(i) is called after each one."""
    import torch

    from .tensor import BlockSparseTensor
    xs, ys = [], []
    for i, (a, b, legs) in enumerate(contractions):
        A = BlockSparseTensor(a.dims, a.data, a.begin, a.end, device=device)
        B = BlockSparseTensor(b.dims, b.data, b.begin, b.end, device=device)
        plan = A.plan_contract(B, legs)
        c = torch.empty(max(plan.out_n_elem, 1), dtype=torch.float64, device=device)
        x, y, _ = record(plan, A.data, B.data, c, reps=reps)
        xs.append(x)
        ys.append(y)
        plan.destroy()
        if progress:
            progress(i + 1)
    return np.concatenate(xs), np.concatenate(ys)


def train(device, n=160, seed=20261203, steps=2000, lr=3e-3, path=DEFAULT_PATH,
          progress=None):
    """Times n training contractions with the autotuner, fits a model on the
    GPU and saves it (parameters, normalisation, the records) to `path`."""
    import time
    t0 = time.perf_counter()
    x, y = collect(training_contractions(n, seed), device, progress=progress)
    t1 = time.perf_counter()
    model = CostModel(seed=0)
    loss = model.fit(x, y, steps=steps, lr=lr)
    t2 = time.perf_counter()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    model.save(path, x=x, y_log_ms=y, loss=np.float32(loss), seed=np.int64(seed),
               n_contractions=np.int64(n), record_seconds=np.float32(t1 - t0),
               fit_seconds=np.float32(t2 - t1))
    return model, loss, x.shape[0], t1 - t0, t2 - t1


def load_default(handle=None, path=DEFAULT_PATH):
    """The committed model trained on this GPU's autotuner records, or None."""
    if not os.path.exists(path):
        return None
    m = CostModel(handle)
    m.load(path)
    return m


if __name__ == "__main__":
    import sys

    import torch
    out = sys.argv[1] if len(sys.argv) > 1 else DEFAULT_PATH
    dev = torch.device("cuda:0")
    def report(i):
        if i % 20 == 0:
            print("recorded %d contractions" % i, flush=True)

    _, loss, n_rows, rec_s, fit_s = train(dev, path=out, progress=report)
    print("cost model: %d records (%.1f s of autotuning), fit %.2f s, normalised MSE %.4f -> %s"
          % (n_rows, rec_s, fit_s, loss, out), flush=True)
