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

    def choose(self, plan):
        """Configures each tile class of `plan` from the predictions; returns
        the two chosen configurations."""
        best = (ctypes.c_int32 * 2)()
        self.h.check(lib().tq_plan_choose_config(plan.ptr, self.ptr, best))
        return [best[0], best[1]]
