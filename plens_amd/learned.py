"""Learned troubled-cell indicator: training data from the plain GPU path, GPU training through the
hand-written kernels of plens_amd/csrc/mlp.hip, and switching a solver handle to the network.

Features of a cell are the (N+1)^3 Legendre modes of rho*p divided by the cell mean (mode 0),
zero-padded to 125; the label is the plain blender value of BlenderCalculator::get_blender
(dgsem/blender_calculator.cc:123-187) before the wall override and the diffusion, which the
solver keeps applying to either value (dgsem/plens.cc:2905-2960).
"""
import ctypes as C

import numpy as np

from . import _lib as L

N_IN = 125
_fp = C.POINTER(C.c_float)


def _f(a):
    return a.ctypes.data_as(_fp)


def init_params(seed, n_in=125, h1=64, h2=64):
    """He-uniform initialisation, packed W1, b1, W2, b2, w3, b3 (FP32)."""
    rng = np.random.default_rng(seed)
    def u(fan_in, shape):
        lim = np.sqrt(6.0 / fan_in)
        return rng.uniform(-lim, lim, shape).astype(np.float32)
    parts = [u(n_in, (h1, n_in)), np.zeros(h1, np.float32), u(h1, (h2, h1)), np.zeros(h2, np.float32),
             u(h2, (h2,)) * np.float32(0.1), np.zeros(1, np.float32)]
    return np.concatenate([p.reshape(-1) for p in parts]).astype(np.float32)


def unpack(P):
    P = np.asarray(P, np.float32)
    o = 0
    out = {}
    for k, shape in (("W1", (64, 125)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("w3", (64,)), ("b3", (1,))):
        n = int(np.prod(shape))
        out[k] = P[o:o + n].reshape(shape).copy()
        o += n
    return out


class MLP:
    def __init__(self, params=None, seed=0):
        self.lib = L.load()
        P = init_params(seed) if params is None else np.ascontiguousarray(params, np.float32)
        assert P.size == self.lib.plens_mlp_n_params()
        self._h = C.c_void_p()
        L.check(self.lib.plens_mlp_create(_f(P), C.byref(self._h)), "plens_mlp_create")

    def params(self):
        P = np.empty(self.lib.plens_mlp_n_params(), np.float32)
        L.check(self.lib.plens_mlp_get_params(self._h, _f(P)), "get_params")
        return P

    def device_params(self):
        return self.lib.plens_mlp_device_bf16_params(self._h)

    def predict(self, X):
        X = np.ascontiguousarray(X, np.float32)
        y = np.empty(X.shape[0], np.float32)
        L.check(self.lib.plens_mlp_predict(self._h, _f(X), X.shape[0], _f(y)), "predict")
        return y

    def grad(self, X, t):
        X = np.ascontiguousarray(X, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.empty(self.lib.plens_mlp_n_params(), np.float32)
        loss = C.c_float()
        L.check(self.lib.plens_mlp_grad(self._h, _f(X), _f(t), X.shape[0], _f(g), C.byref(loss)), "grad")
        return loss.value, g

    def step(self, X, t, lr=1e-3):
        X = np.ascontiguousarray(X, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        L.check(self.lib.plens_mlp_step(self._h, _f(X), _f(t), X.shape[0], lr), "step")

    def train(self, X, t, epochs=5, batch=4096, lr=1e-3, seed=0):
        X = np.ascontiguousarray(X, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        hist = np.zeros(epochs, np.float32)
        L.check(self.lib.plens_mlp_train(self._h, _f(X), _f(t), X.shape[0], epochs, batch, lr, seed, _f(hist)),
                "train")
        return hist

    def close(self):
        if self._h:
            self.lib.plens_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def samples(solver):
    """(features [n_cells][125] float32, plain blender values [n_cells]) of the solver's state."""
    lib = L.load()
    X = np.empty((solver.n_cells, N_IN), np.float32)
    a = np.empty(solver.n_cells)
    solver._check(lib.plens_gpu_indicator_samples(solver._h, _f(X), a.ctypes.data_as(C.POINTER(C.c_double))),
                  "indicator_samples")
    return X, a


def collect(case, steps, every=10):
    """Run the plain GPU path on a case and sample every `every` steps (including the start)."""
    from .solver import GpuPlens
    g = GpuPlens(case)
    Xs, ys = [], []
    done = 0
    while True:
        X, a = samples(g)
        Xs.append(X)
        ys.append(a)
        if done >= steps:
            break
        k = min(every, steps - done)
        g.run(k, case.courant)
        done += k
    g.close()
    return np.concatenate(Xs), np.concatenate(ys).astype(np.float32)


def use_network(solver, mlp, tolerance=0.05, floor=0.02):
    """Switch a solver handle to the network (mlp=None: back to the plain indicator). Network values
    below `floor` count as 0, so that small positive outputs in smooth cells do not turn on the
    subcell blending; |alpha_mlp - alpha_plain| > tolerance falls back to the plain value."""
    ptr = mlp.device_params() if mlp is not None else None
    solver._check(L.load().plens_gpu_set_learned_indicator(solver._h, ptr, tolerance, floor),
                  "set_learned_indicator")


def stats(solver, reset=True):
    ev, fb = C.c_int64(), C.c_int64()
    solver._check(L.load().plens_gpu_indicator_stats(solver._h, C.byref(ev), C.byref(fb), int(reset)),
                  "indicator_stats")
    return ev.value, fb.value
