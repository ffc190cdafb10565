"""The learned (m, s) selector: an FP32 MLP (5 -> 64 -> 64 -> 195) that proposes the Taylor degree m in 4..18 and the
squaring count s in 0..12 for exp(-i t H) psi, run by the library's HIP kernels (dq_mlp_*).

Features: log ||H t||_1, number of diagonals, log2 bandwidth (largest |offset| + 1), log2 N, log10 tol. Labels (for
training): the cheapest (m, s) -- cost m 2^s SpMVs -- whose Taylor action [T_m(Y)]^{2^s}, Y = -i t H / 2^s, is
within tol of exp(-i t H) on every eigenvalue of H, i.e. max over |x| <= rho t of |T_m(-i x / 2^s)^{2^s} - e^{-ix}|
<= tol, with rho a bound on the spectral radius. Proposals are checked a posteriori (dq_evo_probe) and replaced by the
plain norm-based rule when the estimate misses tol.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

from . import capi

M_VALUES = list(range(4, 19))
S_VALUES = list(range(0, 13))
N_CLASSES = len(M_VALUES) * len(S_VALUES)
DIMS = (5, 64, 64, N_CLASSES)
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "selector_weights.npz")


def cls_of(m: int, s: int) -> int:
    return (m - M_VALUES[0]) * len(S_VALUES) + (s - S_VALUES[0])


def ms_of(c: int):
    return M_VALUES[c // len(S_VALUES)], S_VALUES[c % len(S_VALUES)]


def features(norm1: float, t: float, n_diags: int, bandwidth: int, N: int, tol: float) -> np.ndarray:
    return np.array([math.log(max(norm1 * abs(t), 1e-300)), float(n_diags), math.log2(max(bandwidth, 1) + 1.0),
                     math.log2(max(N, 1)), math.log10(tol)], np.float64)


def taylor_error(x, m: int, s: int):
    """|T_m(-i x / 2^s)^{2^s} - e^{-i x}| for real x (array): the error of the scheduled Taylor action on an
    eigencomponent with eigenvalue x / t."""
    x = np.asarray(x, np.float64)
    with np.errstate(over="ignore", invalid="ignore"):
        return _taylor_error(x, m, s)


def _taylor_error(x, m, s):
    y = -1j * x / 2.0 ** s
    term = np.ones_like(y)
    T = np.ones_like(y)
    for k in range(1, m + 1):
        term = term * y / k
        T = T + term
    # T^(2^s) by s squarings
    for _ in range(s):
        T = T * T
    e = np.abs(T - np.exp(-1j * x))
    return np.where(np.isfinite(e), e, np.inf)


def label(rho_t: float, tol: float, n_grid: int = 48):
    """The cheapest (m, s) meeting tol for every |x| <= rho_t (cost m 2^s, ties: fewer sub-steps); None if none."""
    xs = np.linspace(0.0, rho_t, n_grid)
    best = None
    for s in S_VALUES:
        for m in M_VALUES:
            cost = m * 2 ** s
            if best is not None and cost > best[0]:
                continue
            if np.max(taylor_error(xs, m, s)) <= tol:
                if best is None or (cost, s) < (best[0], best[2]):
                    best = (cost, m, s)
                break  # larger m at this s only costs more
    return None if best is None else (best[1], best[2])


_TABLE = None


def _error_table():
    """Cumulative maximum of taylor_error over a fine log grid of x, per (m, s): E[c, i] = max_{x <= X_i} err."""
    global _TABLE
    if _TABLE is None:
        X = np.concatenate([[0.0], np.geomspace(1e-4, 3e4, 3000)])
        E = np.empty((N_CLASSES, X.size))
        for m in M_VALUES:
            for s in S_VALUES:
                E[cls_of(m, s)] = np.maximum.accumulate(taylor_error(X, m, s))
        _TABLE = (X, E)
    return _TABLE


def labels_fast(rho_t, tol):
    """Vectorised ``label`` on the precomputed error table: class index of the cheapest feasible (m, s) per sample
    (cost m 2^s, ties: fewer sub-steps), -1 where no (m, s) of the grid meets tol."""
    X, E = _error_table()
    rho_t = np.asarray(rho_t, np.float64)
    tol = np.asarray(tol, np.float64)
    idx = np.clip(np.searchsorted(X, rho_t, side="left"), 0, X.size - 1)  # first grid point >= rho_t
    err = E[:, idx]  # (classes, samples)
    order = sorted(range(N_CLASSES), key=lambda c: (ms_of(c)[0] * 2 ** ms_of(c)[1], ms_of(c)[1]))
    out = np.full(rho_t.shape, -1, np.int64)
    for c in reversed(order):  # cheapest last, so it wins
        ok = err[c] <= tol
        out[ok] = c
    return out


class Selector:
    """The trained network on the GPU plus the feature standardisation it was trained with."""

    def __init__(self, params=None, mean=None, std=None, seed: int = 0):
        L = capi.lib()
        self._h = ctypes.c_void_p()
        capi.check(L.dq_mlp_create(*DIMS, seed, ctypes.byref(self._h)))
        self.mean = np.zeros(DIMS[0]) if mean is None else np.asarray(mean, np.float64)
        self.std = np.ones(DIMS[0]) if std is None else np.asarray(std, np.float64)
        if params is not None:
            self.set_params(params)

    @classmethod
    def load(cls, path: str = WEIGHTS):
        z = np.load(path)
        return cls(params=z["params"], mean=z["mean"], std=z["std"])

    def save(self, path: str = WEIGHTS, **extra):
        np.savez(path, params=self.get_params(), mean=self.mean, std=self.std, **extra)

    def close(self):
        if self._h:
            capi.lib().dq_mlp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_params(self) -> int:
        return int(capi.lib().dq_mlp_num_params(self._h))

    def get_params(self) -> np.ndarray:
        p = np.empty(self.n_params, np.float32)
        capi.check(capi.lib().dq_mlp_get_params(self._h, capi.ptr(p, ctypes.c_float)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        capi.check(capi.lib().dq_mlp_set_params(self._h, capi.ptr(p, ctypes.c_float)))

    def _x(self, feats, torch):
        z = (np.asarray(feats, np.float64) - self.mean) / self.std
        return torch.as_tensor(z.astype(np.float32), device="cuda").contiguous()

    def train_step(self, x_dev, y_dev, lr: float) -> float:
        loss = ctypes.c_float()
        capi.check(capi.lib().dq_mlp_train_step(self._h, int(y_dev.numel()), x_dev.data_ptr(), y_dev.data_ptr(),
                                                lr, ctypes.byref(loss)))
        return loss.value

    def loss_grad(self, x_dev, y_dev):
        loss = ctypes.c_float()
        g = np.empty(self.n_params, np.float32)
        capi.check(capi.lib().dq_mlp_loss_grad(self._h, int(y_dev.numel()), x_dev.data_ptr(), y_dev.data_ptr(),
                                               ctypes.byref(loss), capi.ptr(g, ctypes.c_float)))
        return loss.value, g

    def predict_classes(self, feats) -> np.ndarray:
        import torch
        x = self._x(feats, torch)
        out = torch.empty(x.shape[0], dtype=torch.int32, device="cuda")
        capi.check(capi.lib().dq_mlp_predict(self._h, x.shape[0], x.data_ptr(), None, out.data_ptr(),
                                             torch.cuda.current_stream().cuda_stream))
        return out.cpu().numpy()

    def predict(self, feats):
        """(m, s) for each feature row."""
        return [ms_of(int(c)) for c in self.predict_classes(np.atleast_2d(feats))]
