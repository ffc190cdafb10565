"""Batched time evolution (the C ABI's dq_evo plan) and the plain norm-based (m, s) rule.

The plan evolves many Hamiltonians, several states each, to their own times t_j with 2^{s_j} sub-steps of a
degree-m_j Taylor action, all classes of equal-size Hamiltonians in one HIP graph (see diaq_amd/csrc/evolve.hip).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import capi


def plain_rule(norm1: float, t: float, tol: float, cutoff: float = 1.0, m_max: int = 64):
    """The plain norm-based choice of (m, s) for exp(-i t H) psi.

    s follows the reference's scaling rule (s = ceil(log2(|t| ||H||_1 / cutoff)) when above the cutoff,
    src/spME_gpu.cu:251-261); m is the smallest degree whose Taylor remainder bound x^{m+1}/(m+1)! (x = |t| ||H||_1
    / 2^s <= cutoff) meets the per-sub-step share tol / 2^s of the tolerance.
    """
    x_full = abs(t) * norm1 + 1e-12
    s = int(math.ceil(math.log2(x_full / cutoff))) if x_full > cutoff else 0
    x = x_full / 2.0 ** s
    budget = tol / 2.0 ** s
    for m in range(1, m_max + 1):
        if (m + 1) * math.log(x) - math.lgamma(m + 2) <= math.log(budget):
            return m, s
    return m_max, s


def taylor_remainder_bound(x: float, m: int) -> float:
    """Upper bound of the Taylor remainder sum_{k>m} x^k/k! for x <= m + 2 (first term times a geometric tail)."""
    first = math.exp((m + 1) * math.log(max(x, 1e-300)) - math.lgamma(m + 2))
    q = x / (m + 2)
    return first / (1 - q) if q < 1 else float("inf")


class BatchedEvolution:
    """One dq_evo plan. ``mats`` are DeviceMatrix objects (or raw dq_devmat handles) of square Hamiltonians."""

    def __init__(self, mats, n_states: int, value_precision: str = "fp32"):
        L = capi.lib()
        handles = [m.handle if hasattr(m, "handle") else int(m) for m in mats]
        arr = (ctypes.c_void_p * len(handles))(*handles)
        self._h = ctypes.c_void_p()
        prec = capi.DQ_F32 if value_precision == "fp32" else capi.DQ_F64
        capi.check(L.dq_evo_create(len(handles), arr, n_states, prec, ctypes.byref(self._h)))
        self.n_jobs = len(handles)
        self.n_states = n_states
        self.sizes = [m.shape[0] if hasattr(m, "shape") else None for m in mats]

    def close(self):
        if self._h:
            capi.lib().dq_evo_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_classes(self) -> int:
        n = ctypes.c_int()
        capi.check(capi.lib().dq_evo_info(self._h, ctypes.byref(n), None, None))
        return n.value

    def set_states(self, job: int, states: np.ndarray):
        """states: (n_states, N) complex array."""
        st = np.ascontiguousarray(states, dtype=np.complex128)
        re = np.ascontiguousarray(st.real)
        im = np.ascontiguousarray(st.imag)
        capi.check(capi.lib().dq_evo_set_states(self._h, job, capi.ptr(re), capi.ptr(im)))

    def get_states(self, job: int, N: int) -> np.ndarray:
        re = np.empty(self.n_states * N)
        im = np.empty(self.n_states * N)
        capi.check(capi.lib().dq_evo_get_states(self._h, job, capi.ptr(re), capi.ptr(im)))
        return (re + 1j * im).reshape(self.n_states, N)

    def set_schedule(self, t, m, s):
        t = np.ascontiguousarray(t, np.float64)
        m = np.ascontiguousarray(m, np.int32)
        s = np.ascontiguousarray(s, np.int32)
        capi.check(capi.lib().dq_evo_set_schedule(self._h, capi.ptr(t), capi.ptr(m, ctypes.c_int32),
                                                  capi.ptr(s, ctypes.c_int32)))

    def tune(self):
        """Per class: (rows per lane, XCD-contiguous tile order, kernel form) chosen by timing. Form: 0 a lane per
        state, 1 the same with the LDS window, 2 four states per lane, 3 and 4 four states per lane with the loads of
        2 or 4 diagonals issued together."""
        rpt = (ctypes.c_int * self.n_classes)()
        xcd = (ctypes.c_int * self.n_classes)()
        lds = (ctypes.c_int * self.n_classes)()
        capi.check(capi.lib().dq_evo_tune(self._h, rpt, xcd, lds))
        return list(zip(list(rpt), list(xcd), list(lds)))

    def probe(self) -> np.ndarray:
        est = np.empty(self.n_jobs)
        capi.check(capi.lib().dq_evo_probe(self._h, capi.ptr(est)))
        return est

    def run(self, use_graph: bool = True, renormalize: bool = True, stream: int = 0) -> float:
        ms = ctypes.c_float()
        capi.check(capi.lib().dq_evo_run(self._h, int(use_graph), int(renormalize), stream, ctypes.byref(ms)))
        return ms.value
