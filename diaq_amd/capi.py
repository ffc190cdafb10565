"""ctypes view of the C ABI in include/diaq_hip.h (libdiaq_hip.so).

Used by the tests, the benchmark and the tuning helpers to call the library exactly as a foreign caller would. The
loader fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdiaq_hip.so")

DQ_OK, DQ_ERR_ARG, DQ_ERR_HIP, DQ_ERR_ALLOC, DQ_ERR_STATE = 0, 1, 2, 3, 4
DQ_F64, DQ_F32 = 0, 1

c_void_p, c_int, c_int32, c_int64, c_uint32, c_uint64, c_double, c_char_p, c_uintptr = (
    ctypes.c_void_p, ctypes.c_int, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double,
    ctypes.c_char_p, ctypes.c_size_t)
P = ctypes.POINTER


class DiaViability(ctypes.Structure):
    _fields_ = [("viable", c_int), ("storage_ok", c_int), ("diagonals_ok", c_int), ("dia_bytes", c_double),
                ("csr_bytes", c_double), ("density", c_double), ("diagonal_budget", c_int64),
                ("reason", ctypes.c_char * 160)]


# name -> (restype, argtypes); every function declared in include/diaq_hip.h
SIGNATURES = {
    "dq_last_error": (c_char_p, []),
    "dq_version": (c_char_p, []),
    "dq_devmat_upload": (c_int, [c_uint32, c_uint32, c_int32, P(c_int32), P(c_int64), P(c_double), P(c_double), c_int,
                                 P(c_void_p)]),
    "dq_devmat_zeros": (c_int, [c_uint32, c_uint32, c_int32, P(c_int32), c_int, P(c_void_p)]),
    "dq_devmat_from_pauli_sum": (c_int, [c_int, c_int, P(ctypes.c_uint64), P(ctypes.c_uint64), P(c_double),
                                         P(c_double), c_int, c_uintptr, P(c_void_p)]),
    "dq_devmat_from_dense_device": (c_int, [c_uint32, c_uint32, c_void_p, c_int64, c_int, c_uintptr, P(c_void_p)]),
    "dq_devmat_to_dense_device": (c_int, [c_void_p, c_void_p, c_int64, c_uintptr]),
    "dq_devmat_set_packed_device": (c_int, [c_void_p, c_void_p, c_void_p, c_uintptr]),
    "dq_kron_apply_device": (c_int, [c_void_p, c_uint32, c_uint32, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                     c_uintptr]),
    "dq_kron_apply_host": (c_int, [c_void_p, c_uint32, c_uint32, P(c_double), P(c_double), P(c_double),
                                   P(c_double)]),
    "dq_gate_plan_create": (c_int, [c_int, c_int, P(c_void_p), P(c_uint32), P(c_void_p)]),
    "dq_gate_plan_passes": (c_int, [c_void_p, P(c_int)]),
    "dq_gate_plan_run": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_uintptr]),
    "dq_gate_plan_destroy": (None, [c_void_p]),
    "dq_devmat_copy": (c_int, [c_void_p, P(c_void_p)]),
    "dq_devmat_convert": (c_int, [c_void_p, c_int, P(c_void_p)]),
    "dq_devmat_download": (c_int, [c_void_p, P(c_int64), P(c_double), P(c_double)]),
    "dq_devmat_get_diagonal": (c_int, [c_void_p, c_int32, P(c_double), P(c_double)]),
    "dq_devmat_destroy": (None, [c_void_p]),
    "dq_devmat_info": (c_int, [c_void_p, P(c_uint32), P(c_uint32), P(c_int32), P(c_int), P(c_int64)]),
    "dq_devmat_offsets": (c_int, [c_void_p, P(c_int32)]),
    "dq_spmv_device": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_uintptr]),
    "dq_spmv_host": (c_int, [c_void_p, P(c_double), P(c_double), P(c_double), P(c_double)]),
    "dq_spmv_ntimes_device": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                                      c_uintptr]),
    "dq_spmv_ntimes_device_noalloc": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p,
                                              c_void_p, c_uintptr]),
    "dq_spmv_ntimes_host": (c_int, [c_void_p, P(c_double), P(c_double), P(c_double), P(c_double), c_int]),
    "dq_spmv_device_banded": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_uintptr]),
    "dq_release_cached_memory": (c_int, []),
    "dq_set_tuning": (c_int, [c_char_p, c_int]),
    "dq_spmv_tune": (c_int, [c_void_p, c_int, c_uintptr, P(ctypes.c_float)]),
    "dq_spmv_tuned_config": (c_int, [c_void_p, P(c_int), P(c_int), P(c_int), P(c_int), P(c_int)]),
    "dq_product_dryrun": (c_int, [c_uint32, c_uint32, c_int32, P(c_int32), c_int32, P(c_int32), P(c_int32),
                                  P(c_int64), P(c_int32), P(c_int64), P(c_int32), P(c_int32)]),
    "dq_product_device": (c_int, [c_void_p, c_void_p, P(c_void_p)]),
    "dq_prune": (c_int, [c_void_p, P(c_void_p)]),
    "dq_estimate_1norm": (c_int, [c_void_p, P(c_double)]),
    "dq_frobenius_norm": (c_int, [c_void_p, P(c_double)]),
    "dq_frobenius_diff": (c_int, [c_void_p, c_void_p, P(c_double)]),
    "dq_trace": (c_int, [c_void_p, P(c_double), P(c_double)]),
    "dq_pow": (c_int, [c_void_p, c_int64, P(c_void_p)]),
    "dq_expm_taylor_scaled": (c_int, [c_void_p, c_int, c_double, c_double, P(c_void_p)]),
    "dq_expm_negative_taylor_scaled": (c_int, [c_void_p, c_double, c_int, c_double, c_double, P(c_void_p), P(c_int),
                                               P(c_int)]),
    "dq_hamsim_host": (c_int, [c_void_p, P(c_double), P(c_double), c_double, c_uint64, c_int, P(c_double),
                               P(c_double)]),
    "dq_hamsim_device": (c_int, [c_void_p, c_void_p, c_void_p, c_double, c_uint64, c_int, c_uintptr]),
    "dq_tensordot": (c_int, [c_void_p, c_int, P(c_uint32), c_void_p, c_int, P(c_uint32), c_int, P(c_uint32),
                             P(c_uint32), P(c_void_p), P(c_uint32), P(c_int)]),
    "dq_hamsim_device_f64state": (c_int, [c_void_p, c_void_p, c_void_p, c_double, c_uint64, c_int, c_uintptr]),
    "dq_spmm_dia_device": (c_int, [c_uintptr, c_uintptr, c_int, c_int64, c_int64, c_uintptr, c_int64, c_uintptr, c_int,
                                   c_uintptr]),
    "dq_dia_viability": (c_int, [c_int64, c_int64, c_int64, c_int64, c_int64, c_int, c_int, P(DiaViability)]),
    "dq_dia_viability_mi355x": (c_int, [c_int64, c_int64, c_int64, c_int64, c_int64, c_int, c_int, P(DiaViability)]),
    "dq_evo_create": (c_int, [c_int, P(c_void_p), c_int, c_int, P(c_void_p)]),
    "dq_evo_destroy": (None, [c_void_p]),
    "dq_evo_info": (c_int, [c_void_p, P(c_int), P(c_int), P(c_int)]),
    "dq_evo_set_states": (c_int, [c_void_p, c_int, P(c_double), P(c_double)]),
    "dq_evo_get_states": (c_int, [c_void_p, c_int, P(c_double), P(c_double)]),
    "dq_evo_set_schedule": (c_int, [c_void_p, P(c_double), P(c_int32), P(c_int32)]),
    "dq_evo_tune": (c_int, [c_void_p, P(c_int), P(c_int), P(c_int)]),
    "dq_evo_run": (c_int, [c_void_p, c_int, c_int, c_uintptr, P(ctypes.c_float)]),
    "dq_evo_probe": (c_int, [c_void_p, P(c_double)]),
    "dq_mlp_create": (c_int, [c_int, c_int, c_int, c_int, c_uint64, P(c_void_p)]),
    "dq_mlp_destroy": (None, [c_void_p]),
    "dq_mlp_num_params": (c_int64, [c_void_p]),
    "dq_mlp_get_params": (c_int, [c_void_p, P(ctypes.c_float)]),
    "dq_mlp_set_params": (c_int, [c_void_p, P(ctypes.c_float)]),
    "dq_mlp_forward": (c_int, [c_void_p, c_int, c_void_p, P(ctypes.c_float)]),
    "dq_mlp_loss_grad": (c_int, [c_void_p, c_int, c_void_p, c_void_p, P(ctypes.c_float), P(ctypes.c_float)]),
    "dq_mlp_train_step": (c_int, [c_void_p, c_int, c_void_p, c_void_p, ctypes.c_float, P(ctypes.c_float)]),
    "dq_mlp_predict": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_uintptr]),
}

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        try:  # share torch's HIP runtime when torch is present (see diaq_amd/__init__.py)
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built (python diaq_amd/build.py)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


class DqError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def check(rc):
    if rc != DQ_OK:
        raise DqError(rc, lib().dq_last_error().decode())
    return rc


def ptr(a, t=ctypes.c_double):
    """ctypes pointer to a contiguous numpy array."""
    return a.ctypes.data_as(P(t))
