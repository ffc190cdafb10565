"""ctypes access to the C oracle (liboracle.so) -- test infrastructure only.

Build with ``make -C oracle`` (``__graft_entry__.build()`` does it). The functions take numpy arrays in the layout
described at the top of ``diaq_oracle.c``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"C oracle not built: {path} (run make -C oracle)")
        L = ctypes.CDLL(path)
        L.oracle_spmv_rows.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int, _i32p, _i64p, _dp, _dp, _dp,
                                       _dp, _dp, _dp]
        L.oracle_spmv_rows.restype = None
        L.oracle_spmv_tasks.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int, _i32p, _i64p, _dp, _dp, _dp,
                                        _dp, _dp, _dp, ctypes.c_int, ctypes.c_int]
        L.oracle_spmv_tasks.restype = ctypes.c_int
        L.oracle_block_size.argtypes = []
        L.oracle_block_size.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t)


def packed(offsets, vals):
    """Concatenate per-diagonal complex arrays into (starts int64, re, im) with diagonals back to back."""
    lens = [len(v) for v in vals]
    starts = np.zeros(len(vals), np.int64)
    if lens:
        starts[1:] = np.cumsum(lens)[:-1]
    cat = np.concatenate(vals) if vals else np.zeros(0, np.complex128)
    return (np.asarray(offsets, np.int32), starts, np.ascontiguousarray(cat.real), np.ascontiguousarray(cat.imag))


def spmv_rows(rows, cols, offsets, starts, re, im, x):
    """y = A x with the GPU row-kernel semantics (diaq_oracle.c: oracle_spmv_rows)."""
    xr = np.ascontiguousarray(np.real(x), np.float64)
    xi = np.ascontiguousarray(np.imag(x), np.float64)
    yr = np.empty(rows, np.float64)
    yi = np.empty(rows, np.float64)
    lib().oracle_spmv_rows(rows, cols, len(offsets), _p(offsets, _i32p), _p(starts, _i64p), _p(re, _dp),
                           _p(im, _dp), _p(xr, _dp), _p(xi, _dp), _p(yr, _dp), _p(yi, _dp))
    return yr + 1j * yi


def spmv_tasks_split(rows, cols, offsets, starts, re, im, xr, xi, yr, yi, block_size=0, n_threads=0):
    """The reference CPU SpMV restated (diaq_oracle.c: oracle_spmv_tasks) on split real/imag arrays."""
    rc = lib().oracle_spmv_tasks(rows, cols, len(offsets), _p(offsets, _i32p), _p(starts, _i64p), _p(re, _dp),
                                 _p(im, _dp), _p(xr, _dp), _p(xi, _dp), _p(yr, _dp), _p(yi, _dp), block_size,
                                 n_threads)
    if rc != 0:
        raise MemoryError("oracle_spmv_tasks: allocation failed")


def block_size():
    return lib().oracle_block_size()
