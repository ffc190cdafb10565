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
        L.oracle_product_tasks.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _i32p, _i64p, _dp, _dp,
                                           _i32p, _i64p, _dp, _dp, ctypes.c_int, _i32p, _i64p, _i32p, _i32p, _i64p,
                                           _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oracle_product_tasks.restype = None
        L.oracle_taylor_step_tasks.argtypes = [ctypes.c_int64, ctypes.c_int, _i32p, _i64p, _dp, _dp, _dp, _dp,
                                               ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int]
        L.oracle_taylor_step_tasks.restype = ctypes.c_int
        L.oracle_kron_apply_tasks.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _i64p, _dp, _dp,
                                              ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int]
        L.oracle_kron_apply_tasks.restype = None
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


def product_tasks(rows_a, cols_a, cols_b, A, B, dry, n_threads=0, res_block=4, elem_block=2048):
    """The reference CPU SpGEMM restated (diaq_oracle.c: oracle_product_tasks). A, B: (offsets int32, starts int64,
    re, im) as from packed(); dry: (res_offsets, pair_start, slot_a, slot_b) of the structural dry-run. Returns the
    result as (res_offsets int32, starts int64, re, im), packed."""
    res, ps, sa, sb = (np.ascontiguousarray(dry[0], np.int32), np.ascontiguousarray(dry[1], np.int64),
                       np.ascontiguousarray(dry[2], np.int32), np.ascontiguousarray(dry[3], np.int32))
    lens = np.array([max(0, min(rows_a, cols_b, rows_a + o if o < 0 else cols_b - o)) for o in res], np.int64)
    st = np.zeros(len(res), np.int64)
    if len(res):
        st[1:] = np.cumsum(lens)[:-1]
    total = int(lens.sum())
    cr = np.zeros(total)
    ci = np.zeros(total)
    lib().oracle_product_tasks(rows_a, cols_a, cols_b, _p(A[0], _i32p), _p(A[1], _i64p), _p(A[2], _dp),
                               _p(A[3], _dp), _p(B[0], _i32p), _p(B[1], _i64p), _p(B[2], _dp), _p(B[3], _dp),
                               len(res), _p(res, _i32p), _p(ps, _i64p), _p(sa, _i32p), _p(sb, _i32p), _p(st, _i64p),
                               _p(cr, _dp), _p(ci, _dp), res_block, elem_block, n_threads)
    return res, st, cr, ci


def taylor_step_tasks(rows, H, psi, dt, m, n_threads=0):
    """One fixed-degree Taylor step with the reference CPU SpMV (diaq_oracle.c: oracle_taylor_step_tasks)."""
    pr = np.ascontiguousarray(np.real(psi), np.float64)
    pi = np.ascontiguousarray(np.imag(psi), np.float64)
    outr, outi = np.empty(rows), np.empty(rows)
    scratch = np.empty(4 * rows)
    rc = lib().oracle_taylor_step_tasks(rows, len(H[0]), _p(H[0], _i32p), _p(H[1], _i64p), _p(H[2], _dp),
                                        _p(H[3], _dp), _p(pr, _dp), _p(pi, _dp), dt, m, _p(outr, _dp), _p(outi, _dp),
                                        _p(scratch, _dp), n_threads)
    if rc != 0:
        raise MemoryError("oracle_taylor_step_tasks: allocation failed")
    return outr + 1j * outi


def kron_apply_tasks(before, G, after, xr, xi, yr, yi, n_threads=0):
    """y += (I (x) G (x) I) x in place on split float64 arrays with the reference's CPU loop nest (diaq_oracle.c:
    oracle_kron_apply_tasks). G is (g, offsets, starts, re, im) as from ``packed``."""
    g, offs, starts, gre, gim = G
    lib().oracle_kron_apply_tasks(before, g, len(offs), _p(offs, _i32p), _p(starts, _i64p), _p(gre, _dp),
                                  _p(gim, _dp), after, _p(xr, _dp), _p(xi, _dp), _p(yr, _dp), _p(yi, _dp), n_threads)
