"""Two of the reference's own path scripts, run through the HIP library with the same inputs and bounds.

- python_tests/test_for_disassembly.py:1-36: SpGEMM of two 1024 x 1024 banded matrices (15 and 8 diagonals, numpy's
  legacy generator seeded 42), against the dense product at rtol = atol = 1e-12.
- python_tests/test_expm_negative.py:29-163: expm_negative at t = 1.8 on its 15 matrices (diagonal, bidiagonal,
  tridiagonal, non-Hermitian 4 x 4 / 6 x 6 cases, the Hadamard gate, dense random 64 x 64 and 128 x 128, and banded
  10-diagonal matrices of 2^10 .. 2^15 rows drawn with seeds 10 .. 15), against scipy's dense expm at the reference's
  rtol = atol = 1e-4. The reference draws the two dense random matrices from numpy's unseeded global generator; here
  they come from np.random.RandomState(0). Beyond the reference's bound, every case is also checked against the
  oracle: offsets exactly, values at 1e-12 (max-abs error over max-abs value) where the dense oracle fits (<= 2^11
  rows); above that, the Taylor term count and squaring count equal the oracle's, the offsets equal the structural
  chain's, and the product of U with random vectors matches scipy's expm_multiply.
"""
import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as O  # noqa: E402

TOL = 1e-12


@pytest.fixture(scope="module")
def d():
    import diaq_amd  # noqa: F401
    import diaq_gpu
    return diaq_gpu


def rel(a, b):
    b = np.asarray(b)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(np.asarray(a) - b)) / (den if den > 0 else 1.0))


# ---------------------------------------------------------------------------------------------------------------------
# python_tests/test_for_disassembly.py
def test_reference_spgemm_disassembly_script(d):
    size = 1024
    rs = np.random.RandomState(42)
    A = np.zeros((size, size), np.complex128)
    B = np.zeros((size, size), np.complex128)
    offs_a = [0, 4, 8, 16, 32, 64, 128, 256, -4, -8, -16, -32, -64, -128, -256]
    offs_b = [1, 16, 32, 512, -1, -16, -32, -512]
    for off in offs_a:
        n = size - abs(off)
        A += np.diag(rs.rand(n) + 1j * rs.rand(n), k=off)
    for off in offs_b:
        n = size - abs(off)
        B += np.diag(rs.rand(n) + 1j * rs.rand(n), k=off)
    C = A @ B
    got = d.product(d.from_numpy(A), d.from_numpy(B))
    np.testing.assert_allclose(d.to_numpy(got), C, rtol=1e-12, atol=1e-12)
    # the offsets of the pruned product are the nonzero diagonals of the dense product (from_numpy(C) in the script)
    assert list(got.dIndices) == list(d.from_numpy(C).dIndices)
    # and the same product on device matrices (no prune): the structural dry-run's offsets
    Dg = d.product(d.to_device(d.from_numpy(A)), d.to_device(d.from_numpy(B)))
    roffs, _ = O.product_dryrun(sorted(offs_a), sorted(offs_b), size, size)
    assert list(Dg.diagonals) == roffs
    assert rel(d.to_numpy(Dg.to_host()), C) < TOL


# ---------------------------------------------------------------------------------------------------------------------
# python_tests/test_expm_negative.py
def band_matrix(n, density=0.01, num_diagonals=10, seed=None):
    """generate_sparse_band_matrix (python_tests/test_expm_negative.py:29-45) with its own seeded legacy generator."""
    rs = np.random.RandomState(seed)
    H = np.zeros((n, n), np.float64)
    half = num_diagonals // 2
    for off in range(-half, half + 1):
        if rs.rand() < density:
            continue
        vals = rs.rand(n - abs(off))
        if off >= 0:
            H[np.arange(n - off), np.arange(off, n)] = vals
        else:
            H[np.arange(-off, n), np.arange(n + off)] = vals
    return H


def small_cases():
    rs = np.random.RandomState(0)
    return [
        np.diag([1.0, 2.0, 3.0, 4.0]),
        np.array([[1, 5, 0, 0], [0, 2, 6, 0], [0, 0, 3, 7], [0, 0, 0, 4]], float),
        np.array([[1, 5, 0, 0], [8, 2, 6, 0], [0, 9, 3, 7], [0, 0, 10, 4]], float),
        np.array([[1, 0, 5, 0], [0, 2, 0, 6], [7, 0, 3, 0], [0, 8, 0, 4]], float),
        np.array([[0.707107, 0.707107], [0.707107, -0.707107]]),
        np.array([[10, 0, 0, 0, 21, 0], [0, 11, 0, 0, 0, 22], [19, 0, 12, 0, 0, 0], [0, 18, 0, 13, 0, 0],
                  [0, 0, 17, 0, 14, 0], [20, 0, 0, 16, 0, 15]], float),
        np.arange(10, 46, dtype=float).reshape(6, 6),
        rs.rand(64, 64),
        rs.rand(128, 128),
    ]


T_REF = 1.8


@pytest.mark.parametrize("case", range(9))
def test_reference_expm_negative_small_cases(d, case):
    import scipy.linalg
    H = small_cases()[case]
    U = d.expm_negative(d.from_numpy(H), T_REF)
    got = d.to_numpy(U)
    np.testing.assert_allclose(got, scipy.linalg.expm(-1j * H * T_REF), rtol=1e-4, atol=1e-4)
    M = O.from_dense(H.astype(complex))
    if H.shape[0] < 64:
        ref = O.expm_negative(M, T_REF)
        roffs, Uref = ref.offsets, O.to_dense(ref)
    else:  # the dense-value twin of the same steps (DIA products in numpy are slow at 127-255 diagonals)
        Uref, roffs, _, _ = O.expm_negative_dense(H.astype(complex), M.offsets, T_REF)
    assert list(U.dIndices) == roffs
    assert rel(got, Uref) < TOL


def expm_negative_capi(D, t):
    from diaq_amd import capi
    out = ctypes.c_void_p()
    terms, s = ctypes.c_int(), ctypes.c_int()
    capi.check(capi.lib().dq_expm_negative_taylor_scaled(D.handle, t, 100, 1.0, 1e-12, ctypes.byref(out),
                                                         ctypes.byref(terms), ctypes.byref(s)))
    return out, terms.value, s.value


def devmat_offsets(h):
    from diaq_amd import capi
    rows, cols, nd, prec = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_int32(), ctypes.c_int()
    stored = ctypes.c_int64()
    capi.check(capi.lib().dq_devmat_info(h, ctypes.byref(rows), ctypes.byref(cols), ctypes.byref(nd),
                                         ctypes.byref(prec), ctypes.byref(stored)))
    offs = np.empty(nd.value, np.int32)
    capi.check(capi.lib().dq_devmat_offsets(h, capi.ptr(offs, ctypes.c_int32)))
    return offs.tolist()


@pytest.mark.parametrize("q", range(10, 16))
def test_reference_expm_negative_band_cases(d, q):
    import scipy.linalg
    import scipy.sparse
    import scipy.sparse.linalg
    from diaq_amd import capi
    n = 1 << q
    H = band_matrix(n, density=0.01, num_diagonals=10, seed=q)
    offs = [o for o in range(-5, 6) if np.any(np.diagonal(H, o))]
    M = O.Dia(n, n, offs, [np.diagonal(H, o).astype(complex) for o in offs])
    D = d.to_device(d.diaQMatrix.from_diagonals([n, n], offs, [np.diagonal(H, o).astype(complex) for o in offs]))
    if n <= 2048:
        U = d.expm_negative(d.from_numpy(H), T_REF)
        got = d.to_numpy(U)
        np.testing.assert_allclose(got, scipy.linalg.expm(-1j * H * T_REF), rtol=1e-4, atol=1e-4)
        Ud, roffs, _, _ = O.expm_negative_dense(H.astype(complex), offs, T_REF)
        assert list(U.dIndices) == roffs
        assert rel(got, Ud) < TOL
        return
    # larger: counts and offsets against the oracle's Taylor loop and structural chain, values through U v
    s_ref = O.scaling_exponent(O.estimate_1norm(M), T_REF, 1.0)
    ts = math.ldexp(T_REF, -s_ref)
    _, m_ref = O.taylor_loop(O.scale_by_imag_copy(M, -ts), n, 100, max(1e-15, 1e-12 * math.ldexp(1.0, -s_ref)))
    h, m, s = expm_negative_capi(D, T_REF)
    try:
        assert (m, s) == (m_ref, s_ref)
        r = [0]
        t_offs = [0]
        for _ in range(m):
            t_offs, _ = O.product_dryrun(t_offs, offs, n, n)
            r = sorted(set(r) | set(t_offs))
        for _ in range(s):
            r, _ = O.product_dryrun(r, r, n, n)
        assert devmat_offsets(h) == r
        rng = np.random.default_rng(q)
        Hs = scipy.sparse.csr_matrix(H)
        for _ in range(2):
            v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
            yr, yi = np.empty(n), np.empty(n)
            vr, vi = np.ascontiguousarray(v.real), np.ascontiguousarray(v.imag)
            P = ctypes.POINTER(ctypes.c_double)
            capi.check(capi.lib().dq_spmv_host(h, vr.ctypes.data_as(P), vi.ctypes.data_as(P), yr.ctypes.data_as(P),
                                               yi.ctypes.data_as(P)))
            ref = scipy.sparse.linalg.expm_multiply(-1j * T_REF * Hs, v)
            np.testing.assert_allclose(yr + 1j * yi, ref, rtol=1e-4, atol=1e-4)
            assert np.linalg.norm(yr + 1j * yi - ref) / np.linalg.norm(ref) < 1e-9
    finally:
        capi.lib().dq_devmat_destroy(h)
