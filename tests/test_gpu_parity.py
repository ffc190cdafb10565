"""GPU parity tests: the HIP path (through the diaq_gpu drop-in and the C ABI) against the oracle, the reference's
fixtures and dense numpy/scipy. Integer results (offset sets) bit-exact; FP64 values within 1e-12 relative
(max-abs error over max-abs value, the reference's own GPU test metric, python_tests/test_device_matrix.py:26)."""
import ctypes
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as O  # noqa: E402
from diaq_amd import hamiltonians as Hm  # noqa: E402

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
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


def banded(n, offsets, rng, cols=None):
    cols = n if cols is None else cols
    A = np.zeros((n, cols), complex)
    for off in offsets:
        k = np.arange(max(0, -off), min(n, cols - off))
        A[k, k + off] = rng.standard_normal(len(k)) + 1j * rng.standard_normal(len(k))
    return A


def dia_of(d, rows, cols, offs, vals):
    return d.diaQMatrix.from_diagonals([rows, cols], list(offs), list(vals))


# ---------------------------------------------------------------------------------------------------------------------
# The reference's DeviceMatrix script (python_tests/test_device_matrix.py:20-51), same inputs and bound
def test_reference_device_matrix_script(d):
    rng = np.random.default_rng(0)
    n = 2048
    A = banded(n, (-3, -1, 0, 2, 5), rng)
    B = banded(n, (-7, -2, 0, 3, 11), rng)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    dA, dB = d.to_device(d.from_numpy(A)), d.DeviceMatrix(d.from_numpy(B))
    worst = 0.0
    worst = max(worst, rel(d.to_numpy(dA.to_host()), A))
    worst = max(worst, rel(d.to_numpy_vector(d.spMV(dA, d.from_numpy_vector(x))), A @ x))
    worst = max(worst, rel(d.to_numpy(d.product(dA, dB).to_host()), A @ B))
    worst = max(worst, rel(d.to_numpy(d.product(d.product(dA, dB), dA).to_host()), A @ B @ A))
    worst = max(worst, rel(d.to_numpy(d.product(d.from_numpy(A), d.from_numpy(B))), A @ B))
    single = banded(n, (4,), rng)
    worst = max(worst, rel(d.to_numpy_vector(d.spMV(d.to_device(d.from_numpy(single)), d.from_numpy_vector(x))),
                           single @ x))
    small = banded(64, (-1, 0, 1), rng) / 4
    worst = max(worst, rel(d.to_numpy(d.pow(d.from_numpy(small), 7)), np.linalg.matrix_power(small, 7)))
    worst = max(worst, rel(d.to_numpy(d.pow(d.from_numpy(small), 8)), np.linalg.matrix_power(small, 8)))
    assert worst < TOL, worst
    import torch
    xr = torch.as_tensor(x.real, device="cuda").contiguous()
    xi = torch.as_tensor(x.imag, device="cuda").contiguous()
    yr, yi = torch.empty_like(xr), torch.empty_like(xi)
    d.spmv_device(dA, xr.data_ptr(), xi.data_ptr(), yr.data_ptr(), yi.data_ptr(),
                  torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert rel((yr + 1j * yi).cpu().numpy(), A @ x) < TOL


# ---------------------------------------------------------------------------------------------------------------------
# SpMV
@pytest.mark.parametrize("shape,offs", [
    ((4096, 4096), list(range(-4, 5))),                 # python_tests/check_gpu_vs_numpy.py:7-10
    ((3000, 5000), [-2999, -7, 0, 1, 4999]),            # extreme offsets, wide
    ((5000, 3000), [-4999, -100, 0, 2999]),             # tall
    ((777, 777), [13]),                                 # one off-diagonal
    ((777, 777), [-776]),                               # one corner element
    ((70000, 70000), [-65536, -1, 0, 3, 40000]),        # multi-tile, unsorted-free
    ((1000, 1000), [300, -1, 0, 1, -300, 17]),          # unsorted slots
])
def test_spmv_vs_oracle(d, shape, offs):
    rng = np.random.default_rng(1)
    A = banded(shape[0], offs, rng, cols=shape[1]) if shape[0] * shape[1] <= 25_000_000 else None
    if A is None:
        vals = [rng.standard_normal(O.diagonal_length(o, *shape)) + 1j * rng.standard_normal(O.diagonal_length(o, *shape))
                for o in offs]
        M = O.Dia(shape[0], shape[1], list(offs), vals)
        H = dia_of(d, shape[0], shape[1], offs, vals)
    else:
        M = O.Dia(shape[0], shape[1], list(offs),
                  [A[O.position_to_rc(o, np.arange(O.diagonal_length(o, *shape)))] for o in offs])
        H = dia_of(d, shape[0], shape[1], offs, M.vals)
    x = rng.standard_normal(shape[1]) + 1j * rng.standard_normal(shape[1])
    ref = O.spmv(M, x)
    got = d.to_numpy_vector(d.spMV(d.to_device(H), d.from_numpy_vector(x)))
    assert rel(got, ref) < TOL
    got_host = d.to_numpy_vector(d.spMV(H, d.from_numpy_vector(x)))  # host API: uploads per call
    assert rel(got_host, ref) < TOL


def test_spmv_known_answer_and_edge_cases(d):
    g = json.load(open(os.path.join(GOLD, "spmv_known.json")))
    A = (np.array(g["A"]["re"]) + 1j * np.array(g["A"]["im"])).reshape(g["A"]["shape"])
    x = np.array(g["x"]["re"]) + 1j * np.array(g["x"]["im"])
    y = d.to_numpy_vector(d.spMV(d.from_numpy(A), d.from_numpy_vector(x)))
    printed = np.array(g["y_printed"]["re"]) + 1j * np.array(g["y_printed"]["im"])
    np.testing.assert_allclose(y, printed, atol=2e-6)
    np.testing.assert_allclose(y, A @ x, rtol=1e-14, atol=1e-15)
    # no diagonals: y = 0; identity: y = x
    Z = d.diaQMatrix([5, 5])
    np.testing.assert_array_equal(d.to_numpy_vector(d.spMV(d.to_device(Z), d.from_numpy_vector(np.ones(5) + 1j))),
                                  np.zeros(5))
    xi = np.arange(9) + 2j
    np.testing.assert_array_equal(d.to_numpy_vector(d.spMV(d.identityMatrix(9), d.from_numpy_vector(xi))), xi)


def test_spmv_ntimes(d):
    rng = np.random.default_rng(2)
    A = banded(300, (-2, 0, 1, 7), rng) / 3
    x = rng.standard_normal(300) + 1j * rng.standard_normal(300)
    for n in (0, 1, 2, 5):
        got = d.to_numpy_vector(d.spMV_ntimes(d.from_numpy(A), d.from_numpy_vector(x), n))
        assert rel(got, np.linalg.matrix_power(A, n) @ x) < TOL
    R = d.from_numpy(banded(200, (0, 5), rng, cols=300))
    with pytest.raises(RuntimeError, match="square"):
        d.spMV_ntimes(R, d.from_numpy_vector(x), 2)
    got = d.to_numpy_vector(d.spMV_ntimes(R, d.from_numpy_vector(x), 1))
    assert rel(got, d.to_numpy(R) @ x) < TOL


def test_spmv_device_capi_streams_and_fp32(d):
    import torch
    from diaq_amd import capi
    rng = np.random.default_rng(3)
    N = 1 << 16
    offs = Hm.random_offsets(N, 24, seed=3)
    vals = [rng.standard_normal(O.diagonal_length(o, N, N)) + 1j * rng.standard_normal(O.diagonal_length(o, N, N))
            for o in offs]
    M = O.Dia(N, N, offs, vals)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    ref = O.spmv(M, x)
    H = dia_of(d, N, N, offs, vals)
    for prec, tol, dt in (("fp64", TOL, torch.float64), ("fp32", 2e-6, torch.float32)):
        D = d.to_device(H, precision=prec)
        assert D.precision == prec and D.stored == M.stored
        xr = torch.as_tensor(x.real, device="cuda").to(dt).contiguous()
        xi = torch.as_tensor(x.imag, device="cuda").to(dt).contiguous()
        yr, yi = torch.empty_like(xr), torch.empty_like(xi)
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            capi.check(capi.lib().dq_spmv_device(D.handle, xr.data_ptr(), xi.data_ptr(), yr.data_ptr(), yi.data_ptr(),
                                                 s.cuda_stream))
        s.synchronize()
        got = (yr.double() + 1j * yi.double()).cpu().numpy()
        assert rel(got, ref) < tol, prec
        for bands in (2, 3, 8):  # column-band passes: same values up to summation order
            capi.check(capi.lib().dq_spmv_device_banded(D.handle, xr.data_ptr(), xi.data_ptr(), yr.data_ptr(),
                                                        yi.data_ptr(), bands, 0))
            torch.cuda.synchronize()
            assert rel((yr.double() + 1j * yi.double()).cpu().numpy(), ref) < tol, (prec, bands)


@pytest.mark.parametrize("shape,offs", [
    ((5000, 5000), [-1024, -300, -2, -1, 0, 1, 2, 5, 300, 1024, 3000]),  # near + far diagonals, ragged last tile
    ((3000, 4500), [-2999, -129, -64, 0, 64, 128, 129, 4499]),           # wide, near diagonals on both halo edges
    ((4500, 3000), [-4499, -2048, -7, 0, 7, 2047, 2999]),                # tall
    ((100, 100), [-99, -3, -1, 0, 1, 3, 99]),                            # smaller than one tile: window off both ends
    ((9000, 9000), list(range(-40, 41, 3))),                              # a band
])
def test_spmv_every_schedule(d, shape, offs):
    """Every (rows per lane, non-temporal, load schedule, LDS x window) combination of the SpMV kernel, forced through
    the process-wide knobs, against the oracle in FP64 and FP32 (Kahan-compensated sums)."""
    import torch
    from diaq_amd import capi
    L = capi.lib()
    rng = np.random.default_rng(11)
    vals = [rng.standard_normal(O.diagonal_length(o, *shape)) + 1j * rng.standard_normal(O.diagonal_length(o, *shape))
            for o in offs]
    M = O.Dia(shape[0], shape[1], list(offs), vals)
    x = rng.standard_normal(shape[1]) + 1j * rng.standard_normal(shape[1])
    ref = O.spmv(M, x)
    H = dia_of(d, shape[0], shape[1], offs, vals)
    try:
        for prec, tol, dt in (("fp64", TOL, torch.float64), ("fp32", 2e-6, torch.float32)):
            D = d.to_device(H, precision=prec)
            xr = torch.as_tensor(x.real, device="cuda").to(dt).contiguous()
            xi = torch.as_tensor(x.imag, device="cuda").to(dt).contiguous()
            for rpt in (1, 2, 4, 8, 16):
                for nt in (0, 1):
                    for pipe in (0, 1):
                        for lds in (0, 1):
                            for xcd in (0, 1):
                                for key, v in (("spmv_rows_per_lane", rpt), ("spmv_nontemporal", nt),
                                               ("spmv_pipelined", pipe), ("spmv_lds", lds), ("spmv_xcd", xcd)):
                                    capi.check(L.dq_set_tuning(key.encode(), v))
                                yr = torch.full((shape[0],), float("nan"), dtype=dt, device="cuda")
                                yi = torch.full((shape[0],), float("nan"), dtype=dt, device="cuda")
                                capi.check(L.dq_spmv_device(D.handle, xr.data_ptr(), xi.data_ptr(), yr.data_ptr(),
                                                            yi.data_ptr(), 0))
                                torch.cuda.synchronize()
                                got = (yr.double() + 1j * yi.double()).cpu().numpy()
                                assert rel(got, ref) < tol, (prec, rpt, nt, pipe, lds, xcd)
    finally:
        for key, v in (("spmv_rows_per_lane", 0), ("spmv_nontemporal", -1), ("spmv_pipelined", -1), ("spmv_lds", -1),
                       ("spmv_xcd", -1)):
            capi.check(L.dq_set_tuning(key.encode(), v))


def test_spmv_tuner_keeps_results(d):
    """dq_spmv_tune picks a schedule per matrix (the LDS window only where near diagonals exist); the tuned SpMV gives
    the oracle's values."""
    import torch
    from diaq_amd import capi
    L = capi.lib()
    rng = np.random.default_rng(12)
    for N, offs, lds_possible in ((1 << 16, list(range(-8, 9)), True), (1 << 16, Hm.random_offsets(1 << 16, 9, 4),
                                                                         False)):
        vals = [rng.standard_normal(O.diagonal_length(o, N, N)) + 1j * rng.standard_normal(O.diagonal_length(o, N, N))
                for o in offs]
        M = O.Dia(N, N, list(offs), vals)
        x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
        D = d.to_device(dia_of(d, N, N, offs, vals))
        best = ctypes.c_float()
        capi.check(L.dq_spmv_tune(D.handle, 2, 0, ctypes.byref(best)))
        cfg = [ctypes.c_int(-5) for _ in range(5)]
        capi.check(L.dq_spmv_tuned_config(D.handle, *[ctypes.byref(c) for c in cfg]))
        assert cfg[0].value in (2, 4, 8, 16) and cfg[1].value in (0, 1) and cfg[2].value in (0, 1)
        assert cfg[3].value in ((0, 1) if lds_possible else (0,)) and cfg[4].value in (0, 1)
        assert best.value > 0
        got = d.to_numpy_vector(d.spMV(D, d.from_numpy_vector(x)))
        assert rel(got, O.spmv(M, x)) < TOL


def test_spmv_full_size_property_2p24(d):
    """N = 2^24 with 64 random offsets (the SpMV benchmark shape) -- values from closed-form functions evaluated on
    both sides; rows are checked on a sample against the oracle rule, and linearity on the whole vector."""
    import torch
    from diaq_amd import capi
    N = 1 << 24
    offs = Hm.random_offsets(N, 64, seed=0)
    lens = [O.diagonal_length(o, N, N) for o in offs]
    assert sum(lens) == 798_254_616
    out = ctypes.c_void_p()
    o32 = np.asarray(offs, np.int32)
    capi.check(capi.lib().dq_devmat_zeros(N, N, len(offs), capi.ptr(o32, ctypes.c_int32), capi.DQ_F64,
                                          ctypes.byref(out)))
    try:
        total = sum(lens)
        pos = torch.cat([torch.arange(L, device="cuda", dtype=torch.float64) for L in lens])
        kid = torch.cat([torch.full((L,), k, device="cuda", dtype=torch.float64) for k, L in enumerate(lens)])
        re = torch.cos(1e-3 * pos + 0.7 * kid)
        im = torch.sin(3e-4 * pos - 0.3 * kid)
        del pos, kid
        capi.check(capi.lib().dq_devmat_set_packed_device(out, re.data_ptr(), im.data_ptr(), 0))
        torch.cuda.synchronize()
        del re, im
        idx = torch.arange(N, device="cuda", dtype=torch.float64)
        x1r, x1i = torch.cos(2e-5 * idx), torch.sin(5e-6 * idx)
        x2r, x2i = torch.sin(7e-6 * idx + 1.0), torch.cos(1e-6 * idx)
        ys = []
        for xr, xi in ((x1r, x1i), (x2r, x2i), (x1r + 2 * x2r, x1i + 2 * x2i)):
            yr, yi = torch.empty_like(xr), torch.empty_like(xi)
            capi.check(capi.lib().dq_spmv_device(out, xr.data_ptr(), xi.data_ptr(), yr.data_ptr(), yi.data_ptr(), 0))
            torch.cuda.synchronize()
            ys.append((yr, yi))
        lin_r = ys[2][0] - (ys[0][0] + 2 * ys[1][0])
        lin_i = ys[2][1] - (ys[0][1] + 2 * ys[1][1])
        scale = float(torch.max(torch.abs(ys[2][0])))
        assert float(torch.max(torch.abs(lin_r))) / scale < 1e-12
        assert float(torch.max(torch.abs(lin_i))) / scale < 1e-12
        # sampled rows against the row rule with the same closed-form values
        rows = np.unique(np.concatenate([np.random.default_rng(0).integers(0, N, 4000), [0, N - 1, N // 2]]))
        x1 = np.cos(2e-5 * np.arange(N)) + 1j * np.sin(5e-6 * np.arange(N))
        expect = np.zeros(len(rows), complex)
        for k, o in enumerate(offs):
            p = rows if o >= 0 else rows + o
            c = rows + o
            ok = (p >= 0) & (p < lens[k]) & (c >= 0) & (c < N)
            a = np.cos(1e-3 * p[ok] + 0.7 * k) + 1j * np.sin(3e-4 * p[ok] - 0.3 * k)
            expect[ok] += a * x1[c[ok]]
        got = (ys[0][0] + 1j * ys[0][1]).cpu().numpy()[rows]
        assert rel(got, expect) < 1e-12
    finally:
        capi.lib().dq_devmat_destroy(out)


# ---------------------------------------------------------------------------------------------------------------------
# SpGEMM
def test_spgemm_reference_known_answer_pairs(d):
    g = json.load(open(os.path.join(GOLD, "spgemm_cases.json")))
    for c in g["cases"]:
        A = (np.array(c["A"]["re"]) + 1j * np.array(c["A"]["im"])).reshape(c["A"]["shape"])
        B = (np.array(c["B"]["re"]) + 1j * np.array(c["B"]["im"])).reshape(c["B"]["shape"])
        P = (np.array(c["product"]["re"]) + 1j * np.array(c["product"]["im"])).reshape(c["product"]["shape"])
        C = d.product(d.from_numpy(A), d.from_numpy(B))
        assert list(C.dIndices) == c["product_offsets_pruned"], c["case"]
        np.testing.assert_allclose(d.to_numpy(C), P, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("shapes,oa,ob", [
    ((1024, 1024, 1024), [-15, -3, 0, 2, 9, 500], [-500, -9, 0, 1, 3]),
    ((600, 900, 700), [-599, -20, 0, 33, 899], [-899, -1, 0, 40, 699]),
    ((4096, 4096, 4096), list(range(-4, 5)), list(range(-4, 5))),
    ((64, 64, 64), [0, 63, -63], [-63, 63, 1]),
    ((5000, 5000, 5000), [-4100, -2049, 0, 2047, 3000], [-3000, -1, 2, 2049]),  # diagonals start mid-chunk
    ((600, 600, 600), list(range(-599, 600, 3)), list(range(-598, 599, 4))),    # > 256 pairs per result diagonal
])
def test_product_device_structure_and_values(d, shapes, oa, ob):
    rng = np.random.default_rng(4)
    m, k, n = shapes
    A = banded(m, oa, rng, cols=k)
    B = banded(k, ob, rng, cols=n)
    MA = O.Dia(m, k, oa, [A[O.position_to_rc(o, np.arange(O.diagonal_length(o, m, k)))] for o in oa])
    MB = O.Dia(k, n, ob, [B[O.position_to_rc(o, np.arange(O.diagonal_length(o, k, n)))] for o in ob])
    ref = O.product(MA, MB)
    C = d.product(d.to_device(dia_of(d, m, k, oa, MA.vals)), d.to_device(dia_of(d, k, n, ob, MB.vals)))
    assert C.diagonals == ref.offsets  # structural result, no pruning
    H = C.to_host()
    assert rel(d.to_numpy(H), O.to_dense(ref)) < TOL
    Pr = d.prune(C)
    assert Pr.diagonals == O.prune(ref).offsets


def test_products_in_a_row_with_equal_offsets(d):
    """Consecutive products whose operands have the same offset lists (the dry-run of the previous call is reused):
    new values, a new shape and a new inner dimension each give the oracle's structure and values."""
    rng = np.random.default_rng(12)
    oa, ob = [-7, -1, 0, 3, 40], [-40, -2, 0, 1, 9]
    for m, k, n in ((300, 300, 300), (300, 300, 300), (303, 303, 303), (200, 260, 180), (200, 310, 180),
                    (300, 300, 300)):
        A = banded(m, oa, rng, cols=k)
        B = banded(k, ob, rng, cols=n)
        MA = O.Dia(m, k, oa, [A[O.position_to_rc(o, np.arange(O.diagonal_length(o, m, k)))] for o in oa])
        MB = O.Dia(k, n, ob, [B[O.position_to_rc(o, np.arange(O.diagonal_length(o, k, n)))] for o in ob])
        ref = O.product(MA, MB)
        C = d.product(d.to_device(dia_of(d, m, k, oa, MA.vals)), d.to_device(dia_of(d, k, n, ob, MB.vals)))
        assert C.diagonals == ref.offsets, (m, k, n)
        assert rel(d.to_numpy(C.to_host()), O.to_dense(ref)) < TOL, (m, k, n)


def test_prune_constructed_cancellation(d):
    A = d.to_device(d.from_numpy(np.array([[1, 1], [0, 1]], complex)))
    B = d.to_device(d.from_numpy(np.array([[1, -1], [0, 1]], complex)))
    C = d.product(A, B)
    assert C.diagonals == [0, 1]
    assert d.prune(C).diagonals == [0]
    assert list(d.product(d.from_numpy(np.array([[1, 1], [0, 1]], complex)),
                          d.from_numpy(np.array([[1, -1], [0, 1]], complex))).dIndices) == [0]


def test_norms_trace(d):
    rng = np.random.default_rng(5)
    A = banded(3000, (-1000, -2, 0, 1, 2999), rng)
    M = O.from_dense(A)
    D = d.to_device(d.from_numpy(A))
    assert abs(d.estimate_1norm(D) - O.estimate_1norm(M)) / O.estimate_1norm(M) < TOL
    assert abs(d.norm(D) - O.frobenius(M)) / O.frobenius(M) < TOL
    assert abs(d.trace(D) - O.trace(M)) / abs(O.trace(M)) < TOL
    assert abs(d.trace(d.from_numpy(A)) - np.trace(A)) / abs(np.trace(A)) < TOL
    rows, cols, offs, vals = Hm.tfim(12, 1.0, 0.7)
    T = d.to_device(dia_of(d, rows, cols, offs, vals))
    assert abs(d.estimate_1norm(T) - 19.4) < 1e-12


def test_pow_vs_oracle(d):
    rng = np.random.default_rng(6)
    small = banded(96, (-2, -1, 0, 1, 3), rng) / 5
    M = O.from_dense(small)
    for n in (0, 1, 2, 3, 6, 7, 8, 13):
        P = d.pow(d.from_numpy(small), n)
        ref = O.pow_gpu(M, n)
        assert list(P.dIndices) == ref.offsets, n
        assert rel(d.to_numpy(P), O.to_dense(ref)) < TOL


def test_tfim20_hh_fp64_fp32(d):
    """TFIM-20 H.H: 727 offsets bit-exact (structural and pruned), values vs the closed form of the diagonal."""
    rows, cols, offs, vals = Hm.tfim(20, 1.0, 0.7)
    H = dia_of(d, rows, cols, offs, vals)
    res_ref, _ = O.product_dryrun(offs, offs, rows, cols)
    out = {}
    for prec in ("fp64", "fp32"):
        D = d.to_device(H, precision=prec)
        C = d.product(D, D)
        assert C.diagonals == res_ref and len(res_ref) == 727
        assert C.stored == 687_341_554
        out[prec] = C
    # (H^2)_{rr} = (diag_r)^2 + n h^2 exactly: check the main diagonal of the FP64 result
    dg = out["fp64"].get_diagonal(0)
    expect = vals[offs.index(0)] ** 2 + 20 * 0.7 ** 2
    assert rel(dg, expect) < TOL
    # FP32 vs FP64: relative Frobenius difference at FP32 rounding level
    err = d.norm_diff(out["fp32"], out["fp64"]) / d.norm(out["fp64"])
    assert 0 < err < 1e-6


@pytest.mark.parametrize("model", ["tfim", "xxz"])
def test_product_zero_block_skip_is_exact(d, model):
    """For operands of 2^16 rows and more the product maps their all-zero 256-row blocks and drops the pairs whose rows
    in a chunk are all zero ("product_skip_zero_blocks"). A dropped pair's products are exact zeros, so the values
    keep their bits with the skip on or off: for H.H (one map) and H.G (two maps, G with other values or other
    offsets), in FP64 and FP32, for every chunk size, and with the map dq_spmv_tune keeps in an FP64 handle. The FP64
    values agree with the oracle's product (src/spGEMM_gpu.cu:25-151) within 1e-12."""
    from diaq_amd import capi
    L = capi.lib()
    q = 16
    if model == "tfim":
        H = Hm.tfim(q, 1.0, 0.7)
        G = Hm.tfim(q, 0.9, 0.4)
    else:
        H = Hm.xxz(q, seed=6)
        G = Hm.tfim(q, 1.1, 0.5)
    ref = {}
    for name, (X, Y) in (("HH", (H, H)), ("HG", (H, G))):
        R = O.product(O.Dia(*X), O.Dia(*Y))
        ref[name] = dict(zip(R.offsets, R.vals))

    def values(C):
        return {k: C.get_diagonal(k) for k in C.diagonals}

    def products(prec):
        DH = d.to_device(dia_of(d, *H), precision=prec)
        DG = d.to_device(dia_of(d, *G), precision=prec)
        return DH, {"HH": values(d.product(DH, DH)), "HG": values(d.product(DH, DG))}

    try:
        for prec in ("fp64", "fp32"):
            capi.check(L.dq_set_tuning(b"product_skip_zero_blocks", 0))
            capi.check(L.dq_set_tuning(b"product_rows_per_lane", 8))
            _, base = products(prec)
            capi.check(L.dq_set_tuning(b"product_skip_zero_blocks", 1))
            for rpt in (4, 8, 16):
                capi.check(L.dq_set_tuning(b"product_rows_per_lane", rpt))
                DH, got = products(prec)
                for name in base:
                    assert list(got[name]) == list(base[name]), (prec, rpt, name)
                    for k in base[name]:
                        np.testing.assert_array_equal(got[name][k], base[name][k], err_msg=f"{prec} {rpt} {name} {k}")
            if prec == "fp64":
                for name in base:
                    for k, v in base[name].items():
                        assert rel(v, ref[name][k]) < TOL, (name, k)
                capi.check(L.dq_set_tuning(b"product_rows_per_lane", 8))
                capi.check(L.dq_spmv_tune(DH.handle, 2, 0, None))  # the handle now keeps its zero-block map
                tuned = values(d.product(DH, DH))
                for k in base["HH"]:
                    np.testing.assert_array_equal(tuned[k], base["HH"][k], err_msg=f"tuned map {k}")
    finally:
        capi.check(L.dq_set_tuning(b"product_skip_zero_blocks", 1))
        capi.check(L.dq_set_tuning(b"product_rows_per_lane", 0))
    with pytest.raises(RuntimeError, match="product_skip_zero_blocks"):
        capi.check(L.dq_set_tuning(b"product_skip_zero_blocks", 2))


# ---------------------------------------------------------------------------------------------------------------------
# SpME
@pytest.mark.parametrize("n,t", [(6, 0.1), (8, 0.7), (9, 2.5)])
def test_expm_negative_vs_oracle(d, n, t):
    rows, cols, offs, vals = Hm.tfim(n, 1.0, 0.7)
    M = O.Dia(rows, cols, offs, vals)
    Ud, roffs, m, s = O.expm_negative_dense(O.to_dense(M), offs, t)
    U = d.expm_negative(dia_of(d, rows, cols, offs, vals), t)
    assert list(U.dIndices) == roffs
    assert rel(d.to_numpy(U), Ud) < TOL


def test_expm_negative_tfim12_config(d):
    """TFIM-12, t = 0.1: s = 1 by the closed form, offsets from the structural chain, values vs dense FP64 expm."""
    import scipy.linalg
    rows, cols, offs, vals = Hm.tfim(12, 1.0, 0.7)
    M = O.Dia(rows, cols, offs, vals)
    Hd = O.to_dense(M)
    Ud, roffs, m, s = O.expm_negative_dense(Hd, offs, 0.1)
    assert s == 1
    U = d.expm_negative(dia_of(d, rows, cols, offs, vals), 0.1)
    assert list(U.dIndices) == roffs
    Ug = d.to_numpy(U)
    assert rel(Ug, Ud) < TOL
    assert np.max(np.abs(Ug - scipy.linalg.expm(-1j * 0.1 * Hd))) < 1e-10


def test_expm_positive_and_single_diagonal(d):
    import scipy.linalg
    rng = np.random.default_rng(7)
    A = banded(40, (-2, 0, 3), rng) * 0.3
    M = O.from_dense(A)
    U = d.expm(d.from_numpy(A))
    ref = O.expm(M)
    assert list(U.dIndices) == ref.offsets
    assert rel(d.to_numpy(U), O.to_dense(ref)) < TOL
    assert np.max(np.abs(d.to_numpy(U) - scipy.linalg.expm(A))) < 1e-10
    Dm = banded(50, (0,), rng)
    np.testing.assert_allclose(d.to_numpy(d.expm_negative(d.from_numpy(Dm), 0.3)),
                               np.diag(np.exp(-1j * 0.3 * np.diag(Dm))), rtol=1e-13)
    np.testing.assert_allclose(d.to_numpy(d.expm(d.from_numpy(Dm))), np.diag(np.exp(np.diag(Dm))), rtol=1e-13)
    S = banded(50, (3,), rng)  # the reference's fast-path quirk for one off-diagonal
    assert rel(d.to_numpy(d.expm_negative(d.from_numpy(S), 0.5)),
               O.to_dense(O.expm_negative(O.from_dense(S), 0.5))) < TOL


# ---------------------------------------------------------------------------------------------------------------------
# Fused evolution
def test_hamsim_vs_oracle(d):
    import scipy.sparse.linalg
    from diaq_amd import capi
    rows, cols, offs, vals = Hm.xxz(10, seed=1)
    M = O.Dia(rows, cols, offs, vals)
    psi = Hm.product_state(10, 2)
    H = dia_of(d, rows, cols, offs, vals)
    try:
        # forced term-kernel schedules (LDS window, XCD order, rows per lane), then the timed choice
        for lds, xcd, rpt in ((0, 0, 2), (1, 0, 4), (0, 1, 8), (1, 1, 2), (-1, -1, 0)):
            capi.check(capi.lib().dq_set_tuning(b"hamsim_lds", lds))
            capi.check(capi.lib().dq_set_tuning(b"spmv_xcd", xcd))
            capi.check(capi.lib().dq_set_tuning(b"hamsim_rows_per_lane", rpt))
            for t, r, m in ((0.5, 5, 12), (1.0, 3, 30)):
                got = d.to_numpy_vector(d.hamiltonian_simulation_cuda(H, d.from_numpy_vector(psi), t, r, m))
                ref = O.hamiltonian_simulation(M, psi, t, r, m)
                assert rel(got, ref) < TOL, (lds, xcd, rpt, t, r, m)
    finally:
        capi.check(capi.lib().dq_set_tuning(b"hamsim_lds", -1))
        capi.check(capi.lib().dq_set_tuning(b"spmv_xcd", -1))
        capi.check(capi.lib().dq_set_tuning(b"hamsim_rows_per_lane", 0))
    exact = scipy.sparse.linalg.expm_multiply(-1j * 1.0 * O.to_dense(M), psi)
    got = d.to_numpy_vector(d.hamiltonian_simulation_cuda(H, d.from_numpy_vector(psi), 1.0, 10, 20))
    assert np.linalg.norm(got - exact) < 1e-10
    same = d.to_numpy_vector(d.hamiltonian_simulation_cuda(H, d.from_numpy_vector(psi), 1.0, 0))
    np.testing.assert_array_equal(same, psi)


@pytest.mark.parametrize("model", ["xxz", "tfim"])
def test_hamsim_zero_block_skip_is_exact(d, model):
    """The fused evolution skips the 256-row blocks of a diagonal whose values are all zero (most of a qubit
    Hamiltonian's flip diagonals). In FP64 a skipped block's products are exact zeros, so the result has the same bits
    as reading every block, under every term-kernel schedule; in FP32 (compensated sums) it stays within FP32
    rounding."""
    import torch
    from diaq_amd import capi
    L = capi.lib()
    n = 14
    rows, cols, offs, vals = Hm.xxz(n, seed=5) if model == "xxz" else Hm.tfim(n, 1.0, 0.7)
    psi = Hm.product_state(n, 6)
    out = {}
    try:
        for prec in ("fp64", "fp32"):
            D = d.to_device(dia_of(d, rows, cols, offs, vals), precision=prec)
            dt = torch.float64 if prec == "fp64" else torch.float32
            for skip in (1, 0):
                for lds, rpt in ((0, 2), (1, 2), (0, 8), (1, 8), (-1, 0)):
                    capi.check(L.dq_set_tuning(b"hamsim_skip_zero_blocks", skip))
                    capi.check(L.dq_set_tuning(b"hamsim_lds", lds))
                    capi.check(L.dq_set_tuning(b"hamsim_rows_per_lane", rpt))
                    pr = torch.as_tensor(psi.real, device="cuda").to(dt).contiguous()
                    pi = torch.as_tensor(psi.imag, device="cuda").to(dt).contiguous()
                    d.hamsim_device(D, pr.data_ptr(), pi.data_ptr(), 0.8, 4, 14)
                    torch.cuda.synchronize()
                    out[prec, skip, lds, rpt] = (pr.double() + 1j * pi.double()).cpu().numpy()
    finally:
        capi.check(L.dq_set_tuning(b"hamsim_skip_zero_blocks", 1))
        capi.check(L.dq_set_tuning(b"hamsim_lds", -1))
        capi.check(L.dq_set_tuning(b"hamsim_rows_per_lane", 0))
    base = out["fp64", 0, 0, 2]
    for key, v in out.items():
        if key[0] == "fp64":
            np.testing.assert_array_equal(v, base, err_msg=str(key))
        else:
            assert rel(v, base) < 5e-6, key
    M = O.Dia(rows, cols, offs, vals)
    assert rel(out["fp64", 1, -1, 0], O.hamiltonian_simulation(M, psi, 0.8, 4, 14)) < TOL
    for key, bad in ((b"hamsim_skip_zero_blocks", -1), (b"hamsim_rows_per_lane", 16), (b"hamsim_lds", 2)):
        with pytest.raises(RuntimeError, match=key.decode()):
            capi.check(L.dq_set_tuning(key, bad))


def test_hamsim_device_in_place(d):
    import torch
    rows, cols, offs, vals = Hm.xxz(12, seed=3)
    M = O.Dia(rows, cols, offs, vals)
    psi = Hm.product_state(12, 4)
    D = d.to_device(dia_of(d, rows, cols, offs, vals))
    pr = torch.as_tensor(psi.real, device="cuda").contiguous()
    pi = torch.as_tensor(psi.imag, device="cuda").contiguous()
    d.hamsim_device(D, pr.data_ptr(), pi.data_ptr(), 0.6, 3, 15)
    torch.cuda.synchronize()
    got = (pr + 1j * pi).cpu().numpy()
    assert rel(got, O.hamiltonian_simulation(M, psi, 0.6, 3, 15)) < TOL
    # FP32 form with the LDS window forced on and off: within the FP32 bound of the FP64 result
    from diaq_amd import capi
    D32 = d.to_device(dia_of(d, rows, cols, offs, vals), precision="fp32")
    try:
        for lds in (0, 1):
            capi.check(capi.lib().dq_set_tuning(b"hamsim_lds", lds))
            qr = torch.as_tensor(psi.real, device="cuda").float().contiguous()
            qi = torch.as_tensor(psi.imag, device="cuda").float().contiguous()
            d.hamsim_device(D32, qr.data_ptr(), qi.data_ptr(), 0.6, 3, 15)
            torch.cuda.synchronize()
            assert rel((qr.double() + 1j * qi.double()).cpu().numpy(), got) < 5e-6, lds
    finally:
        capi.check(capi.lib().dq_set_tuning(b"hamsim_lds", -1))


# ---------------------------------------------------------------------------------------------------------------------
# spmm_dia_device: the reference's four shapes x {complex128, float64} (python_tests/test_spmm_dia_device.py:10-39)
CASES = [(4096, 4096, 64, [-3, -1, 0, 2, 5]), (1000, 1200, 1, [0, 7, -9]), (300, 200, 17, list(range(-40, 41, 3))),
         (64, 64, 3, list(range(-63, 64)))]


@pytest.mark.parametrize("complex_", [True, False])
@pytest.mark.parametrize("case", CASES)
def test_spmm_dia_device(d, case, complex_):
    import torch
    M, K, R, offs = case
    rng = np.random.default_rng(8)
    V = rng.standard_normal((len(offs), M)) + (1j * rng.standard_normal((len(offs), M)) if complex_ else 0)
    r = np.arange(M)
    for k, o in enumerate(offs):
        V[k, ~((r + o >= 0) & (r + o < K))] = 0
    B = rng.standard_normal((K, R)) + (1j * rng.standard_normal((K, R)) if complex_ else 0)
    ref = O.spmm_dia(offs, V, M, K, B)
    dt = torch.complex128 if complex_ else torch.float64
    tV = torch.as_tensor(V, device="cuda").to(dt).contiguous()
    tB = torch.as_tensor(B, device="cuda").to(dt).contiguous()
    tC = torch.empty((M, R), dtype=dt, device="cuda")
    off = torch.tensor(offs, dtype=torch.int64, device="cuda")
    spmm_every_kernel(d, torch, off, tV, tB, tC, offs, M, K, R, complex_, ref)


# (spmm_tiled, spmm_cols_per_lane): the row-chunk kernel (automatic and each width), the one-element-per-lane
# kernel, the LDS-tiled kernel with 1, 2, 4 elements per lane
# (spmm_tiled, spmm_cols_per_lane, spmm_rows_per_lane): every kernel form; -1 rows = the automatic choice
SPMM_KERNELS = [(0, 0, 0), (0, -1, 0), (0, 1, 0), (0, 2, 0), (0, 4, 0), (0, 8, 0), (1, 0, 0), (2, 0, 0), (4, 0, 0),
                (0, 0, 2), (0, 0, 4), (0, 0, 8), (0, 0, -1)]


def spmm_every_kernel(d, torch, off, tV, tB, tC, offs, M, K, R, complex_, ref):
    """Every spmm_dia kernel: within 1e-13 of the oracle and the same bits as each other."""
    from diaq_amd import capi
    first = None
    try:
        for tiled, cols, rows in SPMM_KERNELS:
            capi.check(capi.lib().dq_set_tuning(b"spmm_tiled", tiled))
            capi.check(capi.lib().dq_set_tuning(b"spmm_cols_per_lane", cols))
            capi.check(capi.lib().dq_set_tuning(b"spmm_rows_per_lane", rows))
            tC.fill_(float("nan"))
            d.spmm_dia_device(off.data_ptr(), tV.data_ptr(), len(offs), M, K, tB.data_ptr(), R, tC.data_ptr(),
                              complex_, torch.cuda.current_stream().cuda_stream)
            torch.cuda.synchronize()
            got = tC.cpu().numpy()
            assert rel(got, ref) < 1e-13, (tiled, cols, rows)
            if first is None:
                first = got
            else:
                assert np.array_equal(got, first), (tiled, cols, rows)
    finally:
        capi.check(capi.lib().dq_set_tuning(b"spmm_tiled", 0))
        capi.check(capi.lib().dq_set_tuning(b"spmm_cols_per_lane", 0))
        capi.check(capi.lib().dq_set_tuning(b"spmm_rows_per_lane", -1))


@pytest.mark.parametrize("M,K,R,offs", [
    (5000, 5000, 8, [-4100, -2, 0, 3, 600]),         # near and far diagonals, ragged last tile
    (3000, 2000, 1, [-2999, -700, 0, 511, 512, 1999]),  # one column: tiles of 256..1024 rows, halo edges
    (700, 900, 1500, [-5, 0, 5]),                      # wider than a tile: the flat kernel
    (2048, 2048, 3, [-900, 1200]),                     # no diagonal near the tile: no window
    (4096, 4096, 16, list(range(-20, 21))),            # 41 diagonals: several unrolled groups and a ragged one
    (1025, 1100, 128, [-1024, -3, 0, 9, 1099]),        # wide: 16 chunks of 8 columns per row
    (333, 333, 12, [-5, 1, 4]),                        # 12 columns: chunks of 4
    (1001, 1003, 5, [-7, -6, -5, 10, 11, 30, 1002]),   # runs and jumps of offsets, odd columns, ragged row blocks
    (40000, 40000, 64, list(range(-3, 4))),            # many rows, wide: the automatic row-block form
])
def test_spmm_dia_device_tiling_edges(d, M, K, R, offs):
    import torch
    rng = np.random.default_rng(9)
    V = rng.standard_normal((len(offs), M)) + 1j * rng.standard_normal((len(offs), M))
    r = np.arange(M)
    for k, o in enumerate(offs):
        V[k, ~((r + o >= 0) & (r + o < K))] = 0
    B = rng.standard_normal((K, R)) + 1j * rng.standard_normal((K, R))
    ref = O.spmm_dia(offs, V, M, K, B)
    tV = torch.as_tensor(V, device="cuda").contiguous()
    tB = torch.as_tensor(B, device="cuda").contiguous()
    tC = torch.full((M, R), float("nan"), dtype=torch.complex128, device="cuda")
    off = torch.tensor(offs, dtype=torch.int64, device="cuda")
    spmm_every_kernel(d, torch, off, tV, tB, tC, offs, M, K, R, True, ref)


def test_device_block_cache_reuse(d):
    """Freed device blocks are cached and handed out again: products computed into reused blocks (holding stale
    values) equal the first one bit for bit, with the cache on, off, and after releasing it."""
    from diaq_amd import capi
    L = capi.lib()
    rows, cols, offs, vals = Hm.tfim(14, 1.0, 0.7)
    D = d.to_device(dia_of(d, rows, cols, offs, vals))
    first = d.product(D, D).to_host()
    ref = d.to_numpy(first)
    try:
        for budget in (-1, 0, 64, -1):
            capi.check(L.dq_set_tuning(b"alloc_cache_mb", budget))
            for _ in range(3):
                C = d.product(D, D)
                np.testing.assert_array_equal(d.to_numpy(C.to_host()), ref)
                del C
            capi.check(L.dq_release_cached_memory())
    finally:
        capi.check(L.dq_set_tuning(b"alloc_cache_mb", -1))


def test_dense_path_matches_pair_products(d):
    """Near-dense products inside pow and expm take the dense path (the library's complex MFMA GEMM): the offsets are
    those of the pair-by-pair products and the values agree to rounding, against the pair path and the oracle, for a
    size that is not a multiple of the GEMM tile (300) and in FP32."""
    from diaq_amd import capi
    L = capi.lib()
    rng = np.random.default_rng(21)
    n = 300
    offs = list(range(-299, 300, 2))  # 300 diagonals: squaring them is near-dense
    A = banded(n, offs, rng) / 40
    H = d.from_numpy(A)
    ref = np.linalg.matrix_power(A, 4)
    outs = {}
    try:
        for dense in (0, 1):
            capi.check(L.dq_set_tuning(b"product_dense", dense))
            P = d.pow(H, 4)
            outs[dense] = (list(P.dIndices), d.to_numpy(P))
            assert rel(outs[dense][1], ref) < TOL, dense
        assert outs[0][0] == outs[1][0]
        np.testing.assert_allclose(outs[1][1], outs[0][1], rtol=0, atol=1e-13 * np.abs(ref).max())
        # FP32 operands: widened to FP64 for the dense product, rounded once at the end
        D32 = d.to_device(H, precision="fp32")
        for dense in (0, 1):
            capi.check(L.dq_set_tuning(b"product_dense", dense))
            P32 = d.pow(D32, 4)
            assert list(P32.diagonals) == outs[0][0]
            assert rel(d.to_numpy(P32.to_host()), ref) < 2e-6, dense
        # expm of a 10-qubit TFIM: the final squarings are near-dense
        rows, cols, to, tv = Hm.tfim(10, 1.0, 0.7)
        Ud, roffs, _, s = O.expm_negative_dense(O.to_dense(O.Dia(rows, cols, to, tv)), to, 0.4)
        assert s >= 1
        for dense in (0, 1):
            capi.check(L.dq_set_tuning(b"product_dense", dense))
            U = d.expm_negative(dia_of(d, rows, cols, to, tv), 0.4)
            assert list(U.dIndices) == roffs, dense
            assert rel(d.to_numpy(U), Ud) < TOL, dense
    finally:
        capi.check(L.dq_set_tuning(b"product_dense", 1))


@pytest.mark.parametrize("kind", ["tfim", "xxz", "random"])
def test_pauli_sum_device_builder(d, kind):
    """The device Pauli-sum builder against the host builders: same diagonals once all-zero ones are pruned, same
    values; the structural set before pruning contains them; FP32 form within FP32 rounding."""
    n = 9
    if kind == "tfim":
        terms = Hm.tfim_terms(n, 1.1, 0.6)
        _, _, offs, vals = Hm.tfim(n, 1.1, 0.6)
    elif kind == "xxz":
        terms = Hm.xxz_terms(n, seed=5)
        _, _, offs, vals = Hm.xxz(n, seed=5)
    else:
        _, _, offs, vals, terms = Hm.pauli_sum(n, 30, np.random.default_rng(6))
    x, z, c = Hm.pauli_masks(terms)
    D = d.pauli_sum_device(n, x, z, c)
    assert set(offs) <= set(D.diagonals) and D.diagonals == sorted(D.diagonals)
    P = d.prune(D)
    nonzero = [o for o, v in zip(offs, vals) if np.any(v != 0)]
    assert P.diagonals == nonzero
    ref = O.to_dense(O.Dia(1 << n, 1 << n, offs, vals))
    assert rel(d.to_numpy(P.to_host()), ref) < 1e-15
    D32 = d.pauli_sum_device(n, x, z, c, precision="fp32")
    assert D32.precision == "fp32" and D32.diagonals == D.diagonals
    assert rel(d.to_numpy(d.prune(D32).to_host()), ref) < 1e-6


@pytest.mark.parametrize("rows,cols,offsets,ld_pad", [
    (70, 70, [0, 1, -1, 33, -69, 69], 0),
    (45, 130, [-44, -3, 0, 5, 64, 100, 129], 13),   # wide; a row stride beyond cols
    (200, 37, [-199, -150, -64, -1, 0, 36], 0),     # tall
    (1, 1, [0], 0),
    (33, 65, [], 3),                                 # all zero: no diagonals
    (1000, 1000, list(range(-999, 1000, 37)), 0),    # many tiles, diagonals crossing tile edges
])
@pytest.mark.parametrize("path,tile", [(0, 0), (1, 0), (2, 0), (2, 1), (2, 2), (2, 3)])
def test_dense_device_conversions_vs_oracle(d, rows, cols, offsets, ld_pad, path, tile):
    """from_dense_device: the oracle's diagonal set (ascending nonzero diagonals; -0.0 counts as zero) and the exact
    values; to_dense_device gives back the dense matrix bit for bit and writes nothing outside rows x cols. Each
    kernel family (automatic choice, per-row, tiled with each tile shape)."""
    from diaq_amd import capi
    capi.check(capi.lib().dq_set_tuning(b"dense_convert", path))
    capi.check(capi.lib().dq_set_tuning(b"dense_convert_tile", tile))
    try:
        dense_conversions_case(d, rows, cols, offsets, ld_pad)
    finally:
        capi.check(capi.lib().dq_set_tuning(b"dense_convert", 0))
        capi.check(capi.lib().dq_set_tuning(b"dense_convert_tile", 0))


def dense_conversions_case(d, rows, cols, offsets, ld_pad):
    import torch
    rng = np.random.default_rng(rows * 7 + cols)
    D = banded(rows, offsets, rng, cols)
    if rows >= 70 and cols >= 70:
        k = np.arange(min(rows, cols) - 3)
        D[k + 3, k] = -0.0                      # a diagonal of negative zeros is not stored
        D[(min(rows, cols) - 1) // 2 + 2, (min(rows, cols) - 1) // 2] = 1e-300j  # one tiny imaginary element
    M = O.from_dense(D)
    ld = cols + ld_pad
    Tbig = torch.full((rows, ld), complex(3.0, -5.0), dtype=torch.complex128)
    Tbig[:, :cols] = torch.from_numpy(D)
    T = Tbig.to("cuda")
    A = d.from_dense_device(T.data_ptr(), rows, cols, ld)
    assert A.diagonals == M.offsets
    H = d.to_numpy(A.to_host())
    assert np.array_equal(H, D)
    out = torch.full((rows, ld), complex(7.0, 7.0), dtype=torch.complex128, device="cuda")
    d.to_dense_device(A, out.data_ptr(), ld)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert np.array_equal(got[:, :cols], D)
    assert np.all(got[:, cols:] == complex(7.0, 7.0))
    A32 = d.from_dense_device(T.data_ptr(), rows, cols, ld, precision="fp32")
    assert A32.precision == "fp32" and A32.diagonals == M.offsets
    out.fill_(complex(7.0, 7.0))
    d.to_dense_device(A32, out.data_ptr(), ld)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy()[:, :cols], D.astype(np.complex64).astype(np.complex128))


@pytest.mark.parametrize("path", [0, 2])
def test_dense_device_round_trip_full_size(d, path):
    """8192 x 8192 (1 GiB of complex128) with 300 random diagonals: dense -> DIA -> dense is the identity on the
    device, the diagonal set is exactly the planted one, and SpMV of the converted matrix matches the dense product
    (automatic kernel choice -- the per-row kernels here -- and the tiled kernels)."""
    from diaq_amd import capi
    capi.check(capi.lib().dq_set_tuning(b"dense_convert", path))
    try:
        dense_round_trip_case(d)
    finally:
        capi.check(capi.lib().dq_set_tuning(b"dense_convert", 0))


def dense_round_trip_case(d):
    import torch
    n = 8192
    g = torch.Generator(device="cuda").manual_seed(11)
    offs = sorted(set(np.random.default_rng(3).integers(-(n - 1), n, 300).tolist()) | {0, n - 1, -(n - 1)})
    T = torch.zeros((n, n), dtype=torch.complex128, device="cuda")
    for o in offs:
        L = n - abs(o)
        v = torch.randn(L, dtype=torch.complex128, device="cuda", generator=g)
        idx = torch.arange(L, device="cuda")
        r, c = (idx, idx + o) if o >= 0 else (idx - o, idx)
        T[r, c] = v
    A = d.from_dense_device(T.data_ptr(), n, n)
    assert A.diagonals == offs
    out = torch.empty_like(T)
    d.to_dense_device(A, out.data_ptr())
    torch.cuda.synchronize()
    assert torch.equal(out, T)
    x = torch.randn(n, dtype=torch.complex128, device="cuda", generator=g)
    xr, xi = x.real.contiguous(), x.imag.contiguous()
    yr, yi = torch.empty_like(xr), torch.empty_like(xi)
    d.spmv_device(A, xr.data_ptr(), xi.data_ptr(), yr.data_ptr(), yi.data_ptr(), 0)
    torch.cuda.synchronize()
    ref = (T @ x).cpu().numpy()
    assert rel((yr + 1j * yi).cpu().numpy(), ref) < TOL


@pytest.mark.parametrize("n_times", [0, 1, 2, 3, 6])
def test_spmv_ntimes_device_resident_and_in_place(d, n_times):
    """spMV_ntimes on resident device vectors: the same bits as the host-vector entry point (which runs the same
    kernels), within 1e-12 of the oracle; in place (x == y) for even and odd counts; the num_times rules."""
    import torch
    rng = np.random.default_rng(40 + n_times)
    n = 3000
    A = banded(n, [0, 1, -1, 7, -200], rng) * 0.4
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    dA = d.to_device(d.from_numpy(A))
    ref = O.spmv_ntimes(O.from_dense(A), x, n_times)
    host = d.to_numpy_vector(d.spMV_ntimes(d.from_numpy(A), d.from_numpy_vector(x), n_times))
    xr = torch.tensor(x.real, device="cuda")
    xi = torch.tensor(x.imag, device="cuda")
    yr, yi = torch.empty_like(xr), torch.empty_like(xi)
    wr, wi = torch.empty_like(xr), torch.empty_like(xi)
    d.spmv_ntimes_device(dA, xr.data_ptr(), xi.data_ptr(), yr.data_ptr(), yi.data_ptr(), n_times, wr.data_ptr(),
                         wi.data_ptr())
    torch.cuda.synchronize()
    got = yr.cpu().numpy() + 1j * yi.cpu().numpy()
    assert np.array_equal(got, host)
    assert rel(got, ref) < TOL
    d.spmv_ntimes_device(dA, xr.data_ptr(), xi.data_ptr(), xr.data_ptr(), xi.data_ptr(), n_times, wr.data_ptr(),
                         wi.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(xr.cpu().numpy() + 1j * xi.cpu().numpy(), host)
    if n_times > 1:
        with pytest.raises(RuntimeError, match="work vectors"):
            d.spmv_ntimes_device(dA, xr.data_ptr(), xi.data_ptr(), yr.data_ptr(), yi.data_ptr(), n_times)
    R = d.to_device(d.from_numpy(np.ones((3, 4), complex)))
    if n_times == 0:
        with pytest.raises(RuntimeError, match="requires square"):
            d.spmv_ntimes_device(R, xr.data_ptr(), xi.data_ptr(), yr.data_ptr(), yi.data_ptr(), 0)
    else:  # no diagonals, rectangular: y = 0 for any count, no work vectors needed (src/spMV_ntimes_gpu.cu:98-105)
        Z = d.to_device(d.from_numpy(np.zeros((3, 4), complex)))
        assert Z.num_diags == 0
        zr = torch.full((3,), 7.0, dtype=torch.float64, device="cuda")
        zi = zr.clone()
        d.spmv_ntimes_device(Z, xr.data_ptr(), xi.data_ptr(), zr.data_ptr(), zi.data_ptr(), n_times)
        torch.cuda.synchronize()
        assert not bool(torch.any(zr)) and not bool(torch.any(zi))
        hz = d.to_numpy_vector(d.spMV_ntimes(d.from_numpy(np.zeros((3, 4), complex)), d.from_numpy_vector(x[:4]),
                                             n_times))
        np.testing.assert_array_equal(hz, np.zeros(3))


@pytest.mark.parametrize("model", ["tfim", "xxz"])
def test_spmv_ntimes_zero_block_skip_is_exact(d, model):
    """spMV_ntimes maps its matrix's all-zero 256-row blocks once per call and every product skips them. The skipped
    products are exact zeros, so in FP64 the result has the same bits with the skip on or off, under forced SpMV
    schedules, on resident vectors (in place too) and through the host-vector entry point; a graph-captured call
    (which runs without the map) replays to the same bits."""
    import torch
    from diaq_amd import capi
    L = capi.lib()
    q = 14 if model == "tfim" else 13
    rows, cols, offs, vals = Hm.tfim(q, 1.0, 0.7) if model == "tfim" else Hm.xxz(q, seed=4)
    H = dia_of(d, rows, cols, offs, vals)
    D = d.to_device(H)
    rng = np.random.default_rng(50)
    x = rng.standard_normal(rows) + 1j * rng.standard_normal(rows)
    x /= np.linalg.norm(x)
    ref = O.spmv_ntimes(O.Dia(rows, cols, offs, vals), x, 3)
    sid = torch.cuda.current_stream().cuda_stream

    def device_run(n_times, in_place):
        xr = torch.tensor(x.real, device="cuda")
        xi = torch.tensor(x.imag, device="cuda")
        wr, wi = torch.empty_like(xr), torch.empty_like(xi)
        yr, yi = (xr, xi) if in_place else (torch.empty_like(xr), torch.empty_like(xi))
        d.spmv_ntimes_device(D, xr.data_ptr(), xi.data_ptr(), yr.data_ptr(), yi.data_ptr(), n_times, wr.data_ptr(),
                             wi.data_ptr(), sid)
        torch.cuda.synchronize()
        return yr.cpu().numpy() + 1j * yi.cpu().numpy()

    out = {}
    try:
        for skip in (0, 1):
            capi.check(L.dq_set_tuning(b"spmv_skip_zero_blocks", skip))
            for lds, rpt in ((0, 1), (1, 4), (0, 16), (-1, 0)):
                capi.check(L.dq_set_tuning(b"spmv_lds", lds))
                capi.check(L.dq_set_tuning(b"spmv_rows_per_lane", rpt))
                out[skip, lds, rpt, "dev3"] = device_run(3, False)
                out[skip, lds, rpt, "dev3_in_place"] = device_run(3, True)
            out[skip, "host3"] = d.to_numpy_vector(d.spMV_ntimes(H, d.from_numpy_vector(x), 3))
        capi.check(L.dq_set_tuning(b"spmv_lds", -1))
        capi.check(L.dq_set_tuning(b"spmv_rows_per_lane", 0))
        capi.check(L.dq_spmv_tune(D.handle, 2, 0, None))  # also times the skip schedule (the TFIM has zero blocks)
        out["tuned", "dev3"] = device_run(3, False)
        out["tuned", "dev3_in_place"] = device_run(3, True)
    finally:
        capi.check(L.dq_set_tuning(b"spmv_skip_zero_blocks", 1))
        capi.check(L.dq_set_tuning(b"spmv_lds", -1))
        capi.check(L.dq_set_tuning(b"spmv_rows_per_lane", 0))
    base = out[0, -1, 0, "dev3"]
    for key, v in out.items():
        np.testing.assert_array_equal(v, base, err_msg=str(key))
    assert rel(base, ref) < TOL
    # graph capture: the map is not made while capturing; every replay gives the eager bits
    xr = torch.tensor(x.real, device="cuda")
    xi = torch.tensor(x.imag, device="cuda")
    yr, yi = torch.empty_like(xr), torch.empty_like(xi)
    wr, wi = torch.empty_like(xr), torch.empty_like(xi)
    cs = torch.cuda.Stream()
    cs.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=cs):
        d.spmv_ntimes_device(D, xr.data_ptr(), xi.data_ptr(), yr.data_ptr(), yi.data_ptr(), 3, wr.data_ptr(),
                             wi.data_ptr(), cs.cuda_stream)
    for _ in range(3):
        yr.fill_(7.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(yr.cpu().numpy() + 1j * yi.cpu().numpy(), base)
    with pytest.raises(RuntimeError, match="spmv_skip_zero_blocks"):
        capi.check(L.dq_set_tuning(b"spmv_skip_zero_blocks", 3))


def test_spmv_tuned_zero_block_map_is_exact_and_refreshed_on_rewrite(d):
    """dq_spmv_tune keeps the map of an FP64 matrix's all-zero 256-row blocks, and later SpMVs skip them: the same bits
    as before tuning. Rewriting the values (dq_devmat_set_packed_device) recomputes the map for the new values, so the
    blocks that became nonzero count again. An FP32 matrix keeps the same bits whether tuned or not (its SpMVs do not
    use the map)."""
    import ctypes
    import torch
    from diaq_amd import capi
    L = capi.lib()
    rows, cols, offs, vals = Hm.tfim(13, 1.0, 0.7)
    H = dia_of(d, rows, cols, offs, vals)
    rng = np.random.default_rng(60)
    x = rng.standard_normal(rows) + 1j * rng.standard_normal(rows)
    sid = torch.cuda.current_stream().cuda_stream

    def spmv(D, dt):
        xr = torch.as_tensor(x.real, device="cuda").to(dt).contiguous()
        xi = torch.as_tensor(x.imag, device="cuda").to(dt).contiguous()
        yr, yi = torch.empty_like(xr), torch.empty_like(xi)
        capi.check(L.dq_spmv_device(D.handle, xr.data_ptr(), xi.data_ptr(), yr.data_ptr(), yi.data_ptr(), sid))
        torch.cuda.synchronize()
        return (yr.double() + 1j * yi.double()).cpu().numpy()

    for prec, dt in (("fp64", torch.float64), ("fp32", torch.float32)):
        D = d.to_device(H, precision=prec)
        before = spmv(D, dt)
        capi.check(L.dq_spmv_tune(D.handle, 2, sid, None))
        np.testing.assert_array_equal(spmv(D, dt), before, err_msg=prec)
        if prec == "fp64":
            assert rel(before, O.spmv(O.Dia(rows, cols, offs, vals), x)) < TOL
            # new values, nonzero everywhere (the TFIM's zero blocks included)
            S = D.stored
            g = torch.Generator(device="cuda").manual_seed(61)
            re = torch.randn(S, dtype=torch.float64, device="cuda", generator=g)
            im = torch.randn(S, dtype=torch.float64, device="cuda", generator=g)
            capi.check(L.dq_devmat_set_packed_device(D.handle, re.data_ptr(), im.data_ptr(), sid))
            torch.cuda.synchronize()
            packed = (re + 1j * im).cpu().numpy()
            lens = [len(v) for v in vals]
            new_vals = np.split(packed, np.cumsum(lens)[:-1])
            assert rel(spmv(D, dt), O.spmv(O.Dia(rows, cols, offs, new_vals), x)) < TOL


def test_captured_ntimes_after_value_rewrite_of_a_tuned_matrix(d):
    """A graph captured around spMV_ntimes on a tuned TFIM reads the zero-block map kept in the handle. Rewriting the
    values afterwards so that a zero block becomes nonzero recomputes that map in place, so replays of the old graph
    match the eager run with the new values and the oracle."""
    import torch
    from diaq_amd import capi
    L = capi.lib()
    q = 12
    rows, cols, offs, vals = Hm.tfim(q, 1.0, 0.7)
    D = d.to_device(dia_of(d, rows, cols, offs, vals))
    sid = torch.cuda.current_stream().cuda_stream
    capi.check(L.dq_spmv_tune(D.handle, 2, sid, None))
    rng = np.random.default_rng(80)
    x = rng.standard_normal(rows) + 1j * rng.standard_normal(rows)
    x /= np.linalg.norm(x)
    xr = torch.tensor(x.real, device="cuda")
    xi = torch.tensor(x.imag, device="cuda")
    yr, yi = torch.empty_like(xr), torch.empty_like(xi)
    wr, wi = torch.empty_like(xr), torch.empty_like(xi)
    cs = torch.cuda.Stream()
    cs.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=cs):
        d.spmv_ntimes_device(D, xr.data_ptr(), xi.data_ptr(), yr.data_ptr(), yi.data_ptr(), 3, wr.data_ptr(),
                             wi.data_ptr(), cs.cuda_stream)
    graph.replay()
    torch.cuda.synchronize()
    np.testing.assert_allclose(yr.cpu().numpy() + 1j * yi.cpu().numpy(),
                               O.spmv_ntimes(O.Dia(rows, cols, offs, vals), x, 3), rtol=0, atol=1e-12)
    # the +2^8 diagonal holds -h on the rows whose bit 8 is 0, so its rows 256..511 form an all-zero 256-row block;
    # give that block nonzero values and keep the rest
    k = offs.index(1 << 8)
    assert not np.any(vals[k][256:512])
    new_vals = [np.array(v, copy=True) for v in vals]
    new_vals[k][256:512] = 0.25 + 0.5j
    packed = np.concatenate(new_vals)
    re = torch.tensor(packed.real, device="cuda")
    im = torch.tensor(packed.imag, device="cuda")
    capi.check(L.dq_devmat_set_packed_device(D.handle, re.data_ptr(), im.data_ptr(), sid))
    torch.cuda.synchronize()
    ref = O.spmv_ntimes(O.Dia(rows, cols, offs, new_vals), x, 3)
    out = {}
    for tag in ("replay", "replay2"):
        yr.fill_(7.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        out[tag] = yr.cpu().numpy() + 1j * yi.cpu().numpy()
    d.spmv_ntimes_device(D, xr.data_ptr(), xi.data_ptr(), yr.data_ptr(), yi.data_ptr(), 3, wr.data_ptr(),
                         wi.data_ptr(), sid)
    torch.cuda.synchronize()
    eager = yr.cpu().numpy() + 1j * yi.cpu().numpy()
    assert rel(eager, ref) < TOL
    np.testing.assert_array_equal(out["replay"], eager)
    np.testing.assert_array_equal(out["replay2"], eager)


def test_spmv_tuned_zero_block_map_rectangular_and_ragged(d):
    """A rectangular matrix whose diagonals start and end inside 256-row blocks, with zeroed row and column ranges:
    tuning keeps its zero-block map, and the FP64 SpMV after tuning has the same bits as before and matches the
    oracle; so do spMV_ntimes-style sequences on a square part (host entry point)."""
    import torch
    from diaq_amd import capi
    L = capi.lib()
    rng = np.random.default_rng(70)
    m, n = 3000, 5000
    A = banded(m, [-2900, -300, -1, 0, 7, 700, 4999], rng, cols=n)
    A[256:1100, :] = 0       # rows of several whole blocks, and parts of two more
    A[:, 2000:2600] = 0      # a column range: different rows on each diagonal
    A[2999, :] = 0
    D = d.to_device(d.from_numpy(A))
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    sid = torch.cuda.current_stream().cuda_stream

    def spmv():
        xr = torch.tensor(x.real, device="cuda")
        xi = torch.tensor(x.imag, device="cuda")
        yr = torch.empty(m, dtype=torch.float64, device="cuda")
        yi = torch.empty_like(yr)
        capi.check(L.dq_spmv_device(D.handle, xr.data_ptr(), xi.data_ptr(), yr.data_ptr(), yi.data_ptr(), sid))
        torch.cuda.synchronize()
        return yr.cpu().numpy() + 1j * yi.cpu().numpy()

    before = spmv()
    capi.check(L.dq_spmv_tune(D.handle, 2, sid, None))
    np.testing.assert_array_equal(spmv(), before)
    assert rel(before, A @ x) < TOL
    S = A[:, :m].copy()  # a square part for repeated products
    Hs = d.from_numpy(S)
    xs = x[:m] / np.linalg.norm(x[:m])
    got = {}
    for skip in (0, 1):
        capi.check(L.dq_set_tuning(b"spmv_skip_zero_blocks", skip))
        got[skip] = d.to_numpy_vector(d.spMV_ntimes(Hs, d.from_numpy_vector(xs), 4))
    capi.check(L.dq_set_tuning(b"spmv_skip_zero_blocks", 1))
    np.testing.assert_array_equal(got[0], got[1])
    assert rel(got[1], np.linalg.matrix_power(S, 4) @ xs) < TOL


def test_reference_api_vs_numpy_script(d):
    """python_tests/test_api_vs_numpy.py:12-92 through diaq_gpu: the same generator (seed 3), the same random
    matrices drawn in the same order, and the same checks and bounds (max abs error / max(1, max |expected|) <= 1e-10,
    expm 1e-8) for the functions this module has (round trip, product, spMV, trace, pow, the three Kronecker-times-
    vector functions, expm, expm_negative). The draws for the host-only functions (add, adjoint, kroneckerProduct,
    ...) still happen, so every later input is the reference's."""
    import scipy.linalg
    rng = np.random.default_rng(3)

    def rand(r, c, n_diags=None, fill=0.7):
        A = np.zeros((r, c), complex)
        count = min(n_diags or rng.integers(1, 6), r + c - 1)
        for o in rng.choice(np.arange(-(r - 1), c), size=count, replace=False):
            k = np.arange(max(0, -o), min(r, c - o))
            keep = rng.random(len(k)) < fill
            A[k[keep], k[keep] + o] = rng.standard_normal(keep.sum()) + 1j * rng.standard_normal(keep.sum())
        return A

    worst = {}

    def check(name, got, want, tol=1e-10):
        got, want = np.asarray(got), np.asarray(want)
        if got.shape != want.shape and got.size == want.size:
            got = got.reshape(want.shape)
        assert got.shape == want.shape and np.all(np.isfinite(got)), name
        err = float(np.max(np.abs(got - want), initial=0.0)) / max(1.0, float(np.max(np.abs(want), initial=0.0)))
        worst[name] = max(worst.get(name, 0.0), err)
        assert err <= tol, (name, err)

    M, N = d.from_numpy, lambda m: np.array(d.to_numpy(m))
    V, NV = d.from_numpy_vector, lambda v: np.array(d.to_numpy_vector(v))
    for _ in range(25):
        r, c, k = (int(v) for v in rng.integers(2, 12, size=3))
        A, _A2, B = rand(r, c), rand(r, c), rand(c, k)
        S = rand(r, r)
        x = rng.standard_normal(c) + 1j * rng.standard_normal(c)
        check("round trip", N(M(A)), A)
        check("product", N(d.product(M(A), M(B))), A @ B)
        check("spMV", NV(d.spMV(M(A), V(x))), A @ x)
        check("trace", d.trace(M(S)), np.trace(S))
        check("pow 3", N(d.pow(M(S / 4), 3)), np.linalg.matrix_power(S / 4, 3))
        check("pow 4", N(d.pow(M(S / 4), 4)), np.linalg.matrix_power(S / 4, 4))
        rand(int(rng.integers(2, 5)), int(rng.integers(2, 5)))  # kroneckerProduct's operand (host-only function)
        s1, s2 = (int(v) for v in rng.integers(0, 3, size=2))
        I1, I2 = np.eye(1 << s1), np.eye(1 << s2)
        Sq = rand(r, r)
        big = np.kron(np.kron(I1, Sq), I2)
        xb = rng.standard_normal(big.shape[1]) + 1j * rng.standard_normal(big.shape[1])
        y = V(np.zeros(big.shape[0], complex))
        check("identity_kron_matrix_kron_identity_times_vector",
              NV(d.identity_kron_matrix_kron_identity_times_vector(s1, M(Sq), s2, V(xb), y)), big @ xb)
        y = V(np.zeros(Sq.shape[0] << s1, complex))
        xi = rng.standard_normal(Sq.shape[1] << s1) + 0j
        check("identity_kron_matrix_times_vector", NV(d.identity_kron_matrix_times_vector(s1, M(Sq), V(xi), y)),
              np.kron(I1, Sq) @ xi)
        y = V(np.zeros(Sq.shape[0] << s2, complex))
        xm = rng.standard_normal(Sq.shape[1] << s2) + 0j
        check("matrix_kron_identity_times_vector", NV(d.matrix_kron_identity_times_vector(M(Sq), s2, V(xm), y)),
              np.kron(Sq, I2) @ xm)
    for _ in range(5):
        n = int(rng.integers(2, 10))
        H = rand(n, n, 3)
        H = (H + H.conj().T) / 8
        check("expm", N(d.expm(M(H), 100, 1.0, 1e-14)), scipy.linalg.expm(H), 1e-8)
        check("expm_negative", N(d.expm_negative(M(H), 0.3, 100, 1.0, 1e-14)), scipy.linalg.expm(-1j * 0.3 * H),
              1e-8)
    assert len(worst) == 11  # every named check ran


def test_pow_expm_on_device_matrices(d):
    """pow / expm / expm_negative on a DeviceMatrix return a DeviceMatrix with the same offsets and the same bits as
    the host-matrix forms (the same device computation, without the download)."""
    import scipy.linalg
    rows, cols, offs, vals = Hm.tfim(8, 1.0, 0.7)
    H = dia_of(d, rows, cols, offs, vals)
    D = d.to_device(H)
    for host, dev in ((d.expm_negative(H, 0.3), d.expm_negative(D, 0.3)), (d.expm(H), d.expm(D)),
                      (d.pow(H, 3), d.pow(D, 3))):
        assert isinstance(dev, type(D))
        assert list(host.dIndices) == list(dev.diagonals)
        np.testing.assert_array_equal(d.to_numpy(dev.to_host()), d.to_numpy(host))
    dense = O.to_dense(O.Dia(rows, cols, offs, vals))
    assert rel(d.to_numpy(d.expm_negative(D, 0.3).to_host()), scipy.linalg.expm(-1j * 0.3 * dense)) < 1e-12
