"""CPU tests of the C ABI library: it loads, exports every symbol include/diaq_hip.h declares, and its host-only
entry points (structural dry-run, viability rule, argument checks) agree with the oracle. No GPU compute here."""
import ctypes
import os
import re

import numpy as np
import pytest

from diaq_amd import capi
from diaq_amd import hamiltonians as Hm
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "diaq_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dq_[a-z0-9_]+)\s*\(", text)))


def test_header_and_ctypes_table_agree():
    assert declared_symbols() == sorted(capi.SIGNATURES)


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(capi.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    assert capi.lib().dq_version().startswith(b"diaq-mi355x")


def dryrun_capi(rows, cols, oa, ob):
    L = capi.lib()
    oa = np.asarray(oa, np.int32)
    ob = np.asarray(ob, np.int32)
    n_res, n_pairs = ctypes.c_int32(), ctypes.c_int64()
    capi.check(L.dq_product_dryrun(rows, cols, len(oa), capi.ptr(oa, ctypes.c_int32), len(ob),
                                   capi.ptr(ob, ctypes.c_int32), ctypes.byref(n_res), ctypes.byref(n_pairs), None, None,
                                   None, None))
    res = np.zeros(n_res.value, np.int32)
    ps = np.zeros(n_res.value + 1, np.int64)
    sa = np.zeros(max(1, n_pairs.value), np.int32)
    sb = np.zeros(max(1, n_pairs.value), np.int32)
    capi.check(L.dq_product_dryrun(rows, cols, len(oa), capi.ptr(oa, ctypes.c_int32), len(ob),
                                   capi.ptr(ob, ctypes.c_int32), ctypes.byref(n_res), ctypes.byref(n_pairs),
                                   capi.ptr(res, ctypes.c_int32), capi.ptr(ps, ctypes.c_int64),
                                   capi.ptr(sa, ctypes.c_int32), capi.ptr(sb, ctypes.c_int32)))
    pairs = [list(zip(sa[ps[k]:ps[k + 1]].tolist(), sb[ps[k]:ps[k + 1]].tolist())) for k in range(n_res.value)]
    return res.tolist(), pairs


@pytest.mark.parametrize("rows,cols,oa,ob", [
    (8, 8, [0, -1, 1], [0, 2, -3]),
    (6, 5, [-5, -2, 0, 3, 8], [-7, 0, 4]),       # A is 6 x 9, B is 9 x 5
    (1024, 1024, [3, -1, 0, 5, -1000], [1000, -3, 0]),  # unsorted slots
    (4, 4, [], [0, 1]),
    (1 << 24, 1 << 24, [5, -(1 << 23), 0, 77, -3], [1 << 22, -77, 0, 3]),  # offset range too wide for bucket counts
    (1 << 25, 3, [0, 2, -(1 << 24)], [1, 0, -1, 2]),                      # tall, wide range, out-of-range sums
    (300, 300, list(range(-299, 300, 7)), list(range(-150, 151, 3))),     # many pairs per result diagonal
])
def test_dryrun_bit_exact_vs_oracle(rows, cols, oa, ob):
    res, pairs = dryrun_capi(rows, cols, oa, ob)
    o_res, o_pairs = O.product_dryrun(oa, ob, rows, cols)
    assert res == o_res
    assert pairs == [list(p) for p in o_pairs]


def test_dryrun_tfim20_hh():
    offs = sorted([0] + [s * (1 << k) for k in range(20) for s in (1, -1)])
    res, pairs = dryrun_capi(1 << 20, 1 << 20, offs, offs)
    assert len(res) == 727 and sum(len(p) for p in pairs) == 41 * 41 - 2
    o_res, _ = O.product_dryrun(offs, offs, 1 << 20, 1 << 20)
    assert res == o_res


@pytest.mark.parametrize("args", [(4096, 4096, 5, 5 * 4096, 1, 16, 8), (4096, 4096, 200, 200 * 4096, 1, 16, 8),
                                  (4096, 4096, 9, 9 * 4096, 64, 16, 8), (1000, 1000, 4, 400, 1, 16, 8),
                                  (65536, 65536, 1024, 1024 * 60000, 16, 8, 4)])
def test_dia_viability_vs_oracle(args):
    v = capi.DiaViability()
    capi.check(capi.lib().dq_dia_viability(*args, ctypes.byref(v)))
    o = O.dia_viability(*args)
    assert bool(v.viable) == o["viable"] and bool(v.storage_ok) == o["storage_ok"]
    assert bool(v.diagonals_ok) == o["diagonals_ok"] and v.diagonal_budget == o["diagonal_budget"]
    assert v.dia_bytes == o["dia_bytes"] and v.csr_bytes == o["csr_bytes"] and v.density == o["density"]
    assert v.reason.decode() == o["reason"]


def test_argument_errors_need_no_gpu():
    L = capi.lib()
    out = ctypes.c_void_p()
    offs = np.array([5], np.int32)
    rc = L.dq_devmat_zeros(4, 4, 1, capi.ptr(offs, ctypes.c_int32), capi.DQ_F64, ctypes.byref(out))
    assert rc == capi.DQ_ERR_ARG and b"outside the matrix" in L.dq_last_error()
    rc = L.dq_devmat_zeros(4, 4, 1, capi.ptr(np.array([0], np.int32), ctypes.c_int32), 7, ctypes.byref(out))
    assert rc == capi.DQ_ERR_ARG
    assert L.dq_set_tuning(b"spmv_rows_per_lane", 3) == capi.DQ_ERR_ARG
    assert L.dq_set_tuning(b"no_such_knob", 1) == capi.DQ_ERR_ARG
    for key, bad in ((b"evo_lanes", 9), (b"evo_diag_batch", 3), (b"product_skip_zero_blocks", 2)):
        assert L.dq_set_tuning(key, bad) == capi.DQ_ERR_ARG and key in L.dq_last_error()
    assert L.dq_spmv_device(None, None, None, None, None, 0) == capi.DQ_ERR_ARG
    assert L.dq_spmm_dia_device(0, 0, -1, 4, 4, 0, 1, 0, 0, 0) == capi.DQ_ERR_ARG


def test_binding_shape_errors_before_any_device_work():
    import diaq_amd  # noqa: F401
    import diaq_gpu as d
    A = d.from_numpy(np.eye(4, dtype=complex))
    B = d.from_numpy(np.eye(5, dtype=complex))
    with pytest.raises(RuntimeError, match="dimension mismatch"):
        d.product(A, B)
    with pytest.raises(RuntimeError, match="dimension mismatch"):
        d.spMV(A, d.from_numpy_vector(np.ones(5, complex)))
    with pytest.raises(RuntimeError, match="num_times must be >= 0"):
        d.spMV_ntimes(A, d.from_numpy_vector(np.ones(4, complex)), -1)
    with pytest.raises(RuntimeError, match="negative exponent"):
        d.pow(A, -1)
    R = d.from_numpy(np.ones((3, 4), complex))
    with pytest.raises(RuntimeError, match="must be square"):
        d.hamiltonian_simulation_cuda(R, d.from_numpy_vector(np.ones(4, complex)), 1.0, 1)
    with pytest.raises(ValueError):
        d.to_device(A, precision="fp16")


def test_binding_host_container_matches_reference_layout():
    import diaq_amd  # noqa: F401
    import diaq_gpu as d
    rng = np.random.default_rng(0)
    D = rng.standard_normal((20, 20)) * (rng.random((20, 20)) < 0.2) + 0j
    A = d.from_numpy(D)
    M = O.from_dense(D)
    assert list(A.dIndices) == M.offsets
    starts, total = O.padded_offsets(M.offsets, 20, 20)
    assert list(A.diagOffsets) == starts and A.totalLength == total
    np.testing.assert_array_equal(d.to_numpy(A), D)
    assert list(d.diaQMatrix(7, 5).dIndices) == O.symmetric_order(5)
    I = d.identityMatrix(6)
    np.testing.assert_array_equal(d.to_numpy(I), np.eye(6))
    v = d.from_numpy_vector(np.arange(5) + 1j)
    np.testing.assert_array_equal(d.to_numpy_vector(v), np.arange(5) + 1j)


def test_pauli_builder_argument_errors_need_no_gpu():
    L = capi.lib()
    out = ctypes.c_void_p()
    x = np.array([1, 2], np.uint64)
    z = np.array([0, 8], np.uint64)  # bit 3 beyond 3 qubits
    cr = np.array([1.0, 0.5])
    ci = np.zeros(2)
    args = (capi.ptr(x, ctypes.c_uint64), capi.ptr(z, ctypes.c_uint64), capi.ptr(cr), capi.ptr(ci))
    rc = L.dq_devmat_from_pauli_sum(3, 2, *args, capi.DQ_F64, 0, ctypes.byref(out))
    assert rc == capi.DQ_ERR_ARG and b"beyond n_qubits" in L.dq_last_error()
    rc = L.dq_devmat_from_pauli_sum(0, 2, *args, capi.DQ_F64, 0, ctypes.byref(out))
    assert rc == capi.DQ_ERR_ARG and b"n_qubits" in L.dq_last_error()
    rc = L.dq_devmat_from_pauli_sum(3, 2, *args, 7, 0, ctypes.byref(out))
    assert rc == capi.DQ_ERR_ARG


def test_pauli_terms_match_the_chain_builders():
    """tfim_terms / xxz_terms describe the same matrices as tfim() / xxz() (checked through the numpy Pauli-sum
    builder's rule on dense matrices)."""
    n = 5
    for terms, dia in ((Hm.tfim_terms(n, 1.3, 0.4), Hm.tfim(n, 1.3, 0.4)), (Hm.xxz_terms(n, seed=2), Hm.xxz(n, seed=2))):
        N = 1 << n
        dense = np.zeros((N, N), complex)
        r = np.arange(N)
        for c, ops in terms:
            x, z, _ = Hm.pauli_masks([(c, ops)])
            par = np.bitwise_count(r & z[0]) & 1
            phase = (1j ** int(np.bitwise_count(np.uint64(x[0] & z[0])))) * (1.0 - 2.0 * par)
            dense[r ^ x[0], r] += c * phase
        ref = np.zeros((N, N), complex)
        for o, v in zip(dia[2], dia[3]):
            k = np.arange(len(v))
            rows, cols = (k, k + o) if o >= 0 else (k - o, k)
            ref[rows, cols] = v
        np.testing.assert_allclose(dense, ref, atol=1e-14)


def test_dense_conversion_argument_errors_need_no_gpu():
    L = capi.lib()
    out = ctypes.c_void_p()
    assert L.dq_devmat_from_dense_device(4, 4, None, 4, capi.DQ_F64, 0, ctypes.byref(out)) == capi.DQ_ERR_ARG
    assert b"dense is NULL" in L.dq_last_error()
    assert L.dq_devmat_from_dense_device(4, 4, 4096, 3, capi.DQ_F64, 0, ctypes.byref(out)) == capi.DQ_ERR_ARG
    assert b"ld must be" in L.dq_last_error()
    assert L.dq_devmat_from_dense_device(4, 4, 4104, 4, capi.DQ_F64, 0, ctypes.byref(out)) == capi.DQ_ERR_ARG
    assert b"aligned" in L.dq_last_error()
    assert L.dq_devmat_from_dense_device(4, 4, 4096, 4, 9, 0, ctypes.byref(out)) == capi.DQ_ERR_ARG
    assert L.dq_devmat_to_dense_device(None, 4096, 4, 0) == capi.DQ_ERR_ARG


def test_gate_plan_argument_errors_need_no_gpu():
    L = capi.lib()
    out = ctypes.c_void_p()
    q = (ctypes.c_uint32 * 1)(0)
    assert L.dq_gate_plan_create(12, 0, None, q, ctypes.byref(out)) == capi.DQ_ERR_ARG
    gates = (ctypes.c_void_p * 1)(None)
    assert L.dq_gate_plan_create(12, 1, gates, q, ctypes.byref(out)) == capi.DQ_ERR_ARG
    assert b"NULL gate" in L.dq_last_error()
    assert L.dq_set_tuning(b"gates_tile_qubits", 13) == capi.DQ_ERR_ARG
    n = ctypes.c_int()
    assert L.dq_gate_plan_passes(None, ctypes.byref(n)) == capi.DQ_ERR_ARG
    assert L.dq_gate_plan_run(None, None, None, 1, 0) == capi.DQ_ERR_ARG


def test_permutation_and_dense_matrix_classes():
    """permutationMatrix (src/specialMatrices.hpp:26-51) and the denseMatrix container (src/denseMatrix.hpp)."""
    import diaq_amd  # noqa: F401
    import diaq_gpu as d
    order = [3, 0, 4, 1, 2, 7, 6, 5]
    P = d.permutationMatrix(8, order)
    dense = np.zeros((8, 8))
    dense[np.arange(8), order] = 1
    np.testing.assert_array_equal(d.to_numpy(P), dense)
    assert list(P.dIndices) == O.from_dense(dense).offsets  # ascending nonzero diagonals, as the dense route gives
    rng = np.random.default_rng(4)
    R = rng.standard_normal((6, 6)) * (rng.random((6, 6)) < 0.4)
    I = rng.standard_normal((6, 6)) * (rng.random((6, 6)) < 0.4)
    D = d.denseMatrix(6, R.ravel(), I.ravel())
    assert list(D.matrix_shape) == [6, 6] and list(D.shape) == [6, 6]
    assert D.get(2, 3) == complex(R[2, 3], I[2, 3]) and D.getItem([2, 3]) == complex(R[2, 3], I[2, 3])
    D.set(1, 1, 2 - 3j)
    R[1, 1], I[1, 1] = 2, -3
    A = D.convertToDiaQ()
    np.testing.assert_array_equal(d.to_numpy(A), R + 1j * I)
    assert list(A.dIndices) == O.from_dense(R + 1j * I).offsets
    assert abs(D.sparsity() - 100.0 * np.mean((R == 0) & (I == 0))) < 1e-12
    B = D.convertToDiaQ([4, 9])  # the same entries read as a 4 x 9 matrix
    np.testing.assert_array_equal(d.to_numpy(B), (R + 1j * I).reshape(4, 9))
    with pytest.raises(ValueError):
        D.convertToDiaQ([5, 5])
    with pytest.raises(ValueError):
        d.denseMatrix(3, np.ones(8))


def test_host_dia_sparsity_dense_and_scaled_add():
    """diaQMatrix.sparsity / convertToDenseMatrix / addMatrixScaled (src/diaQMatrix.hpp:1061-1142, 1361-1375)."""
    import diaq_amd  # noqa: F401
    import diaq_gpu as d
    rng = np.random.default_rng(5)
    A = (rng.standard_normal((7, 7)) + 1j * rng.standard_normal((7, 7))) * (rng.random((7, 7)) < 0.3)
    B = (rng.standard_normal((7, 7)) + 1j * rng.standard_normal((7, 7))) * (rng.random((7, 7)) < 0.3)
    dA, dB = d.from_numpy(A), d.from_numpy(B)
    assert abs(dA.sparsity() - (1 - np.count_nonzero(A) / 49)) < 1e-15
    Dn = dA.convertToDenseMatrix()
    assert list(Dn.matrix_shape) == [7, 7] and Dn.get(3, 4) == A[3, 4]
    dA.addMatrixScaled(dB, 0.5, -2.0)
    np.testing.assert_allclose(d.to_numpy(dA), A + (0.5 - 2j) * B, rtol=0, atol=1e-15)
    assert list(dA.dIndices) == sorted(set(O.from_dense(A).offsets) | set(O.from_dense(B).offsets))
