"""Pins for the CPU oracle: every oracle function is checked against something other than itself -- the reference's
own fixtures (tests/golden/), dense numpy/scipy, closed forms, or brute force on tiny inputs."""
import json
import math
import os

import numpy as np
import pytest
import scipy.linalg

from oracle import oracle as O
from diaq_amd import hamiltonians as Hm

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rand_dia(rows, cols, offsets, rng, fill=1.0):
    vals = []
    for d in offsets:
        n = O.diagonal_length(d, rows, cols)
        v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        if fill < 1.0:
            v[rng.random(n) > fill] = 0
        vals.append(v)
    return O.Dia(rows, cols, list(offsets), vals)


def dec(e):
    return (np.array(e["re"]) + 1j * np.array(e["im"])).reshape(e["shape"])


# ---------------------------------------------------------------------------------------------------------------------
def test_diagonal_length_brute():
    for rows, cols in [(5, 5), (3, 6), (6, 2), (1, 7), (9, 4)]:
        I, J = np.indices((rows, cols))
        for d in range(-(rows + 1), cols + 2):
            assert O.diagonal_length(d, rows, cols) == int(np.sum(J - I == d)), (rows, cols, d)


def test_padded_offsets_closed_form():
    offs = [0, -1, 1, -5, 7]
    starts, total = O.padded_offsets(offs, 20, 20)
    assert all(s % 8 == 0 for s in starts)
    for k in range(len(offs) - 1):
        assert starts[k + 1] - starts[k] == 8 * math.ceil(O.diagonal_length(offs[k], 20, 20) / 8)
    assert total == starts[-1] + O.diagonal_length(offs[-1], 20, 20)


def test_symmetric_order():
    assert O.symmetric_order(5) == [0, -1, 1, -2, 2]


def test_dense_roundtrip_ascending():
    rng = np.random.default_rng(1)
    for shape in [(7, 7), (4, 9), (9, 4)]:
        A = rng.standard_normal(shape) * (rng.random(shape) < 0.3)
        M = O.from_dense(A)
        assert M.offsets == sorted(M.offsets)
        assert np.array_equal(O.to_dense(M), A)
        for d in M.offsets:
            assert np.any(np.diagonal(A, d) != 0)


# ---------------------------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("shape,offs", [((64, 64), [-3, -1, 0, 2, 5]), ((50, 70), [-9, 0, 7, 30]),
                                        ((70, 50), [-40, -2, 0, 3]), ((33, 33), [4]), ((33, 33), [-32])])
def test_spmv_vs_dense(shape, offs):
    rng = np.random.default_rng(2)
    M = rand_dia(*shape, offs, rng)
    x = rng.standard_normal(shape[1]) + 1j * rng.standard_normal(shape[1])
    np.testing.assert_allclose(O.spmv(M, x), O.to_dense(M) @ x, rtol=1e-13, atol=1e-13)


def test_spmv_known_answer_spmv2():
    g = json.load(open(os.path.join(GOLD, "spmv_known.json")))
    A = dec(g["A"])
    M = O.from_dense(A)
    assert M.offsets == g["printed_offsets"]
    x = np.array(g["x"]["re"]) + 1j * np.array(g["x"]["im"])
    y = O.spmv(M, x)
    printed = np.array(g["y_printed"]["re"]) + 1j * np.array(g["y_printed"]["im"])
    np.testing.assert_allclose(y, printed, atol=2e-6)  # six printed digits


def test_spmv_identity_and_paulis():
    n = 5
    N = 1 << n
    x = Hm.random_state(N, 3)
    np.testing.assert_array_equal(O.spmv(O.identity(N), x), x)
    # X on qubit 2 as a DIA matrix from the TFIM builder with J = 0, h = -1: sum_k X_k
    rows, cols, offs, vals = Hm.tfim(n, J=0.0, h=-1.0)
    M = O.Dia(rows, cols, offs, vals)
    idx = np.arange(N)
    expect = sum(x[idx ^ (1 << k)] for k in range(n))
    np.testing.assert_allclose(O.spmv(M, x), expect, atol=1e-14)


def test_spmv_ntimes_rules():
    rng = np.random.default_rng(4)
    M = rand_dia(16, 16, [-1, 0, 2], rng)
    x = rng.standard_normal(16) + 0j
    np.testing.assert_allclose(O.spmv_ntimes(M, x, 3), np.linalg.matrix_power(O.to_dense(M), 3) @ x, rtol=1e-12)
    np.testing.assert_array_equal(O.spmv_ntimes(M, x, 0), x)
    R = rand_dia(8, 16, [0, 3], rng)
    with pytest.raises(RuntimeError):
        O.spmv_ntimes(R, x, 2)
    with pytest.raises(RuntimeError):
        O.spmv_ntimes(M, x, -1)
    np.testing.assert_array_equal(O.spmv_ntimes(O.Dia(16, 16, [], []), x, 2), np.zeros(16))


# ---------------------------------------------------------------------------------------------------------------------
def test_spgemm_reference_known_answer_pairs():
    g = json.load(open(os.path.join(GOLD, "spgemm_cases.json")))
    assert len(g["cases"]) == 51
    for c in g["cases"]:
        A, B, P = dec(c["A"]), dec(c["B"]), dec(c["product"])
        C = O.product_host(O.from_dense(A), O.from_dense(B))
        assert C.offsets == c["product_offsets_pruned"], c["case"]
        np.testing.assert_allclose(O.to_dense(C), P, rtol=1e-13, atol=1e-13)


def test_dryrun_brute_force_structure():
    rng = np.random.default_rng(5)
    for rows, inner, cols in [(8, 8, 8), (6, 9, 5), (12, 5, 12)]:
        oa = sorted(rng.choice(np.arange(-(rows - 1), inner), 4, replace=False).tolist())
        ob = sorted(rng.choice(np.arange(-(inner - 1), cols), 3, replace=False).tolist())
        A = rand_dia(rows, inner, oa, rng)
        B = rand_dia(inner, cols, ob, rng)
        offs, pairs = O.product_dryrun(A.offsets, B.offsets, rows, cols)
        # brute force: which (a, b) slot pairs have a common index (i, k, j) in the dense product
        P = np.abs(O.to_dense(A)) @ np.abs(O.to_dense(B))  # no cancellation with |.|
        nz = [d for d in range(-(rows - 1), cols) if np.any(np.diagonal(P, d) != 0)]
        assert set(nz) <= set(offs)
        assert offs == sorted(offs)
        for d, pl in zip(offs, pairs):
            assert pl == sorted(pl)  # A-slot order, then B-slot order
            for sa, sb in pl:
                assert A.offsets[sa] + B.offsets[sb] == d
            if d not in nz:  # a structural diagonal with no overlapping positions stays all zero
                C = O.product(A, B)
                assert not np.any(C.vals[C.offsets.index(d)])
        total = sum(len(p) for p in pairs)
        assert total == sum(1 for a in A.offsets for b in B.offsets if -(rows - 1) <= a + b <= cols - 1)


def test_prune_constructed_cancellation():
    A = O.from_dense(np.array([[1, 1], [0, 1]], complex))
    B = O.from_dense(np.array([[1, -1], [0, 1]], complex))
    C = O.product(A, B)
    assert C.offsets == [0, 1]
    assert O.prune(C).offsets == [0]
    np.testing.assert_array_equal(O.to_dense(O.prune(C)), np.eye(2))


def test_tfim20_hh_dryrun_closed_form():
    offs = [0] + [s * (1 << k) for k in range(20) for s in (1, -1)]
    N = 1 << 20
    res, pairs = O.product_dryrun(sorted(offs), sorted(offs), N, N)
    assert len(res) == 727
    # 41 x 41 = 1681 pairs, of which +-2^19 + +-2^19 = +-2^20 fall outside [-(N-1), N-1]
    assert sum(len(p) for p in pairs) == 41 * 41 - 2
    assert Hm.stored_elements(N, N, res) == 687_341_554


# ---------------------------------------------------------------------------------------------------------------------
def test_union_add_norms_trace():
    rng = np.random.default_rng(6)
    A = rand_dia(20, 20, [-2, 0, 5], rng)
    B = rand_dia(20, 20, [-7, 0, 1, 5], rng)
    C = O.add_union(A, B)
    assert C.offsets == [-7, -2, 0, 1, 5]
    np.testing.assert_allclose(O.to_dense(C), O.to_dense(A) + O.to_dense(B), rtol=1e-15)
    assert O.frobenius(A) == pytest.approx(np.linalg.norm(O.to_dense(A)), rel=1e-14)
    assert O.trace(A) == pytest.approx(np.trace(O.to_dense(A)), rel=1e-14)
    D = O.to_dense(A)
    assert O.estimate_1norm(A) == pytest.approx(np.max(np.sum(np.abs(D.real) + np.abs(D.imag), axis=0)), rel=1e-14)


def test_tfim12_norm_and_scaling_closed_form():
    rows, cols, offs, vals = Hm.tfim(12, J=1.0, h=0.7)
    M = O.Dia(rows, cols, offs, vals)
    assert len(offs) == 25 and M.stored == 94_210
    assert O.estimate_1norm(M) == pytest.approx(11 * 1.0 + 12 * 0.7, rel=1e-14)  # 19.4
    assert O.scaling_exponent(O.estimate_1norm(M), 0.1, 1.0) == 1


def test_xxz20_structure():
    n = 20
    N = 1 << n
    offs = [0] + [s * (1 << i) for i in range(n - 1) for s in (1, -1)]
    assert len(offs) == 39 and Hm.stored_elements(N, N, offs) == 39_845_890


@pytest.mark.parametrize("builder", ["tfim", "xxz"])
def test_hamiltonian_builders_vs_kron(builder):
    n = 5
    X = np.array([[0, 1], [1, 0]], complex)
    Y = np.array([[0, -1j], [1j, 0]])
    Z = np.diag([1.0, -1.0]).astype(complex)

    def op(single, i):  # qubit i is bit i: kron order puts qubit n-1 first
        mats = [np.eye(2)] * n
        mats = list(mats)
        mats[n - 1 - i] = single
        out = np.ones((1, 1))
        for m in mats:
            out = np.kron(out, m)
        return out

    if builder == "tfim":
        J, h = 1.3, 0.7
        H = sum(-J * op(Z, i) @ op(Z, i + 1) for i in range(n - 1)) + sum(-h * op(X, i) for i in range(n))
        r, c, offs, vals = Hm.tfim(n, J, h)
    else:
        f = np.random.default_rng(0).uniform(-1, 1, n)
        H = sum(1.0 * (op(X, i) @ op(X, i + 1) + op(Y, i) @ op(Y, i + 1)) + 0.8 * op(Z, i) @ op(Z, i + 1)
                for i in range(n - 1)) + sum(f[i] * op(Z, i) for i in range(n))
        r, c, offs, vals = Hm.xxz(n, 1.0, 0.8, fields=f)
    M = O.Dia(r, c, offs, vals)
    np.testing.assert_allclose(O.to_dense(M), H, atol=1e-14)
    assert O.from_dense(H).offsets == offs


def test_product_state_matches_kron():
    n = 4
    psi = Hm.product_state(n, 7)
    assert np.linalg.norm(psi) == pytest.approx(1.0, rel=1e-14)
    rng = np.random.default_rng(7)
    th = rng.uniform(0, np.pi, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    for idx in range(1 << n):
        amp = 1.0 + 0j
        for i in range(n):
            amp *= np.cos(th[i]) if ((idx >> i) & 1) == 0 else np.exp(1j * ph[i]) * np.sin(th[i])
        assert psi[idx] == pytest.approx(amp, abs=1e-15)


# ---------------------------------------------------------------------------------------------------------------------
def test_pow_vs_matrix_power():
    rng = np.random.default_rng(8)
    A = rand_dia(48, 48, [-1, 0, 1], rng)
    A = O.Dia(48, 48, A.offsets, [v / 4 for v in A.vals])
    D = O.to_dense(A)
    for n in [0, 1, 2, 3, 7, 8]:
        np.testing.assert_allclose(O.to_dense(O.pow_gpu(A, n)), np.linalg.matrix_power(D, n), rtol=1e-12,
                                   atol=1e-14)


def test_expm_negative_vs_scipy_tfim6():
    rows, cols, offs, vals = Hm.tfim(6, 1.0, 0.7)
    M = O.Dia(rows, cols, offs, vals)
    for t in [0.05, 0.1, 0.7]:
        U = O.expm_negative(M, t)
        ref = scipy.linalg.expm(-1j * t * O.to_dense(M))
        np.testing.assert_allclose(O.to_dense(U), ref, atol=1e-11)


def test_expm_vs_scipy_and_single_diag_quirk():
    rng = np.random.default_rng(9)
    A = rand_dia(24, 24, [-2, 0, 3], rng)
    A = O.Dia(24, 24, A.offsets, [v * 0.3 for v in A.vals])
    np.testing.assert_allclose(O.to_dense(O.expm(A)), scipy.linalg.expm(O.to_dense(A)), rtol=1e-10, atol=1e-12)
    # one main diagonal: elementwise exp, a closed form
    D = rand_dia(10, 10, [0], rng)
    np.testing.assert_allclose(O.to_dense(O.expm_negative(D, 0.3)), np.diag(np.exp(-1j * 0.3 * D.vals[0])),
                               rtol=1e-14)
    np.testing.assert_allclose(O.to_dense(O.expm(D)), np.diag(np.exp(D.vals[0])), rtol=1e-14)
    # one off-diagonal: the reference writes it onto the main diagonal (kept quirk)
    S = rand_dia(10, 10, [3], rng)
    U = O.to_dense(O.expm_negative(S, 0.5))
    exp_diag = np.ones(10, complex)
    exp_diag[:7] = np.exp(-1j * 0.5 * S.vals[0])
    np.testing.assert_allclose(U, np.diag(exp_diag), rtol=1e-14)


def test_taylor_action_and_hamsim():
    rows, cols, offs, vals = Hm.xxz(6, seed=1)
    H = O.Dia(rows, cols, offs, vals)
    psi = Hm.product_state(6, 2)
    out = O.hamiltonian_simulation(H, psi, t=1.0, r=10, max_iterations=30)
    ref = scipy.linalg.expm(-1j * 1.0 * O.to_dense(H)) @ psi
    np.testing.assert_allclose(out, ref, atol=1e-12)
    assert np.linalg.norm(out) == pytest.approx(1.0, abs=1e-12)
    np.testing.assert_array_equal(O.hamiltonian_simulation(H, psi, 1.0, 0), psi)


# ---------------------------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("M,K,R,offs", [(40, 40, 5, [-3, -1, 0, 2, 5]), (30, 36, 1, [0, 7, -9]),
                                        (20, 14, 3, list(range(-13, 14, 3)))])
def test_spmm_dia_vs_dense(M, K, R, offs):
    rng = np.random.default_rng(10)
    V = rng.standard_normal((len(offs), M)) + 1j * rng.standard_normal((len(offs), M))
    A = np.zeros((M, K), complex)
    r = np.arange(M)
    for k, o in enumerate(offs):
        ok = (r + o >= 0) & (r + o < K)
        V[k, ~ok] = 0
        A[r[ok], r[ok] + o] = V[k, ok]
    B = rng.standard_normal((K, R)) + 1j * rng.standard_normal((K, R))
    np.testing.assert_allclose(O.spmm_dia(offs, V, M, K, B), A @ B, rtol=1e-13, atol=1e-13)


def test_dia_viability_cases():
    v = O.dia_viability(4096, 4096, 5, 5 * 4096, 1)
    assert v["viable"] and v["diagonal_budget"] == 128
    v = O.dia_viability(4096, 4096, 200, 200 * 4096, 1)
    assert not v["diagonals_ok"] and v["reason"].startswith("diagonals")
    v = O.dia_viability(4096, 4096, 9, 9 * 4096, 64)
    assert v["diagonal_budget"] == 8 and not v["viable"]
    v = O.dia_viability(1000, 1000, 4, 400, 1)
    assert not v["storage_ok"]
    # dia bytes 16*5*4096 + 40 vs csr 24*nnz + 8*4097: 0.9 threshold by hand
    assert v["dia_bytes"] == 16 * 4 * 1000 + 32 and v["csr_bytes"] == 24 * 400 + 8 * 1001


# ---------------------------------------------------------------------------------------------------------------------
def test_c_oracle_spmv_matches_numpy():
    from oracle import coracle
    rng = np.random.default_rng(11)
    for shape, offs in [((1000, 1000), [-300, -1, 0, 1, 17, 999]), ((500, 800), [-499, 0, 250, 799]),
                        ((800, 500), [-799, -10, 0, 499])]:
        M = rand_dia(*shape, offs, rng)
        x = rng.standard_normal(shape[1]) + 1j * rng.standard_normal(shape[1])
        o, st, re, im = coracle.packed(M.offsets, M.vals)
        ref = O.spmv(M, x)
        np.testing.assert_allclose(coracle.spmv_rows(shape[0], shape[1], o, st, re, im, x), ref, rtol=1e-13,
                                   atol=1e-13)
        yr, yi = np.empty(shape[0]), np.empty(shape[0])
        coracle.spmv_tasks_split(shape[0], shape[1], o, st, re, im, np.ascontiguousarray(x.real),
                                 np.ascontiguousarray(x.imag), yr, yi, block_size=64, n_threads=3)
        np.testing.assert_allclose(yr + 1j * yi, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("shapes,oa,ob", [
    ((300, 300, 300), [-299, -40, -1, 0, 1, 7, 299], [-299, -5, 0, 2, 150, 299]),
    ((200, 350, 120), [-199, -3, 0, 4, 349], [-349, -20, 0, 6, 119]),   # rectangular: A 200x350, B 350x120
    ((5000, 5000, 5000), [-2100, -1, 0, 1, 2100], [-4097, 0, 3, 2049]),  # several element blocks per diagonal
])
def test_c_oracle_product_tasks_matches_dense(shapes, oa, ob):
    """The restated reference CPU SpGEMM (tiles of 4 result diagonals x 2048 positions) against dense products."""
    from oracle import coracle
    m, k, n = shapes
    rng = np.random.default_rng(12)
    A = rand_dia(m, k, oa, rng)
    B = rand_dia(k, n, ob, rng)
    res, pairs = O.product_dryrun(oa, ob, m, n)
    ps = np.zeros(len(res) + 1, np.int64)
    ps[1:] = np.cumsum([len(p) for p in pairs])
    sa = np.array([a for p in pairs for a, _ in p], np.int32)
    sb = np.array([b for p in pairs for _, b in p], np.int32)
    out = coracle.product_tasks(m, k, n, coracle.packed(A.offsets, A.vals), coracle.packed(B.offsets, B.vals),
                                (res, ps, sa, sb), n_threads=3)
    C = O.Dia(m, n, list(out[0]), [out[2][s:s + O.diagonal_length(o, m, n)] + 1j * out[3][s:s + O.diagonal_length(o, m, n)]
                                   for o, s in zip(out[0], out[1])])
    if m * n <= 4_000_000:
        ref = O.to_dense(A) @ O.to_dense(B)
        np.testing.assert_allclose(O.to_dense(C), ref, rtol=1e-12, atol=1e-12)
    else:  # large: against the oracle's pair-by-pair product
        R = O.product(A, B)
        assert R.offsets == C.offsets
        for a, b in zip(R.vals, C.vals):
            np.testing.assert_allclose(b, a, rtol=1e-12, atol=1e-12)


def test_c_oracle_taylor_step_matches_oracle():
    """One fixed-degree Taylor step with the restated reference CPU SpMV against the oracle's evolution rule."""
    from oracle import coracle
    rows, cols, offs, vals = Hm.xxz(9, seed=4)
    M = O.Dia(rows, cols, offs, vals)
    psi = Hm.product_state(9, 3)
    got = coracle.taylor_step_tasks(rows, coracle.packed(offs, vals), psi, 0.3, 14, n_threads=2)
    ref = O.hamiltonian_simulation(M, psi, 0.3, 1, 14)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-13)


def test_c_oracle_block_size_rule():
    from oracle import coracle
    b = coracle.block_size()
    assert b % 8 == 0 and b > 0
    l1 = os.sysconf("SC_LEVEL1_DCACHE_SIZE") if "SC_LEVEL1_DCACHE_SIZE" in os.sysconf_names else 0
    if l1 == 49152:
        assert b == 968  # floor(floor(49152 * 0.95 / 48) / 8) * 8


def test_dense_expm_oracle_matches_dia_oracle():
    rows, cols, offs, vals = Hm.tfim(5, 1.0, 0.7)
    M = O.Dia(rows, cols, offs, vals)
    U = O.expm_negative(M, 0.3)
    Ud, roffs, m, s = O.expm_negative_dense(O.to_dense(M), offs, 0.3)
    assert roffs == U.offsets and s == O.scaling_exponent(O.estimate_1norm(M), 0.3, 1.0)
    np.testing.assert_allclose(Ud, O.to_dense(U), atol=1e-13)
    np.testing.assert_allclose(Ud, scipy.linalg.expm(-1j * 0.3 * O.to_dense(M)), atol=1e-11)


def test_spmv_benchmark_offsets_stored_count():
    """The SpMV benchmark's offsets (numpy default_rng(0), 64 distinct in [-N/2, N/2]) store 798,254,616 elements."""
    N = 1 << 24
    offs = Hm.random_offsets(N, 64, seed=0)
    assert len(set(offs)) == 64 and all(-N // 2 <= o <= N // 2 for o in offs)
    assert Hm.stored_elements(N, N, offs) == 798_254_616


def test_taylor_action_ms_oracle():
    rows, cols, offs, vals = Hm.tfim(5, 1.1, 0.9)
    H = O.to_dense(O.Dia(rows, cols, offs, vals))
    psi = Hm.random_state(rows, 1)
    exact = scipy.linalg.expm(-1j * 1.3 * H) @ psi
    np.testing.assert_allclose(O.taylor_action_ms(H, psi, 1.3, 30, 3), exact, atol=1e-13)
    # one step, degree m: the truncated series on each eigencomponent (closed form)
    w, V = np.linalg.eigh(H)
    c = V.conj().T @ psi
    from math import factorial
    poly = sum((-1j * 0.4 * w) ** k / factorial(k) for k in range(0, 4))
    np.testing.assert_allclose(O.taylor_action_ms(H, psi, 0.4, 3, 0), V @ (poly * c), atol=1e-14)
    poly2 = sum((-1j * 0.2 * w) ** k / factorial(k) for k in range(0, 3)) ** 2
    np.testing.assert_allclose(O.taylor_action_ms(H, psi, 0.4, 2, 1), V @ (poly2 * c), atol=1e-14)


def test_mlp_gradient_vs_finite_differences():
    rng = np.random.default_rng(12)
    dims = (5, 7, 6, 9)
    n = 7 * 5 + 7 + 6 * 7 + 6 + 9 * 6 + 9
    p = rng.standard_normal(n) * 0.5
    x = rng.standard_normal((11, 5))
    y = rng.integers(0, 9, 11)
    loss, g = O.mlp_loss_grad(p, dims, x, y)

    def L(q):
        z, _, _ = O.mlp_forward(q, dims, x)
        zm = z - z.max(1, keepdims=True)
        return float(np.mean(-(zm[np.arange(11), y] - np.log(np.exp(zm).sum(1)))))

    assert loss == pytest.approx(L(p), rel=1e-12)
    h = 1e-6
    fd = np.array([(L(p + h * e) - L(p - h * e)) / (2 * h) for e in np.eye(n)])
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-9)


def test_adam_step_closed_forms():
    """Adam from zero moments (b1 = 0.9, b2 = 0.999, eps = 1e-8, lr = 0.1), worked by hand.

    Step 1: m = (1-b1) g and v = (1-b2) g^2, so the bias-corrected moments are g and g^2 and the update is
    -lr g / (|g| + eps) = -0.0999999995 for g = 2. Step 2 with g = 0: m = b1 (1-b1) g1 and v = b2 (1-b2) g1^2, corrected
    by 1 - b1^2 = (1-b1)(1+b1) and 1 - b2^2, giving b1 g1 / (1+b1) and b2 g1^2 / (1+b2); the update is
    -lr (0.9/1.9) 2 / (sqrt(0.999/1.999) 2 + eps) = -0.0670058249397... A constant gradient keeps both corrected
    moments at g and g^2, so every step moves by lr g / (|g| + eps). Dropping the bias correction would give
    lr (0.1 g) / (sqrt(0.001) |g|) = 0.316 at step 1; correcting with step - 1 divides by zero."""
    p = np.array([1.0, -3.0, 0.5])
    g1 = np.array([2.0, -2.0, 0.0])
    m = np.zeros(3)
    v = np.zeros(3)
    O.adam_step(p, g1, m, v, 1, 0.1)
    np.testing.assert_allclose(p, [1.0 - 0.0999999995, -3.0 + 0.0999999995, 0.5], rtol=0, atol=1e-15)
    O.adam_step(p, np.zeros(3), m, v, 2, 0.1)
    np.testing.assert_allclose(p, [1.0 - 0.0999999995 - 0.06700582493973309,
                                   -3.0 + 0.0999999995 + 0.06700582493973309, 0.5], rtol=0, atol=1e-15)
    q = np.array([0.0])
    mq, vq = np.zeros(1), np.zeros(1)
    for step in range(1, 6):
        O.adam_step(q, np.array([-4.0]), mq, vq, step, 0.01)
        assert q[0] == pytest.approx(step * 0.01 * 4.0 / (4.0 + 1e-8), rel=1e-12)


def test_kron_apply_vs_numpy_kron():
    """The reference's own check (python_tests/test_api_vs_numpy.py:61-80): random square gates of size 2..11,
    before/after in 0..2, against np.kron(np.kron(I, G), I) @ x; y accumulates."""
    rng = np.random.default_rng(12)
    for _ in range(25):
        g = int(rng.integers(2, 12))
        b, a = (int(v) for v in rng.integers(0, 3, size=2))
        D = (rng.standard_normal((g, g)) + 1j * rng.standard_normal((g, g))) * (rng.random((g, g)) < 0.6)
        G = O.from_dense(D)
        big = np.kron(np.kron(np.eye(1 << b), D), np.eye(1 << a))
        x = rng.standard_normal(big.shape[1]) + 1j * rng.standard_normal(big.shape[1])
        y0 = rng.standard_normal(big.shape[0]) + 1j * rng.standard_normal(big.shape[0])
        np.testing.assert_allclose(O.kron_apply(b, G, a, x, y0), y0 + big @ x, rtol=0, atol=1e-12)
        np.testing.assert_allclose(O.identity_kron_matrix_times_vector(b, G, x[:g << b], np.zeros(g << b)),
                                   np.kron(np.eye(1 << b), D) @ x[:g << b], rtol=0, atol=1e-12)
        np.testing.assert_allclose(O.matrix_kron_identity_times_vector(G, a, x[:g << a], np.zeros(g << a)),
                                   np.kron(D, np.eye(1 << a)) @ x[:g << a], rtol=0, atol=1e-12)


def test_kron_apply_prime_gate_case():
    """cpp_tests/test_main.cpp:2231-2262: the 4 x 4 prime gate, (G (x) I_4) ones(16) -- each output is its row sum."""
    P = np.array([[2, 3, 5, 7], [11, 13, 17, 19], [23, 29, 31, 37], [41, 43, 47, 53]], complex)
    y = O.matrix_kron_identity_times_vector(O.from_dense(P), 2, np.ones(16, complex), np.zeros(16, complex))
    np.testing.assert_array_equal(y, np.repeat(P.sum(axis=1), 4))
    # a Hadamard on the first of 4 qubits applied to |0000> (test_main.cpp:2291-2330)
    H = np.array([[1, 1], [1, -1]], complex) / np.sqrt(2)
    x = np.zeros(16, complex)
    x[0] = 1
    y = O.matrix_kron_identity_times_vector(O.from_dense(H), 3, x, np.zeros(16, complex))
    ref = np.zeros(16, complex)
    ref[0] = ref[8] = 1 / np.sqrt(2)
    np.testing.assert_allclose(y, ref, atol=1e-15)


def test_c_kron_apply_tasks_matches_numpy_oracle():
    """The C restatement of the reference's Kronecker loop (the CPU baseline) against oracle.kron_apply."""
    from oracle import coracle as C
    rng = np.random.default_rng(21)
    for g, b, a in ((2, 3, 4), (4, 0, 5), (5, 2, 0), (8, 1, 3), (11, 2, 2)):
        D = (rng.standard_normal((g, g)) + 1j * rng.standard_normal((g, g))) * (rng.random((g, g)) < 0.6)
        M = O.from_dense(D)
        n = g << (b + a)
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        y0 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        offs, starts, re, im = C.packed(M.offsets, M.vals)
        yr, yi = y0.real.copy(), y0.imag.copy()
        C.kron_apply_tasks(b, (g, offs, starts, re, im), a, np.ascontiguousarray(x.real), np.ascontiguousarray(x.imag),
                           yr, yi)
        np.testing.assert_allclose(yr + 1j * yi, O.kron_apply(b, M, a, x, y0), rtol=0, atol=1e-12)


@pytest.mark.parametrize("sa,sb,axes", [((4, 6), (6, 5), ([1], [0])), ((2, 3, 4), (4, 3, 5), ([1, 2], [1, 0])),
                                        ((3, 4), (3, 4), ([0, 1], [0, 1])), ((2, 2, 2, 2), (2, 2), ([3], [1])),
                                        ((5, 3), (2, 7), ([], []))])
def test_tensordot_vs_numpy(sa, sb, axes):
    rng = np.random.default_rng(31)
    A = rng.standard_normal(sa) + 1j * rng.standard_normal(sa)
    B = rng.standard_normal(sb) + 1j * rng.standard_normal(sb)
    A[rng.random(sa) < 0.4] = 0
    B[rng.random(sb) < 0.4] = 0
    want = np.tensordot(A, B, axes=axes)
    got = O.tensordot(A, B, axes)
    dense = want.reshape(got.rows, got.cols)
    np.testing.assert_allclose(O.to_dense(got), dense, atol=1e-13)
    assert got.offsets == O.from_dense(dense).offsets
