"""GPU tests of planned gate sequences (dq_gate_plan): the fused passes give the same bits as applying the gates one
by one with the single-gate kernel (itself checked against the oracle in test_gpu_kron.py), and agree with the
oracle's Kronecker restatement on small states."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as O  # noqa: E402


@pytest.fixture(scope="module")
def d():
    import diaq_amd  # noqa: F401
    import diaq_gpu
    return diaq_gpu


def random_sequence(d, rng, n, n_gates, precision="fp64", diag_frac=0.3):
    seq, host = [], []
    for _ in range(n_gates):
        k = int(rng.integers(1, 5))
        q = int(rng.integers(0, n - k + 1))
        g = 1 << k
        if rng.random() < diag_frac:
            D = np.diag(np.exp(1j * rng.standard_normal(g)))
        else:
            D = (rng.standard_normal((g, g)) + 1j * rng.standard_normal((g, g))) * (rng.random((g, g)) < 0.5)
            D += np.eye(g)
            D /= np.linalg.norm(D, 2)
        M = O.from_dense(D)
        G = d.to_device(d.diaQMatrix.from_diagonals([g, g], M.offsets, M.vals), precision=precision)
        seq.append((G, q))
        host.append((M, k, q))
    return seq, host


def state(torch, n, seed, dtype=None):
    import torch as _t
    dtype = dtype or _t.float64
    g = torch.Generator(device="cuda").manual_seed(seed)
    return (torch.randn(1 << n, dtype=dtype, device="cuda", generator=g),
            torch.randn(1 << n, dtype=dtype, device="cuda", generator=g))


def sequential(d, seq, n, vr, vi):
    for G, q in seq:
        k = int(np.log2(G.shape[0]))
        d.kron_apply_device(G, n - q - k, q, vr.data_ptr(), vi.data_ptr(), vr.data_ptr(), vi.data_ptr(), False)


@pytest.mark.parametrize("kernel", [0, 1])
@pytest.mark.parametrize("tile", [10, 11, 12])
@pytest.mark.parametrize("n,n_gates,seed", [(12, 25, 1), (14, 40, 2), (16, 60, 3), (13, 1, 4), (15, 8, 5)])
def test_gate_plan_bitwise_equals_single_gates(d, kernel, tile, n, n_gates, seed):
    """Both pass kernels (LDS tiles; amplitudes in registers with shuffles) at each tile size."""
    import torch
    from diaq_amd import capi
    capi.check(capi.lib().dq_set_tuning(b"gates_tile_qubits", tile))
    capi.check(capi.lib().dq_set_tuning(b"gates_kernel", kernel))
    try:
        rng = np.random.default_rng(seed * 10 + tile)
        seq, _ = random_sequence(d, rng, n, n_gates)
        P = d.GatePlan(n, seq)
        assert 1 <= P.n_passes <= n_gates
        ar, ai = state(torch, n, seed)
        br, bi = ar.clone(), ai.clone()
        P.run(ar.data_ptr(), ai.data_ptr())
        sequential(d, seq, n, br, bi)
        torch.cuda.synchronize()
        assert torch.equal(ar, br) and torch.equal(ai, bi)
    finally:
        capi.check(capi.lib().dq_set_tuning(b"gates_tile_qubits", 11))
        capi.check(capi.lib().dq_set_tuning(b"gates_kernel", 0))


def test_gate_plan_vs_oracle_and_repetitions(d):
    import torch
    n = 12
    rng = np.random.default_rng(7)
    seq, host = random_sequence(d, rng, n, 20)
    P = d.GatePlan(n, seq)
    vr, vi = state(torch, n, 8)
    x = vr.cpu().numpy() + 1j * vi.cpu().numpy()
    P.run(vr.data_ptr(), vi.data_ptr(), repetitions=3)
    torch.cuda.synchronize()
    ref = x
    for _ in range(3):
        for M, k, q in host:
            ref = O.kron_apply(n - q - k, M, q, ref, np.zeros_like(ref))
    got = vr.cpu().numpy() + 1j * vi.cpu().numpy()
    assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-12


@pytest.mark.parametrize("kernel", [0, 1])
def test_gate_plan_fp32_bitwise(d, kernel):
    import torch
    from diaq_amd import capi
    capi.check(capi.lib().dq_set_tuning(b"gates_kernel", kernel))
    try:
        fp32_case(d, torch)
    finally:
        capi.check(capi.lib().dq_set_tuning(b"gates_kernel", 0))


def fp32_case(d, torch):
    n = 14
    rng = np.random.default_rng(9)
    seq, _ = random_sequence(d, rng, n, 30, precision="fp32")
    P = d.GatePlan(n, seq)
    ar, ai = state(torch, n, 10, torch.float32)
    br, bi = ar.clone(), ai.clone()
    P.run(ar.data_ptr(), ai.data_ptr())
    sequential(d, seq, n, br, bi)
    torch.cuda.synchronize()
    assert torch.equal(ar, br) and torch.equal(ai, bi)


@pytest.mark.parametrize("kernel", [0, 1])
def test_trotter_step_plan_full_size(d, kernel):
    """Ten first-order Trotter steps of TFIM-24 (27 + 28 gates per step in the 28-qubit bench; 23 + 24 here) as one
    plan run with repetitions: the same bits as the gates one by one, norm kept to 1e-12, and the ZZ layer fused
    into the first pass (passes << gates); both pass kernels."""
    import torch
    from diaq_amd import capi
    capi.check(capi.lib().dq_set_tuning(b"gates_kernel", kernel))
    try:
        trotter_case(d, torch)
    finally:
        capi.check(capi.lib().dq_set_tuning(b"gates_kernel", 0))


def trotter_case(d, torch):
    n = 24
    dt, J, h = 0.05, 1.0, 0.7
    zz = np.exp(-1j * dt * J * np.array([1.0, -1.0, -1.0, 1.0]))
    c, s_ = np.cos(dt * h), np.sin(dt * h)
    GZ = d.to_device(d.diaQMatrix.from_diagonals([4, 4], [0], [zz]))
    GX = d.to_device(d.diaQMatrix.from_diagonals([2, 2], [-1, 0, 1], [np.array([-1j * s_]), np.array([c, c]) + 0j,
                                                                       np.array([-1j * s_])]))
    seq = [(GZ, q) for q in range(n - 1)] + [(GX, q) for q in range(n)]
    P = d.GatePlan(n, seq)
    assert P.n_passes <= 5
    ar, ai = state(torch, n, 11)
    nrm = torch.sqrt((ar * ar + ai * ai).sum())
    ar /= nrm
    ai /= nrm
    br, bi = ar.clone(), ai.clone()
    P.run(ar.data_ptr(), ai.data_ptr(), repetitions=10)
    for _ in range(10):
        sequential(d, seq, n, br, bi)
    torch.cuda.synchronize()
    assert torch.equal(ar, br) and torch.equal(ai, bi)
    assert abs(float(torch.sqrt((ar * ar + ai * ai).sum())) - 1.0) < 1e-12


def test_gate_plan_errors(d):
    G2 = d.to_device(d.from_numpy(np.eye(2, dtype=complex)))
    G3 = d.to_device(d.from_numpy(np.eye(3, dtype=complex)))
    G32 = d.to_device(d.from_numpy(np.eye(2, dtype=complex)), precision="fp32")
    with pytest.raises(RuntimeError, match="n_qubits"):
        d.GatePlan(8, [(G2, 0)])
    with pytest.raises(RuntimeError, match="2\\^k"):
        d.GatePlan(12, [(G3, 0)])
    with pytest.raises(RuntimeError, match="precision"):
        d.GatePlan(12, [(G2, 0), (G32, 1)])
    with pytest.raises(RuntimeError, match="beyond"):
        d.GatePlan(12, [(G2, 12)])


@pytest.mark.parametrize("low", [4, 6])
def test_gate_plan_low_qubits_bitwise(d, low):
    """Other low-qubit counts per tile (the default is 5): the same bits as gate by gate."""
    import torch
    from diaq_amd import capi
    capi.check(capi.lib().dq_set_tuning(b"gates_low_qubits", low))
    try:
        rng = np.random.default_rng(30 + low)
        n = 15
        seq, _ = random_sequence(d, rng, n, 40)
        P = d.GatePlan(n, seq)
        ar, ai = state(torch, n, 31)
        br, bi = ar.clone(), ai.clone()
        P.run(ar.data_ptr(), ai.data_ptr())
        sequential(d, seq, n, br, bi)
        torch.cuda.synchronize()
        assert torch.equal(ar, br) and torch.equal(ai, bi)
    finally:
        capi.check(capi.lib().dq_set_tuning(b"gates_low_qubits", 5))
