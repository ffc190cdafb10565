"""GPU parity tests of the gate-application kernels, y (+)= (I_{2^before} (x) G (x) I_{2^after}) x: the reference's
host-vector API (src/spKron.cpp:286-571, bound at include/pythonWrappers.cpp:409-455) and the device entry point,
against the oracle (oracle.kron_apply, itself pinned against np.kron). FP64 within 1e-12 relative; FP32 within 2e-6."""
import ctypes

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


def random_gate(rng, g, density=0.6, offsets=None):
    D = (rng.standard_normal((g, g)) + 1j * rng.standard_normal((g, g))) * (rng.random((g, g)) < density)
    D[0, 0] = 1.0 + 0.5j  # at least the main diagonal
    M = O.from_dense(D)
    if offsets == "shuffled":  # slots out of ascending order
        order = rng.permutation(len(M.offsets))
        M = O.Dia(g, g, [M.offsets[k] for k in order], [M.vals[k] for k in order])
    return M


def dev_gate(d, M, precision="fp64"):
    return d.to_device(d.diaQMatrix.from_diagonals([M.rows, M.cols], M.offsets, M.vals), precision=precision)


def split(torch, v, dtype):
    return (torch.tensor(v.real, dtype=dtype, device="cuda").contiguous(),
            torch.tensor(v.imag, dtype=dtype, device="cuda").contiguous())


def test_reference_host_api_vs_numpy_kron(d):
    """python_tests/test_api_vs_numpy.py:61-80 through the drop-in names: 25 random square gates (2..11), before and
    after in 0..2, y starting at zero."""
    rng = np.random.default_rng(4)
    for _ in range(25):
        g = int(rng.integers(2, 12))
        s1, s2 = (int(v) for v in rng.integers(0, 3, size=2))
        Sq = (rng.standard_normal((g, g)) + 1j * rng.standard_normal((g, g))) * (rng.random((g, g)) < 0.5)
        big = np.kron(np.kron(np.eye(1 << s1), Sq), np.eye(1 << s2))
        xb = rng.standard_normal(big.shape[1]) + 1j * rng.standard_normal(big.shape[1])
        y = d.from_numpy_vector(np.zeros(big.shape[0], complex))
        got = d.to_numpy_vector(d.identity_kron_matrix_kron_identity_times_vector(s1, d.from_numpy(Sq), s2,
                                                                                 d.from_numpy_vector(xb), y))
        assert rel(got, big @ xb) < TOL
        xi = rng.standard_normal(g << s1) + 0j
        y = d.from_numpy_vector(np.zeros(g << s1, complex))
        got = d.to_numpy_vector(d.identity_kron_matrix_times_vector(s1, d.from_numpy(Sq), d.from_numpy_vector(xi), y))
        assert rel(got, np.kron(np.eye(1 << s1), Sq) @ xi) < TOL
        xm = rng.standard_normal(g << s2) + 0j
        y = d.from_numpy_vector(np.zeros(g << s2, complex))
        got = d.to_numpy_vector(d.matrix_kron_identity_times_vector(d.from_numpy(Sq), s2, d.from_numpy_vector(xm), y))
        assert rel(got, np.kron(Sq, np.eye(1 << s2)) @ xm) < TOL


def test_host_api_accumulates_and_checks_lengths(d):
    rng = np.random.default_rng(5)
    M = random_gate(rng, 4)
    x = rng.standard_normal(32) + 1j * rng.standard_normal(32)
    y0 = rng.standard_normal(32) + 1j * rng.standard_normal(32)
    G = d.diaQMatrix.from_diagonals([4, 4], M.offsets, M.vals)
    got = d.to_numpy_vector(d.identity_kron_matrix_kron_identity_times_vector(1, G, 2, d.from_numpy_vector(x),
                                                                             d.from_numpy_vector(y0)))
    assert rel(got, O.kron_apply(1, M, 2, x, y0)) < TOL
    with pytest.raises(RuntimeError, match="dimension mismatch"):
        d.identity_kron_matrix_kron_identity_times_vector(1, G, 1, d.from_numpy_vector(x), d.from_numpy_vector(y0))


@pytest.mark.parametrize("g", [2, 3, 4, 5, 8, 11, 16, 17, 24])
@pytest.mark.parametrize("before,after", [(0, 0), (2, 0), (0, 3), (1, 7), (3, 6)])
@pytest.mark.parametrize("slots", ["ascending", "shuffled"])
def test_device_kron_apply_vs_oracle(d, g, before, after, slots):
    """Register path (g <= 16) and element path (g > 16), coalesced (2^after >= 64) and strided columns, slots in
    ascending and in shuffled order, accumulate on and off; FP64 against the oracle."""
    import torch
    from diaq_amd import capi
    rng = np.random.default_rng(g * 100 + before * 10 + after)
    M = random_gate(rng, g, offsets=slots)
    n = g << (before + after)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y0 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = O.kron_apply(before, M, after, x, y0)
    G = dev_gate(d, M)
    xr, xi = split(torch, x, torch.float64)
    for acc in (1, 0):
        yr, yi = split(torch, y0, torch.float64)
        capi.check(capi.lib().dq_kron_apply_device(G.handle, before, after, xr.data_ptr(), xi.data_ptr(),
                                                   yr.data_ptr(), yi.data_ptr(), acc, 0))
        torch.cuda.synchronize()
        got = yr.cpu().numpy() + 1j * yi.cpu().numpy()
        want = ref if acc else ref - y0
        assert rel(got, want) < TOL, acc


@pytest.mark.parametrize("g", [2, 4, 8, 16])
def test_device_kron_apply_in_place_and_fp32(d, g):
    """x == y (in place) on the register path, through the binding on a side stream; the FP32 form within 2e-6."""
    import torch
    rng = np.random.default_rng(g)
    M = random_gate(rng, g)
    before, after = 3, 8
    n = g << (before + after)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = O.kron_apply(before, M, after, x, np.zeros(n, complex))
    G = dev_gate(d, M)
    vr, vi = split(torch, x, torch.float64)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        d.kron_apply_device(G, before, after, vr.data_ptr(), vi.data_ptr(), vr.data_ptr(), vi.data_ptr(), False,
                            s.cuda_stream)
    s.synchronize()
    assert rel(vr.cpu().numpy() + 1j * vi.cpu().numpy(), ref) < TOL
    G32 = dev_gate(d, M, "fp32")
    xr, xi = split(torch, x, torch.float32)
    yr, yi = torch.zeros_like(xr), torch.zeros_like(xi)
    d.kron_apply_device(G32, before, after, xr.data_ptr(), xi.data_ptr(), yr.data_ptr(), yi.data_ptr(), True, 0)
    torch.cuda.synchronize()
    assert rel(yr.cpu().double().numpy() + 1j * yi.cpu().double().numpy(), ref) < 2e-6


def test_device_kron_apply_errors(d):
    import torch
    rng = np.random.default_rng(9)
    G = dev_gate(d, random_gate(rng, 20))
    v = torch.zeros(20 << 2, dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError, match="in place"):
        d.kron_apply_device(G, 1, 1, v.data_ptr(), v.data_ptr(), v.data_ptr(), v.data_ptr())
    R = d.to_device(d.from_numpy(np.ones((3, 4), complex)))
    with pytest.raises(RuntimeError, match="square"):
        d.kron_apply_device(R, 0, 0, v.data_ptr(), v.data_ptr(), v.data_ptr(), v.data_ptr())


def test_trotter_layer_full_size(d):
    """A 26-qubit state (2^26 amplitudes): a layer of two-qubit ZZ rotations and one-qubit X rotations applied in
    place on the device keeps the norm (unitary gates, within 1e-12), and undoing the layer with the adjoint gates
    in reverse order gives back the state; one gate checked against a reshape-and-multiply of the whole vector."""
    import torch
    n = 26
    N = 1 << n
    g = torch.Generator(device="cuda").manual_seed(1)
    vr = torch.randn(N, dtype=torch.float64, device="cuda", generator=g)
    vi = torch.randn(N, dtype=torch.float64, device="cuda", generator=g)
    nrm = torch.sqrt((vr * vr + vi * vi).sum())
    vr /= nrm
    vi /= nrm
    r0, i0 = vr.clone(), vi.clone()
    t = 0.173
    zz = np.diag(np.exp(-1j * t * np.array([1, -1, -1, 1])))
    rx = np.array([[np.cos(t), -1j * np.sin(t)], [-1j * np.sin(t), np.cos(t)]])
    gates = [(dev_gate(d, O.from_dense(zz)), dev_gate(d, O.from_dense(zz.conj().T)), n - 2 - q, q)
             for q in range(0, n - 1, 5)]
    gates += [(dev_gate(d, O.from_dense(rx)), dev_gate(d, O.from_dense(rx.conj().T)), n - 1 - q, q)
              for q in range(0, n, 3)]
    for G, _, before, after in gates:
        d.kron_apply_device(G, before, after, vr.data_ptr(), vi.data_ptr(), vr.data_ptr(), vi.data_ptr(), False)
    torch.cuda.synchronize()
    assert abs(float(torch.sqrt((vr * vr + vi * vi).sum())) - 1.0) < 1e-12
    for _, Gh, before, after in reversed(gates):
        d.kron_apply_device(Gh, before, after, vr.data_ptr(), vi.data_ptr(), vr.data_ptr(), vi.data_ptr(), False)
    torch.cuda.synchronize()
    assert float(torch.max(torch.abs(vr - r0)).item()) < 1e-12 and float(torch.max(torch.abs(vi - i0)).item()) < 1e-12
    # one gate against the reshaped product on the CPU
    x = (r0.cpu().numpy() + 1j * i0.cpu().numpy())
    yr, yi = torch.empty_like(r0), torch.empty_like(i0)
    before, after = 7, n - 9
    d.kron_apply_device(gates[0][0], before, after, r0.data_ptr(), i0.data_ptr(), yr.data_ptr(), yi.data_ptr(), False)
    torch.cuda.synchronize()
    ref = np.einsum("rc,bca->bra", zz, x.reshape(1 << before, 4, 1 << after)).reshape(-1)
    assert rel(yr.cpu().numpy() + 1j * yi.cpu().numpy(), ref) < TOL
