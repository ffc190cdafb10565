"""The learned (m, s) selector: builder, labels and the network against the oracle (CPU parts), and the HIP training
kernels against the numpy twin (GPU parts)."""
import numpy as np
import pytest

from oracle import oracle as O
from diaq_amd import hamiltonians as Hm
from diaq_amd import selector as S


def test_pauli_sum_builder_vs_kron():
    rng = np.random.default_rng(3)
    n = 6
    N, _, offs, vals, strings = Hm.pauli_sum(n, 15, rng)
    X = np.array([[0, 1], [1, 0]], complex)
    Y = np.array([[0, -1j], [1j, 0]])
    Z = np.diag([1.0, -1.0]).astype(complex)
    H = np.zeros((N, N), complex)
    for c, ops in strings:
        M = np.ones((1, 1))
        for q in reversed(range(n)):  # qubit q is bit q: the highest qubit is the leftmost factor
            M = np.kron(M, {"X": X, "Y": Y, "Z": Z}.get(ops.get(q), np.eye(2)))
        H += c * M
    np.testing.assert_allclose(O.to_dense(O.Dia(N, N, offs, vals)), H, atol=1e-14)
    assert offs == O.from_dense(H).offsets


def test_taylor_error_is_the_scheduled_action_error():
    rng = np.random.default_rng(4)
    lam = rng.uniform(-3, 3, 16)
    H = np.diag(lam).astype(complex)
    psi = np.eye(16)[:, 5].astype(complex)
    for m, s, t in [(4, 0, 0.5), (6, 2, 2.0), (9, 3, 1.7)]:
        got = np.abs(O.taylor_action_ms(H, psi, t, m, s)[5] - np.exp(-1j * lam[5] * t))
        assert S.taylor_error(lam[5] * t, m, s) == pytest.approx(got, rel=1e-9, abs=1e-15)


@pytest.mark.parametrize("rho_t,tol", [(0.05, 1e-10), (2.0, 1e-8), (20.0, 1e-6), (150.0, 1e-5)])
def test_label_is_cheapest_feasible_by_brute_force(rho_t, tol):
    xs = np.linspace(0, rho_t, 48)
    feas = [(m * 2 ** s, s, m) for m in S.M_VALUES for s in S.S_VALUES if np.max(S.taylor_error(xs, m, s)) <= tol]
    cost, s, m = min(feas)
    assert S.label(rho_t, tol) == (m, s)


def test_class_index_round_trip():
    assert S.N_CLASSES == 195
    for c in range(S.N_CLASSES):
        assert S.cls_of(*S.ms_of(c)) == c


@pytest.mark.gpu
def test_mlp_kernels_vs_numpy_twin():
    import torch
    import diaq_amd  # noqa: F401
    rng = np.random.default_rng(5)
    sel = S.Selector(seed=7)
    p0 = sel.get_params().astype(np.float64)
    B = 96
    x = rng.standard_normal((B, 5)).astype(np.float32)
    y = rng.integers(0, S.N_CLASSES, B).astype(np.int32)
    xd = torch.as_tensor(x, device="cuda")
    yd = torch.as_tensor(y, device="cuda")
    loss, g = sel.loss_grad(xd, yd)
    l_ref, g_ref = O.mlp_loss_grad(p0, S.DIMS, x.astype(np.float64), y)
    assert loss == pytest.approx(l_ref, rel=1e-5)
    assert np.max(np.abs(g - g_ref)) <= 1e-5 * max(1.0, np.max(np.abs(g_ref)))
    # Adam steps against the numpy twin
    p = p0.copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for step in range(1, 4):
        _, gr = O.mlp_loss_grad(p, S.DIMS, x.astype(np.float64), y)
        O.adam_step(p, gr, m, v, step, 1e-2)
        sel.train_step(xd, yd, 1e-2)
    np.testing.assert_allclose(sel.get_params(), p, atol=2e-5)
    # inference: argmax of the logits
    z, _, _ = O.mlp_forward(sel.get_params().astype(np.float64), S.DIMS, x.astype(np.float64))
    cls = sel.predict_classes(x)  # features are fed raw (mean 0, std 1)
    agree = np.mean(cls == z.argmax(1))
    assert agree >= 0.99  # ties within float32 rounding may flip a class


@pytest.mark.gpu
def test_trained_selector_loads_and_predicts_feasible_schedules():
    import os
    import diaq_amd  # noqa: F401
    if not os.path.exists(S.WEIGHTS):
        pytest.skip("no trained weights in the tree")
    sel = S.Selector.load()
    feats = [S.features(50.0, 0.1, 21, 2 ** 11, 2 ** 12, 1e-6), S.features(10.0, 1.0, 41, 2 ** 15, 2 ** 16, 1e-8)]
    for m, s in sel.predict(np.array(feats)):
        assert m in S.M_VALUES and s in S.S_VALUES


def test_fast_labels_agree_with_reference_labels():
    rng = np.random.default_rng(6)
    rho_t = np.exp(rng.uniform(np.log(1e-3), np.log(300), 40))
    tol = 10.0 ** rng.uniform(-10, -4, 40)
    fast = S.labels_fast(rho_t, tol)
    for x, e, c in zip(rho_t, tol, fast):
        ref = S.label(x, e, n_grid=400)
        if ref is None:
            assert c == -1
        else:
            m, s = S.ms_of(int(c))
            # the table is evaluated at the next grid point >= rho_t: the same or a slightly more conservative choice
            assert m * 2 ** s >= ref[0] * 2 ** ref[1]
            assert m * 2 ** s <= 2 * ref[0] * 2 ** ref[1]
