"""The benchmark configurations at their real sizes, asserted rather than only reported.

- XXZ-20 (N = 2^20, Jxy = 1, Jz = 0.8, fields h_i ~ U[-1, 1] seed 0), product state seed 0, dt = 0.05: two fused
  FP64 steps against the C oracle (the reference's CPU SpMV inside the same fixed-degree Taylor step); the 100-step
  evolution at the reference's m = 100 keeps the norm, and the plain-rule degree agrees with it; the precision form
  (FP32 terms, FP64 state) within 1e-6 of that FP64 result.
- The N = 2^24, 64-diagonal SpMV in FP32 (the FP64 form is in test_gpu_parity.py).
- The batched form: 256 TFIM Hamiltonians (n ~ U{10..18}) x 8 states with the learned (m, s), the probe and the
  plain-rule fallback, run as one captured graph: every state within 1e-6 of an FP64 reference, graph replays equal
  to the eager run bit for bit, and fewer SpMVs than the plain rule.
- The learned selector on held-out Pauli sums (n <= 12): every final state within tol of scipy's expm_multiply, the
  fallback count, and the cost against the plain rule.
- FP32 expm_negative (FP32 Taylor terms, FP64 squarings) against the FP64 dense oracle.
"""
import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as O  # noqa: E402
from diaq_amd import hamiltonians as Hm  # noqa: E402

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


def norm_rule_m(norm1_dt, tol, m_max=100):
    """Smallest m with (||H dt||_1)^(m+1) / (m+1)! <= tol."""
    for m in range(1, m_max + 1):
        if (m + 1) * math.log(norm1_dt) - math.lgamma(m + 2) <= math.log(tol):
            return m
    return m_max


# ---------------------------------------------------------------------------------------------------------------------
# XXZ-20 fused evolution
@pytest.fixture(scope="module")
def xxz20(d):
    import torch
    n = 20
    rows, cols, offs, vals = Hm.xxz(n, 1.0, 0.8, seed=0)
    H = d.to_device(d.diaQMatrix.from_diagonals([rows, cols], offs, vals))
    psi0 = Hm.product_state(n, seed=0)
    # the reference's default degree: truncation far below FP64 rounding, the FP64 oracle of the precision form
    pr = torch.as_tensor(psi0.real, device="cuda").contiguous()
    pi = torch.as_tensor(psi0.imag, device="cuda").contiguous()
    d.hamsim_device(H, pr.data_ptr(), pi.data_ptr(), 5.0, 100, 100)
    torch.cuda.synchronize()
    ref = (pr + 1j * pi).cpu().numpy()
    return dict(rows=rows, offs=offs, vals=vals, H=H, psi0=psi0, ref=ref, norm1=d.estimate_1norm(H))


def test_xxz20_two_fp64_steps_vs_c_oracle(d, xxz20):
    import torch
    from oracle import coracle
    dt = 0.05
    m = norm_rule_m(xxz20["norm1"] * dt, 1e-8)
    assert m == 18
    pr = torch.as_tensor(xxz20["psi0"].real, device="cuda").contiguous()
    pi = torch.as_tensor(xxz20["psi0"].imag, device="cuda").contiguous()
    d.hamsim_device(xxz20["H"], pr.data_ptr(), pi.data_ptr(), 2 * dt, 2, m)
    torch.cuda.synchronize()
    got = (pr + 1j * pi).cpu().numpy()
    Hp = coracle.packed(xxz20["offs"], xxz20["vals"])
    ref = coracle.taylor_step_tasks(xxz20["rows"], Hp, xxz20["psi0"], dt, m)
    ref = coracle.taylor_step_tasks(xxz20["rows"], Hp, ref, dt, m)
    assert rel(got, ref) < TOL


def test_xxz20_100_steps_norm_and_degree(d, xxz20):
    import torch
    ref = xxz20["ref"]
    assert abs(np.linalg.norm(ref) - 1.0) < 1e-12
    m = norm_rule_m(xxz20["norm1"] * 0.05, 1e-12)
    pr = torch.as_tensor(xxz20["psi0"].real, device="cuda").contiguous()
    pi = torch.as_tensor(xxz20["psi0"].imag, device="cuda").contiguous()
    d.hamsim_device(xxz20["H"], pr.data_ptr(), pi.data_ptr(), 5.0, 100, m)
    torch.cuda.synchronize()
    got = (pr + 1j * pi).cpu().numpy()
    assert np.linalg.norm(got - ref) < 1e-10
    assert abs(np.linalg.norm(got) - 1.0) < 1e-10


def test_xxz20_precision_form_vs_fp64(d, xxz20):
    """FP32 values and Taylor terms, each step summed into an FP64 state, renormalised in FP64 at the end: within
    1e-6 (L2, unit state) of the FP64 evolution at m = 100, at the plain-rule degree for a 1e-8 per-step budget."""
    import torch
    H32 = xxz20["H"].astype("fp32")
    m = norm_rule_m(xxz20["norm1"] * 0.05, 1e-8)
    pr = torch.as_tensor(xxz20["psi0"].real, device="cuda").contiguous()
    pi = torch.as_tensor(xxz20["psi0"].imag, device="cuda").contiguous()
    d.hamsim_device_fp64_state(H32, pr.data_ptr(), pi.data_ptr(), 5.0, 100, m)
    torch.cuda.synchronize()
    got = (pr + 1j * pi).cpu().numpy()
    err = float(np.linalg.norm(got - xxz20["ref"]))
    print(f"XXZ-20 precision form: L2 error {err:.2e} against the FP64 evolution (m = 100)")
    assert err < 1e-6
    assert abs(np.linalg.norm(got) - np.linalg.norm(xxz20["psi0"])) < 1e-12
    # the FP32-state entry: the same evolution, the state rounded to FP32 once at the end
    qr = torch.as_tensor(xxz20["psi0"].real, device="cuda").float().contiguous()
    qi = torch.as_tensor(xxz20["psi0"].imag, device="cuda").float().contiguous()
    d.hamsim_device(H32, qr.data_ptr(), qi.data_ptr(), 5.0, 100, m)
    torch.cuda.synchronize()
    got32 = (qr.double() + 1j * qi.double()).cpu().numpy()
    assert float(np.linalg.norm(got32 - got)) < 1e-6
    # the host-vector entry point with an FP32 H runs the same precision form: the same bits
    from diaq_amd import capi
    N = xxz20["rows"]
    hr, hi = np.ascontiguousarray(xxz20["psi0"].real), np.ascontiguousarray(xxz20["psi0"].imag)
    outr, outi = np.empty(N), np.empty(N)
    capi.check(capi.lib().dq_hamsim_host(H32.handle, capi.ptr(hr), capi.ptr(hi), 5.0, 100, m, capi.ptr(outr),
                                         capi.ptr(outi)))
    np.testing.assert_array_equal(outr + 1j * outi, got)
    # and the reference's host-matrix drop-in (FP64) gives the m = 100 FP64 result
    host = d.to_numpy_vector(d.hamiltonian_simulation_cuda(
        d.diaQMatrix.from_diagonals([N, N], xxz20["offs"], xxz20["vals"]), d.from_numpy_vector(xxz20["psi0"]),
        5.0, 100, 100))
    np.testing.assert_array_equal(host, xxz20["ref"])


# ---------------------------------------------------------------------------------------------------------------------
# N = 2^24 SpMV in FP32
def test_spmv_fp32_full_size_property_2p24(d):
    """The SpMV benchmark matrix (N = 2^24, 64 random offsets) in FP32: closed-form values, sampled rows against the
    row rule evaluated in FP64 on the same FP32-rounded values and x, and linearity on the whole vector."""
    import torch
    from diaq_amd import capi
    N = 1 << 24
    offs = Hm.random_offsets(N, 64, seed=0)
    lens = [O.diagonal_length(o, N, N) for o in offs]
    out = ctypes.c_void_p()
    o32 = np.asarray(offs, np.int32)
    capi.check(capi.lib().dq_devmat_zeros(N, N, len(offs), capi.ptr(o32, ctypes.c_int32), capi.DQ_F32,
                                          ctypes.byref(out)))
    try:
        pos = torch.cat([torch.arange(L, device="cuda", dtype=torch.float64) for L in lens])
        kid = torch.cat([torch.full((L,), k, device="cuda", dtype=torch.float64) for k, L in enumerate(lens)])
        re = torch.cos(1e-3 * pos + 0.7 * kid).float()
        im = torch.sin(3e-4 * pos - 0.3 * kid).float()
        del pos, kid
        capi.check(capi.lib().dq_devmat_set_packed_device(out, re.data_ptr(), im.data_ptr(), 0))
        torch.cuda.synchronize()
        del re, im
        idx = torch.arange(N, device="cuda", dtype=torch.float64)
        x1r, x1i = torch.cos(2e-5 * idx).float(), torch.sin(5e-6 * idx).float()
        x2r, x2i = torch.sin(7e-6 * idx + 1.0).float(), torch.cos(1e-6 * idx).float()
        ys = []
        for xr, xi in ((x1r, x1i), (x2r, x2i), (x1r + 2 * x2r, x1i + 2 * x2i)):
            yr, yi = torch.empty_like(xr), torch.empty_like(xi)
            capi.check(capi.lib().dq_spmv_device(out, xr.data_ptr(), xi.data_ptr(), yr.data_ptr(), yi.data_ptr(), 0))
            torch.cuda.synchronize()
            ys.append((yr.double(), yi.double()))
        scale = float(torch.max(torch.abs(ys[2][0])))
        # x1 + 2 x2 is itself rounded to FP32, so linearity holds to FP32 rounding of the inputs and the row sums
        assert float(torch.max(torch.abs(ys[2][0] - (ys[0][0] + 2 * ys[1][0])))) / scale < 1e-5
        assert float(torch.max(torch.abs(ys[2][1] - (ys[0][1] + 2 * ys[1][1])))) / scale < 1e-5
        rows = np.unique(np.concatenate([np.random.default_rng(1).integers(0, N, 4000), [0, N - 1, N // 2]]))
        ar = np.arange(N, dtype=np.float64)
        x1 = (np.cos(2e-5 * ar).astype(np.float32).astype(np.float64)
              + 1j * np.sin(5e-6 * ar).astype(np.float32).astype(np.float64))
        expect = np.zeros(len(rows), complex)
        for k, o in enumerate(offs):
            p = rows if o >= 0 else rows + o
            c = rows + o
            ok = (p >= 0) & (p < lens[k]) & (c >= 0) & (c < N)
            a = (np.cos(1e-3 * p[ok] + 0.7 * k).astype(np.float32).astype(np.float64)
                 + 1j * np.sin(3e-4 * p[ok] - 0.3 * k).astype(np.float32).astype(np.float64))
            expect[ok] += a * x1[c[ok]]
        got = (ys[0][0] + 1j * ys[0][1]).cpu().numpy()[rows]
        err = rel(got, expect)
        print(f"FP32 SpMV N=2^24: max rel error {err:.2e} on {len(rows)} sampled rows")
        assert err < 1e-6
    finally:
        capi.lib().dq_devmat_destroy(out)


# ---------------------------------------------------------------------------------------------------------------------
# Batched form: 256 TFIM x 8 states, learned (m, s), one graph
def test_batched_256_tfim_learned_schedule(d):
    import scipy.sparse
    import scipy.sparse.linalg
    import torch
    from diaq_amd import evolution as ev
    from diaq_amd import selector as S
    jobs = Hm.batched_tfim_workload(256, 8, seed=0)
    assert {j["n"] for j in jobs} == set(range(10, 19))
    Ds, mats = [], []
    for j in jobs:
        r, c, offs, vals = Hm.tfim(j["n"], j["J"], j["h"])
        mats.append((r, offs, vals))
        Ds.append(d.to_device(d.diaQMatrix.from_diagonals([r, c], offs, vals)))
    tol = 1e-6
    norms = [d.estimate_1norm(D) for D in Ds]
    plain = [ev.plain_rule(nm, j["t"], tol) for nm, j in zip(norms, jobs)]
    sel = S.Selector.load()
    B = ev.BatchedEvolution(Ds, 8, "fp32")

    def load():
        for q, j in enumerate(jobs):
            B.set_states(q, j["states"])

    feats = np.array([S.features(nm, j["t"], len(D.diagonals), max(abs(o) for o in D.diagonals), D.shape[0], tol / 4)
                      for nm, j, D in zip(norms, jobs, Ds)])
    prop = sel.predict(feats)
    load()
    B.set_schedule([j["t"] for j in jobs], [m for m, s in plain], [s for m, s in plain])
    B.tune()
    load()
    B.set_schedule([j["t"] for j in jobs], [m for m, s in prop], [s for m, s in prop])
    est = B.probe()
    final = [p if e <= tol / 4 else q for p, q, e in zip(prop, plain, est)]
    fallbacks = sum(1 for p, f in zip(prop, final) if p != f)
    B.set_schedule([j["t"] for j in jobs], [m for m, s in final], [s for m, s in final])
    load()
    B.run(use_graph=True, renormalize=True)
    first = [B.get_states(q, 1 << j["n"]) for q, j in enumerate(jobs)]
    for graph in (True, False):  # a replay of the captured graph, then the same launches without a graph
        load()
        B.run(use_graph=graph, renormalize=True)
        for q, j in enumerate(jobs):
            np.testing.assert_array_equal(B.get_states(q, 1 << j["n"]), first[q], err_msg=f"job {q} graph={graph}")
    # FP64 references: scipy's expm_multiply for n <= 12, the FP64 fused path (pinned against the oracle) at a 1e-12
    # budget above
    worst = 0.0
    sid = torch.cuda.current_stream().cuda_stream
    for q, (j, D, (r, offs, vals)) in enumerate(zip(jobs, Ds, mats)):
        if j["n"] <= 12:
            rows_, cols_, data = [], [], []
            for dd, v in zip(offs, vals):
                p = np.arange(len(v))
                rows_.append(p if dd >= 0 else p - dd)
                cols_.append(p + dd if dd >= 0 else p)
                data.append(v)
            Hs = scipy.sparse.csr_matrix((np.concatenate(data), (np.concatenate(rows_), np.concatenate(cols_))),
                                         shape=(r, r))
            ref = scipy.sparse.linalg.expm_multiply(-1j * j["t"] * Hs, j["states"].T).T
        else:
            m, s = ev.plain_rule(norms[q], j["t"], 1e-12)
            ref = []
            for st in j["states"]:
                pr = torch.as_tensor(st.real, device="cuda").contiguous()
                pi = torch.as_tensor(st.imag, device="cuda").contiguous()
                d.hamsim_device(D, pr.data_ptr(), pi.data_ptr(), j["t"], 1 << s, m, sid)
                ref.append((pr + 1j * pi).cpu().numpy())
            ref = np.array(ref)
        worst = max(worst, float(np.max(np.linalg.norm(first[q] - ref, axis=1))))
    spmv_learned = sum(m * 2 ** s for m, s in final)
    spmv_plain = sum(m * 2 ** s for m, s in plain)
    print(f"batched 256 TFIM: max state error {worst:.2e}, fallbacks {fallbacks}, SpMVs per state "
          f"{spmv_learned} learned vs {spmv_plain} plain")
    assert worst <= 1e-6
    assert fallbacks <= 8
    assert spmv_learned < spmv_plain / 2


# ---------------------------------------------------------------------------------------------------------------------
# Learned selector: what the network gains over the plain rule on held-out Pauli sums
def heldout_pauli_cases(n_cases=200, n_ham=200, per_ham=10, seed=9000, n_max=12):
    """The held-out draw of the selector's training script (seed 9000: n ~ U{8..16}, 10-60 strings, per Hamiltonian
    per_ham (t, tol) pairs with t ~ log-uniform[1e-3, 10], tol ~ log-uniform[1e-10, 1e-4]); the Hamiltonians with
    n <= n_max, in draw order, until they hold n_cases cases."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_ham):
        if len(out) * per_ham >= n_cases:
            break
        n = int(rng.integers(8, 16, endpoint=True))
        P = int(rng.integers(10, 60, endpoint=True))
        N, _, offs, vals, _ = Hm.pauli_sum(n, P, rng)
        ts = np.exp(rng.uniform(np.log(1e-3), np.log(10.0), per_ham))
        tols = 10.0 ** rng.uniform(-10, -4, per_ham)
        if n <= n_max:
            out.append((n, offs, vals, ts, tols))
    return out


def test_learned_selector_gain_on_heldout_pauli_sums(d):
    import scipy.sparse
    import scipy.sparse.linalg
    import torch
    from diaq_amd import evolution as ev
    from diaq_amd import selector as S
    sel = S.Selector.load()
    cases = heldout_pauli_cases()
    sid = torch.cuda.current_stream().cuda_stream
    cost_learned, cost_plain, fallbacks, ratios, count = 0, 0, 0, [], 0
    for h, (n, offs, vals, ts, tols) in enumerate(cases):
        if count >= 200:
            break
        N = 1 << n
        D = d.to_device(d.diaQMatrix.from_diagonals([N, N], offs, vals))
        norm1 = d.estimate_1norm(D)
        bw = max(abs(o) for o in offs)
        rows_, cols_, data = [], [], []
        for dd, v in zip(offs, vals):
            p = np.arange(len(v))
            rows_.append(p if dd >= 0 else p - dd)
            cols_.append(p + dd if dd >= 0 else p)
            data.append(v)
        Hs = scipy.sparse.csr_matrix((np.concatenate(data), (np.concatenate(rows_), np.concatenate(cols_))),
                                     shape=(N, N))
        feats = np.array([S.features(norm1, t, len(offs), bw, N, tol) for t, tol in zip(ts, tols)])
        props = sel.predict(feats)
        B1 = ev.BatchedEvolution([D], n_states=1, value_precision="fp64")
        for k, (t, tol, (m, s)) in enumerate(zip(ts, tols, props)):
            psi = Hm.random_state(N, seed=1000 * h + k)
            B1.set_states(0, psi[None, :])
            B1.set_schedule([t], [m], [s])
            est = B1.probe()[0]
            mp, sp = ev.plain_rule(norm1, t, tol)
            if not est <= tol:
                fallbacks += 1
                m, s = mp, sp
            cost_learned += m * 2 ** s
            cost_plain += mp * 2 ** sp
            pr = torch.as_tensor(psi.real, device="cuda").contiguous()
            pi = torch.as_tensor(psi.imag, device="cuda").contiguous()
            d.hamsim_device(D, pr.data_ptr(), pi.data_ptr(), t, 1 << s, m, sid)
            got = (pr + 1j * pi).cpu().numpy()
            ref = scipy.sparse.linalg.expm_multiply(-1j * t * Hs, psi)
            ratios.append(float(np.linalg.norm(got - ref)) / tol)
            count += 1
        B1.close()
    ratios = np.array(ratios)
    print(f"selector on {count} held-out cases: fallbacks {fallbacks}, error/tol max {ratios.max():.2f} median "
          f"{np.median(ratios):.2e}, SpMVs learned {cost_learned} vs plain {cost_plain} "
          f"({cost_plain / cost_learned:.2f}x)")
    assert count == 200
    assert ratios.max() <= 1.0
    assert fallbacks <= 20
    assert cost_learned * 2 < cost_plain


# ---------------------------------------------------------------------------------------------------------------------
# FP32 expm_negative
@pytest.mark.parametrize("n,t", [(9, 2.5), (12, 0.1)])
def test_expm_negative_fp32_vs_fp64_dense_oracle(d, n, t):
    """FP32 values and Taylor terms, FP64 1-norm and squarings: the offsets equal the FP64 result's, and the values
    are within 4e-7 * (2^s + 1) (max-abs error over max-abs value) of the FP64 dense oracle. FP32 rounding of U's
    Taylor sum (~1e-7 relative) grows at most 2^s-fold through the squarings; the final rounding to FP32 adds one
    more unit."""
    rows, cols, offs, vals = Hm.tfim(n, 1.0, 0.7)
    M = O.Dia(rows, cols, offs, vals)
    Ud, roffs, _, s = O.expm_negative_dense(O.to_dense(M), offs, t)
    D32 = d.to_device(d.diaQMatrix.from_diagonals([rows, cols], offs, vals), precision="fp32")
    U = d.expm_negative(D32, t)
    assert U.precision == "fp32"
    assert list(U.diagonals) == roffs
    err = rel(d.to_numpy(U.to_host()), Ud)
    print(f"FP32 expm_negative TFIM-{n}, t = {t}: s = {s}, max rel error {err:.2e} against the FP64 dense oracle")
    assert err < 4e-7 * (2 ** s + 1)
