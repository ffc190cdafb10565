"""Train the learned (m, s) selector on the GPU and evaluate it on held-out cases.

    python -m diaq_amd.selector_train [--train-ham 2000] [--test-ham 200] [--per-ham 10] [--out DIR]

Data: random Pauli sums on 8-16 qubits with 10-60 random 2-to-4-local strings and N(0, 1) coefficients (seeded),
built in DIA and uploaded; for each Hamiltonian the device computes ||H||_1 (dq_estimate_1norm) and bounds its
spectrum by Lanczos with full reorthogonalisation on dq_spmv_device (Ritz values widened by their residual bounds);
each Hamiltonian then gets per_ham cases with t ~ log-uniform[1e-3, 10] and tol ~ log-uniform[1e-10, 1e-4]. A
case's label is the cheapest (m, s) meeting tol on the whole spectral interval (selector.labels_fast); its features
are those of selector.features. 2000 Hamiltonians x 10 cases = 20k training cases, 200 x 10 = 2k held out.

Training: the library's HIP kernels (dq_mlp_train_step: forward, softmax cross-entropy backward, Adam), minibatches
of 256, cosine-decayed learning rate.

Evaluation on the held-out cases: exact accuracy; feasibility of the proposals on the spectral criterion; cost
m 2^s against the plain norm-based rule; the a-posteriori probe's fallback count; and, for cases with n <= 12, the
state error of the FP64 fused evolution run with the proposal (after the probe's fallback) against scipy's FP64
expm_multiply.
Writes selector_weights.npz (next to this file unless --out) and a JSON report.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np


def log(*a):
    print(*a, flush=True)


def lanczos_bounds(torch, d, capi, D, N, k=40, seed=0):
    """Lower and upper bounds of the spectrum of the Hermitian DeviceMatrix D: extreme Ritz values of k Lanczos steps
    (full reorthogonalisation), widened by their residual norms |beta_k s_k|."""
    L = capi.lib()
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    v = torch.complex(torch.randn(N, dtype=torch.float64, device="cuda", generator=g),
                      torch.randn(N, dtype=torch.float64, device="cuda", generator=g))
    v = v / torch.linalg.vector_norm(v)
    k = min(k, N)
    V = torch.empty((k, N), dtype=torch.complex128, device="cuda")
    alphas, betas = [], []
    wr = torch.empty(N, dtype=torch.float64, device="cuda")
    wi = torch.empty(N, dtype=torch.float64, device="cuda")
    sid = torch.cuda.current_stream().cuda_stream
    for j in range(k):
        V[j] = v
        vr, vi = v.real.contiguous(), v.imag.contiguous()
        capi.check(L.dq_spmv_device(D.handle, vr.data_ptr(), vi.data_ptr(), wr.data_ptr(), wi.data_ptr(), sid))
        w = torch.complex(wr, wi)
        a = torch.vdot(v, w).real
        w = w - V[: j + 1].T @ (V[: j + 1].conj() @ w)  # full reorthogonalisation
        w = w - V[: j + 1].T @ (V[: j + 1].conj() @ w)
        b = torch.linalg.vector_norm(w)
        alphas.append(float(a))
        betas.append(float(b))
        if float(b) < 1e-12:
            break
        v = w / b
    n = len(alphas)
    T = np.diag(alphas) + np.diag(betas[: n - 1], 1) + np.diag(betas[: n - 1], -1)
    th, S = np.linalg.eigh(T)
    res = np.abs(betas[n - 1] * S[-1, :])
    return float(th[0] - res[0]), float(th[-1] + res[-1])


def build_cases(torch, d, capi, Hm, S, n_ham, per_ham, seed, tag, want_ham=False):
    rng = np.random.default_rng(seed)
    feats, labels, meta, hams = [], [], [], []
    t0 = time.time()
    for h in range(n_ham):
        n = int(rng.integers(8, 16, endpoint=True))
        P = int(rng.integers(10, 60, endpoint=True))
        N, _, offs, vals, _ = Hm.pauli_sum(n, P, rng)
        D = d.to_device(d.diaQMatrix.from_diagonals([N, N], offs, vals))
        norm1 = d.estimate_1norm(D)
        lo, hi = lanczos_bounds(torch, d, capi, D, N, seed=seed + h)
        rho = max(abs(lo), abs(hi))
        bw = max(abs(o) for o in offs) if offs else 0
        ts = np.exp(rng.uniform(np.log(1e-3), np.log(10.0), per_ham))
        tols = 10.0 ** rng.uniform(-10, -4, per_ham)
        lab = S.labels_fast(rho * ts, tols)
        for t, tol, c in zip(ts, tols, lab):
            feats.append(S.features(norm1, t, len(offs), bw, N, tol))
            labels.append(int(c))
            meta.append(dict(ham=h, n=n, P=P, norm1=norm1, rho=rho, t=float(t), tol=float(tol)))
        if want_ham:
            hams.append((n, offs, vals))
        if h % 100 == 0:
            log(f"[selector] {tag}: {h}/{n_ham} Hamiltonians, {time.time() - t0:.0f}s")
    return np.array(feats), np.array(labels), meta, hams


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train-ham", type=int, default=2000)
    ap.add_argument("--test-ham", type=int, default=200)
    ap.add_argument("--per-ham", type=int, default=10)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--out", default=None)
    ap.add_argument("--eval-states", type=int, default=400, help="held-out cases (n <= 12) run on the GPU")
    a = ap.parse_args()
    import torch
    import diaq_amd  # noqa: F401
    import diaq_gpu as d
    from diaq_amd import capi, evolution as ev, hamiltonians as Hm, selector as S

    out_dir = a.out or os.path.dirname(os.path.abspath(S.__file__))
    Xtr, ytr, _, _ = build_cases(torch, d, capi, Hm, S, a.train_ham, a.per_ham, 1000, "train")
    Xte, yte, mte, hte = build_cases(torch, d, capi, Hm, S, a.test_ham, a.per_ham, 9000, "held-out", want_ham=True)
    keep = ytr >= 0
    log(f"[selector] {keep.sum()} / {len(ytr)} training cases have a feasible (m, s) in the grid")
    Xtr, ytr = Xtr[keep], ytr[keep]
    mean, std = Xtr.mean(0), Xtr.std(0) + 1e-12
    sel = S.Selector(mean=mean, std=std, seed=1)
    xd = torch.as_tensor(((Xtr - mean) / std).astype(np.float32), device="cuda")
    yd = torch.as_tensor(ytr.astype(np.int32), device="cuda")
    B = 256
    steps_per_epoch = max(1, len(ytr) // B)
    total = a.epochs * steps_per_epoch
    rng = np.random.default_rng(0)
    t0 = time.time()
    step = 0
    for ep in range(a.epochs):
        perm = torch.as_tensor(rng.permutation(len(ytr)), device="cuda")
        losses = []
        for i in range(steps_per_epoch):
            idx = perm[i * B:(i + 1) * B]
            lr = 3e-3 * 0.5 * (1 + math.cos(math.pi * step / total)) + 1e-5
            losses.append(sel.train_step(xd[idx].contiguous(), yd[idx].contiguous(), lr))
            step += 1
        if ep % 10 == 0 or ep == a.epochs - 1:
            log(f"[selector] epoch {ep}: loss {np.mean(losses):.4f} ({time.time() - t0:.0f}s)")
    train_acc = float(np.mean(sel.predict_classes(Xtr) == ytr))

    # held-out evaluation
    pred = sel.predict_classes(Xte)
    valid = yte >= 0
    acc = float(np.mean(pred[valid] == yte[valid]))
    rho_t = np.array([m["rho"] * m["t"] for m in mte])
    tol = np.array([m["tol"] for m in mte])
    X, E = S._error_table()
    idx = np.clip(np.searchsorted(X, rho_t), 0, X.size - 1)
    feasible = E[pred, idx] <= tol
    cost_l = np.array([S.ms_of(int(c))[0] * 2 ** S.ms_of(int(c))[1] for c in pred])
    plain = [ev.plain_rule(m["norm1"], m["t"], m["tol"]) for m in mte]
    cost_p = np.array([pm * 2 ** ps for pm, ps in plain])
    cost_opt = np.array([S.ms_of(int(c))[0] * 2 ** S.ms_of(int(c))[1] if c >= 0 else np.nan for c in yte])
    log(f"[selector] held-out: accuracy {acc:.3f}, feasible {feasible.mean():.3f}, mean cost learned "
        f"{cost_l.mean():.1f} vs plain {cost_p.mean():.1f} vs optimal {np.nanmean(cost_opt):.1f}")

    # run a sample of held-out cases (n <= 12) with the proposal, probe first; compare with expm_multiply
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    small = [i for i, m in enumerate(mte) if m["n"] <= 12][: a.eval_states]
    errs, fallbacks, final_errs = [], 0, []
    by_ham = {}
    for i in small:
        by_ham.setdefault(mte[i]["ham"], []).append(i)
    for h, cases in by_ham.items():
        n, offs, vals = hte[h]
        N = 1 << n
        D = d.to_device(d.diaQMatrix.from_diagonals([N, N], offs, vals))
        rows, cols, data = [], [], []
        for dd, v in zip(offs, vals):
            p = np.arange(len(v))
            rows.append(p if dd >= 0 else p - dd)
            cols.append(p + dd if dd >= 0 else p)
            data.append(v)
        Hs = sp.csr_matrix((np.concatenate(data), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
        for i in cases:
            m_, s_ = S.ms_of(int(pred[i]))
            t, tl = mte[i]["t"], mte[i]["tol"]
            psi = Hm.random_state(N, seed=i)
            B1 = ev.BatchedEvolution([D], n_states=1, value_precision="fp64")
            B1.set_states(0, psi[None, :])
            B1.set_schedule([t], [m_], [s_])
            est = B1.probe()[0]
            B1.close()
            if not est <= tl:
                fallbacks += 1
                m_, s_ = ev.plain_rule(mte[i]["norm1"], t, tl)
            # truncation error of the schedule: the FP64 fused evolution (2^s steps of degree m) vs expm_multiply
            pr = torch.as_tensor(psi.real, device="cuda").contiguous()
            pi = torch.as_tensor(psi.imag, device="cuda").contiguous()
            d.hamsim_device(D, pr.data_ptr(), pi.data_ptr(), t, 1 << s_, m_, torch.cuda.current_stream().cuda_stream)
            got = (pr + 1j * pi).cpu().numpy()
            ref = spla.expm_multiply(-1j * t * Hs, psi)
            final_errs.append(float(np.linalg.norm(got - ref)))
            errs.append(final_errs[-1] / tl)
    final_errs = np.array(final_errs)
    errs = np.array(errs)
    log(f"[selector] ran {len(small)} held-out cases (n <= 12): fallbacks {fallbacks}, error/tol median "
        f"{np.median(errs):.2e} max {errs.max():.2e}")
    sel.save(os.path.join(out_dir, "selector_weights.npz"))
    report = dict(train_cases=int(len(ytr)), train_hamiltonians=a.train_ham, heldout_cases=int(len(yte)),
                  heldout_hamiltonians=a.test_ham, train_accuracy=train_acc, heldout_accuracy=acc,
                  heldout_feasible_fraction=float(feasible.mean()), mean_cost_learned=float(cost_l.mean()),
                  mean_cost_plain=float(cost_p.mean()), mean_cost_optimal=float(np.nanmean(cost_opt)),
                  cost_ratio_plain_over_learned=float(cost_p.sum() / cost_l.sum()),
                  run_cases=len(small), run_fallbacks=fallbacks,
                  run_error_over_tol_median=float(np.median(errs)), run_error_over_tol_max=float(errs.max()),
                  run_max_state_error=float(final_errs.max()),
                  epochs=a.epochs, batch=B, optimizer="Adam (0.9, 0.999), lr 3e-3 cosine")
    with open(os.path.join(out_dir, "selector_report.json"), "w") as f:
        json.dump(report, f, indent=1)
    log(json.dumps(report))


if __name__ == "__main__":
    sys.exit(main())
