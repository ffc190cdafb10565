#!/usr/bin/env python3
"""Benchmark of the DiaQ GPU path on one MI355X.

Headline (one JSON line on stdout): DIA SpMV y = A x at N = 2^24 with 64 diagonals at seeded uniform-random offsets
in [-N/2, N/2] (numpy default_rng(0)), complex128 split real/imaginary values drawn on the device (torch, seed 0),
everything resident in HBM. A step is one SpMV; value = compulsory bytes (16*S + 16*N + 16*N, S = stored elements)
per second of the K timed steps, in GB/s; roofline.frac = that over the MI355X HBM peak (8 TB/s).

Also reported: the reference-style CPU SpMV (the C oracle's restatement of the reference's OpenMP/SIMD path) timed on
this host on the same matrix; the time evolution of the 20-qubit XXZ chain (100 steps, dt = 0.05) by the plain FP64
fused path; and the batched evolution of 256 TFIM Hamiltonians x 8 states -- evolutions/s of the batched FP32 plan
with learned (m, s) against the plain FP64 path, with state errors against an FP64 reference.

    python bench.py [--steps K] [--warmup W] [--no-cpu] [--no-evolution] [--no-batched]

Two measurement tools live here as well: `python bench.py viability-sweep` (the DIA-vs-CSR sweep behind the MI355X
viability budgets) and `python bench.py rocprof-summary` (the summaries under profiles/ from rocprofv3's output).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the CPU baselines' OpenMP placement (as the reference's runs: one thread per core, close binding); read by the
# OpenMP runtime when the C oracle first runs, so it is set before anything loads it
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

HBM_PEAK_GBS = 8000.0
METRIC = "SpMV GB/s vs HBM peak; evolutions/s batched-fp32 vs plain fp64 at 1e-6"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_threads() -> int:
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def spmv_section(args, torch, d, capi, Hm, fp32=True):
    N = 1 << 24
    offs = Hm.random_offsets(N, 64, seed=0)
    lens = [Hm.stored_elements(N, N, [o]) for o in offs]
    S = int(sum(lens))
    log(f"[bench] spmv N=2^24 diagonals=64 S={S}")
    L = capi.lib()
    import ctypes
    out = ctypes.c_void_p()
    o32 = np.asarray(offs, np.int32)
    capi.check(L.dq_devmat_zeros(N, N, len(offs), capi.ptr(o32, ctypes.c_int32), capi.DQ_F64, ctypes.byref(out)))
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    re = torch.randn(S, dtype=torch.float64, device="cuda", generator=g)
    im = torch.randn(S, dtype=torch.float64, device="cuda", generator=g)
    capi.check(L.dq_devmat_set_packed_device(out, re.data_ptr(), im.data_ptr(), 0))
    xr = torch.randn(N, dtype=torch.float64, device="cuda", generator=g)
    xi = torch.randn(N, dtype=torch.float64, device="cuda", generator=g)
    yr, yi = torch.empty_like(xr), torch.empty_like(xi)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    sid = stream.cuda_stream
    # one-time schedule selection for this matrix (outside the timed region, like a library's plan step)
    best = ctypes.c_float()
    capi.check(L.dq_spmv_tune(out, 3, sid, ctypes.byref(best)))
    cfg = [ctypes.c_int() for _ in range(5)]
    capi.check(L.dq_spmv_tuned_config(out, *[ctypes.byref(c) for c in cfg]))
    tuned = dict(rows_per_lane=cfg[0].value, nontemporal=cfg[1].value, pipelined=cfg[2].value, lds_window=cfg[3].value,
                 xcd_tiles=cfg[4].value, tune_ms=best.value)
    log(f"[bench] spmv tuned schedule: {tuned}")

    def step():
        capi.check(L.dq_spmv_device(out, xr.data_ptr(), xi.data_ptr(), yr.data_ptr(), yi.data_ptr(), sid))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # an event between consecutive launches on the launch stream: the K steps' total and each launch's own time
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for k in range(args.steps):
        step()
        ev[k + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = ev[0].elapsed_time(ev[-1]) / args.steps
    per_launch = [ev[k].elapsed_time(ev[k + 1]) for k in range(args.steps)]
    ms_median = float(np.median(per_launch))
    bytes_alg = 16 * S + 16 * N + 16 * N
    gbs = bytes_alg / (ms * 1e-3) / 1e9
    log(f"[bench] spmv fp64: {ms:.3f} ms/step (median launch {ms_median:.3f} ms)  {gbs:.1f} GB/s  "
        f"({gbs / HBM_PEAK_GBS:.3f} of peak)  wall {wall:.2f}s")
    res = dict(N=N, S=S, offsets=offs, ms=ms, gbs=gbs, bytes=bytes_alg, tuned=tuned, ms_median=ms_median,
               ms_min=float(min(per_launch)), ms_max=float(max(per_launch)))

    # The same product through the host-vector entry point, for callers that hand over host buffers: x goes up and y
    # comes back over PCIe on every call (pageable numpy arrays, as the reference's spmv_cuda(DeviceMatrix, x)).
    # Context only, never the headline; the full bench only, so that profiling runs time the device entry alone.
    try:
        if getattr(args, "profile", None) is not None:
            raise RuntimeError("skipped in profiling runs")
        hx = [t.cpu().numpy() for t in (xr, xi)]
        hy = [np.empty(N), np.empty(N)]
        args_h = [capi.ptr(a, ctypes.c_double) for a in (*hx, *hy)]
        capi.check(L.dq_spmv_host(out, *args_h))
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            capi.check(L.dq_spmv_host(out, *args_h))
        ms_h = (time.perf_counter() - t0) / reps * 1e3
        same = bool(np.array_equal(hy[0], yr.cpu().numpy()) and np.array_equal(hy[1], yi.cpu().numpy()))
        res["host_vectors"] = dict(ms=ms_h, gbs_alg=bytes_alg / (ms_h * 1e-3) / 1e9, pcie_bytes=4 * 8 * N,
                                   same_bits_as_device_entry=same, reps=reps)
        log(f"[bench] spmv fp64 with host vectors (x up, y down over PCIe per call): {ms_h:.2f} ms = "
            f"{res['host_vectors']['gbs_alg']:.0f} GB/s of algorithmic bytes; same bits {same}")
        del hx, hy
    except Exception as e:  # context only
        res["host_vectors"] = dict(error=str(e))
        log(f"[bench] spmv host vectors failed: {e}")

    # FP32 form of the same matrix (precision form): half the bytes
    try:
        if not fp32:
            raise RuntimeError("skipped")
        out32 = ctypes.c_void_p()
        capi.check(L.dq_devmat_convert(out, capi.DQ_F32, ctypes.byref(out32)))
        capi.check(L.dq_spmv_tune(out32, 3, sid, ctypes.byref(best)))
        xr32, xi32 = xr.float(), xi.float()
        yr32, yi32 = torch.empty_like(xr32), torch.empty_like(xi32)
        for _ in range(args.warmup):
            capi.check(L.dq_spmv_device(out32, xr32.data_ptr(), xi32.data_ptr(), yr32.data_ptr(), yi32.data_ptr(), sid))
        torch.cuda.synchronize()
        e0, e1 = ev[0], ev[1]
        e0.record(stream)
        for _ in range(args.steps):
            capi.check(L.dq_spmv_device(out32, xr32.data_ptr(), xi32.data_ptr(), yr32.data_ptr(), yi32.data_ptr(), sid))
        e1.record(stream)
        torch.cuda.synchronize()
        ms32 = e0.elapsed_time(e1) / args.steps
        b32 = 8 * S + 8 * N + 8 * N
        y64 = torch.complex(yr, yi)
        err32 = float(torch.linalg.vector_norm(torch.complex(yr32.double(), yi32.double()) - y64) /
                      torch.linalg.vector_norm(y64))
        res["fp32"] = dict(ms=ms32, gbs=b32 / (ms32 * 1e-3) / 1e9, rel_l2_err_vs_fp64=err32)
        log(f"[bench] spmv fp32: {ms32:.3f} ms/step  {res['fp32']['gbs']:.1f} GB/s  rel err vs fp64 {err32:.1e}")
        L.dq_devmat_destroy(out32)
        del xr32, xi32, yr32, yi32
    except Exception as e:  # the headline stays valid without the FP32 figure
        res["fp32"] = dict(error=str(e))

    if getattr(args, "vendor", True):
        try:
            res["vendor_csr"] = vendor_csr_spmv(args, torch, re, im, lens, offs, xr, xi, yr, yi, N, stream)
        except Exception as e:  # context only: the headline stays valid without it
            res["vendor_csr"] = dict(error=str(e))
            log(f"[bench] vendor CSR SpMV failed: {e}")
    if not args.no_cpu:
        try:
            res["cpu"] = cpu_baseline(args, torch, re, im, lens, offs, xr, xi, yr, yi, N)
        except Exception as e:
            res["cpu"] = dict(error=str(e))
            log(f"[bench] cpu baseline failed: {e}")
    del re, im
    L.dq_devmat_destroy(out)
    torch.cuda.synchronize()
    return res


def vendor_csr_spmv(args, torch, re, im, lens, offs, xr, xi, yr, yi, N, stream):
    """The same matrix in CSR (int32 indices, complex128 values) multiplied by the vendor library: torch's sparse CSR
    matmul, i.e. hipSPARSE / rocSPARSE SpMV/SpMM on ROCm. Context for the DIA kernel, not a baseline of the
    reference (whose GPU path is DIA only). GB/s is quoted on the DIA algorithmic bytes of the same product, so the
    two rates compare times directly; the CSR kernel itself must move 20 B per nonzero."""
    dev = "cuda"
    parts_r, parts_c = [], []
    for o, ln in zip(offs, lens):
        rlo = -o if o < 0 else 0
        r = torch.arange(rlo, rlo + ln, dtype=torch.int64, device=dev)
        parts_r.append(r)
        parts_c.append(r + o)
    rows = torch.cat(parts_r)
    cols = torch.cat(parts_c)
    del parts_r, parts_c
    order = torch.argsort(rows * N + cols)
    cols_s = cols[order].to(torch.int32)
    vals_s = torch.complex(re[order], im[order])
    crow = torch.zeros(N + 1, dtype=torch.int32, device=dev)
    crow[1:] = torch.cumsum(torch.bincount(rows, minlength=N), 0).to(torch.int32)
    del rows, cols, order
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")  # torch marks sparse CSR as beta
        A = torch.sparse_csr_tensor(crow, cols_s, vals_s, size=(N, N))
    x = torch.complex(xr, xi).unsqueeze(1)
    y = A @ x
    torch.cuda.synchronize()
    ref = torch.complex(yr, yi)
    err = float((y[:, 0] - ref).abs().max() / ref.abs().max())
    reps = max(1, min(args.steps, 10))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        y = A @ x
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    S = int(sum(lens))
    gbs = (16 * S + 32 * N) / (ms * 1e-3) / 1e9
    log(f"[bench] vendor CSR SpMV (torch sparse CSR, int32 indices): {ms:.3f} ms = {gbs:.0f} GB/s on the DIA bytes; "
        f"max rel diff vs DIA {err:.1e}")
    del A, cols_s, vals_s, crow, x, y
    torch.cuda.empty_cache()
    return dict(ms=ms, gbs_on_dia_bytes=gbs, csr_bytes=20 * S + 4 * (N + 1) + 32 * N, max_rel_diff_vs_dia=err)


def cpu_baseline(args, torch, re, im, lens, offs, xr, xi, yr, yi, N):
    """The reference's CPU SpMV as restated by the C oracle (2-D diagonal x tile tasks, thread-local y, merge), on
    this host's cores, on the same matrix and x as the GPU run."""
    from oracle import coracle
    nt = cpu_threads()
    log(f"[bench] cpu baseline: copying the matrix to the host, {nt} threads")
    h_re = re.cpu().numpy()
    h_im = im.cpu().numpy()
    starts = np.zeros(len(lens), np.int64)
    starts[1:] = np.cumsum(lens)[:-1]
    o32 = np.asarray(offs, np.int32)
    h_xr, h_xi = xr.cpu().numpy(), xi.cpu().numpy()
    o_yr, o_yi = np.empty(N), np.empty(N)
    coracle.spmv_tasks_split(N, N, o32, starts, h_re, h_im, h_xr, h_xi, o_yr, o_yi, 0, nt)  # warmup
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        coracle.spmv_tasks_split(N, N, o32, starts, h_re, h_im, h_xr, h_xi, o_yr, o_yi, 0, nt)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    S = int(sum(lens))
    gbs = (16 * S + 32 * N) / t / 1e9
    g = torch.complex(yr, yi).cpu().numpy()
    c = o_yr + 1j * o_yi
    agree = float(np.max(np.abs(g - c)) / np.max(np.abs(c)))
    log(f"[bench] cpu baseline: {t * 1e3:.1f} ms/SpMV  {gbs:.1f} GB/s  GPU-vs-CPU max rel diff {agree:.1e}")
    return dict(value=gbs, unit="GB/s", cores=nt, kind="port", ms=t * 1e3, block_size=coracle.block_size(),
                gpu_vs_cpu_max_rel_diff=agree,
                sample="the full N=2^24, 64-diagonal SpMV (same matrix and x as the GPU run), median of 5 after 1 warmup")


def norm_rule_m(norm1_dt: float, tol: float, m_max: int = 100) -> int:
    """Smallest m with (||H dt||_1)^(m+1) / (m+1)! <= tol (the plain norm rule for a fixed-degree Taylor step)."""
    for m in range(1, m_max + 1):
        if (m + 1) * math.log(norm1_dt) - math.lgamma(m + 2) <= math.log(tol):
            return m
    return m_max


def evolution_section(args, torch, d, capi, Hm, steps=100, fp32=True, m100=True):
    n = 20
    rows, cols, offs, vals = Hm.xxz(n, 1.0, 0.8, seed=0)
    H = d.to_device(d.diaQMatrix.from_diagonals([rows, cols], offs, vals))
    psi0 = Hm.product_state(n, seed=0)
    dt = 0.05
    norm1 = d.estimate_1norm(H)
    # one-time schedule selection for this Hamiltonian (a plan step, outside the timed evolution)
    import ctypes
    best = ctypes.c_float()
    capi.check(capi.lib().dq_spmv_tune(H.handle, 3, 0, ctypes.byref(best)))
    cfg = [ctypes.c_int() for _ in range(5)]
    capi.check(capi.lib().dq_spmv_tuned_config(H.handle, *[ctypes.byref(c) for c in cfg]))
    tol_step = 1e-6 / steps  # per-step truncation budget for a 1e-6 total
    m_plain = norm_rule_m(norm1 * dt, tol_step)
    out = dict(workload=f"XXZ-{n} (Jxy=1, Jz=0.8, h_i~U[-1,1] seed 0), {steps} steps of dt={dt}, product state seed 0",
               norm1=norm1, m_plain=m_plain, tol_step=tol_step,
               schedule=dict(lds_window=cfg[3].value, xcd_tiles=cfg[4].value))
    sid = torch.cuda.current_stream().cuda_stream
    # the first evolution of a matrix times its Taylor-term schedule (kept in the handle): a one-step plan call on a
    # scratch state, outside the timed evolutions, as the SpMV tuning above
    scratch = [torch.as_tensor(v, device="cuda").contiguous() for v in (psi0.real, psi0.imag)]
    d.hamsim_device(H, scratch[0].data_ptr(), scratch[1].data_ptr(), dt, 1, m_plain, sid)
    for label, m in (("plain_fp64", m_plain), ("reference_m100", 100))[: 2 if m100 else 1]:
        pr = torch.as_tensor(psi0.real, device="cuda").contiguous()
        pi = torch.as_tensor(psi0.imag, device="cuda").contiguous()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        d.hamsim_device(H, pr.data_ptr(), pi.data_ptr(), steps * dt, steps, m, sid)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        out[label] = dict(m=m, seconds=t, evolutions_per_s=1.0 / t,
                          norm=float(torch.sqrt(torch.sum(pr * pr + pi * pi))))
        log(f"[bench] evolution {label}: m={m} {t:.3f} s/evolution")
        if label == "plain_fp64":
            ref_state = torch.complex(pr, pi)
        else:
            oracle_state = torch.complex(pr, pi)  # m = 100: truncation far below FP64 rounding
            out["plain_vs_m100_l2"] = float(torch.linalg.vector_norm(oracle_state - ref_state))
    S = H.stored
    out["bytes_per_term_all_stored"] = 16 * S + 64 * rows
    # the term kernel skips the 256-row blocks of a diagonal whose values are all zero: they count for no bytes
    out["bytes_per_term"] = 16 * live_stored(rows, cols, offs, vals, 256) + 64 * rows
    out["plain_fp64"]["gbs"] = out["bytes_per_term"] * m_plain * steps / out["plain_fp64"]["seconds"] / 1e9
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = evolution_cpu_baseline(torch, d, H, offs, vals, rows, psi0, dt, m_plain, steps, sid,
                                                         out["bytes_per_term"])
        except Exception as e:
            out["cpu_baseline"] = dict(error=str(e))
            log(f"[bench] evolution cpu baseline failed: {e}")
    if not fp32:
        return out
    # precision form: FP32 values and Taylor terms (Kahan-compensated row sums), each step summed into the FP64 state,
    # renormalised in FP64 at the end
    H32 = H.astype("fp32")
    pr = torch.as_tensor(psi0.real, device="cuda").contiguous()
    pi = torch.as_tensor(psi0.imag, device="cuda").contiguous()
    scratch = [v.clone() for v in (pr, pi)]
    d.hamsim_device_fp64_state(H32, scratch[0].data_ptr(), scratch[1].data_ptr(), dt, 1, m_plain, sid)  # schedule
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    d.hamsim_device_fp64_state(H32, pr.data_ptr(), pi.data_ptr(), steps * dt, steps, m_plain, sid)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    st = torch.complex(pr, pi)
    b32 = 8 * live_stored(rows, cols, offs, vals, 256) + 32 * rows
    out["fp32"] = dict(m=m_plain, seconds=t, evolutions_per_s=1.0 / t, gbs=b32 * m_plain * steps / t / 1e9,
                       l2_err_vs_plain_fp64=float(torch.linalg.vector_norm(st - ref_state)),
                       speedup_vs_plain_fp64=out["plain_fp64"]["seconds"] / t,
                       state="FP64 (terms in FP32, each step summed in FP64, FP64 renormalisation)")
    if m100:
        out["fp32"]["l2_err_vs_fp64_oracle_m100"] = float(torch.linalg.vector_norm(st - oracle_state))
    log(f"[bench] evolution fp32: m={m_plain} {t:.3f} s/evolution, error vs FP64 {out['fp32']['l2_err_vs_plain_fp64']:.1e}"
        + (f", vs the FP64 m=100 oracle {out['fp32']['l2_err_vs_fp64_oracle_m100']:.1e}" if m100 else ""))
    return out


def evolution_cpu_baseline(torch, d, H, offs, vals, rows, psi0, dt, m, steps, sid, bytes_per_term):
    """One Taylor step of the same evolution on this host's cores with the reference's CPU SpMV (restated by the C
    oracle), scaled to the full number of steps; checked against one GPU step from the same state."""
    from oracle import coracle
    nt = cpu_threads()
    Hp = coracle.packed(offs, vals)
    coracle.taylor_step_tasks(rows, Hp, psi0, dt, m, n_threads=nt)  # warmup
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        cpu = coracle.taylor_step_tasks(rows, Hp, psi0, dt, m, n_threads=nt)
        times.append(time.perf_counter() - t0)
    t1 = float(np.median(times))
    pr = torch.as_tensor(psi0.real, device="cuda").contiguous()
    pi = torch.as_tensor(psi0.imag, device="cuda").contiguous()
    d.hamsim_device(H, pr.data_ptr(), pi.data_ptr(), dt, 1, m, sid)
    torch.cuda.synchronize()
    g = (pr + 1j * pi).cpu().numpy()
    diff = float(np.max(np.abs(g - cpu)) / np.max(np.abs(cpu)))
    per_ev = t1 * steps
    log(f"[bench] evolution cpu baseline: {t1:.2f} s per step, {per_ev:.1f} s per evolution (x{steps}), {nt} threads, "
        f"GPU-vs-CPU max rel diff {diff:.1e}")
    return dict(value=bytes_per_term * m / t1 / 1e9, unit="GB/s", seconds_per_step=t1, seconds_per_evolution=per_ev,
                cores=nt, kind="port",
                sample=f"1 of the {steps} steps (m={m} terms; median of 3 after 1 warmup), scaled by {steps}",
                gpu_vs_cpu_max_rel_diff=diff)


def ntimes_section(args, torch, d, capi, Hm, n=24, n_times=10):
    """spMV_ntimes on resident vectors for the 24-qubit TFIM itself (built on the device; its flip diagonals are zero
    on alternate runs of 2^k rows): ten products per call, with the zero-block skip and without it. Bytes per product
    count only the 256-row blocks holding a nonzero, plus x read and y written."""
    import ctypes
    x_m, z_m, c_m = Hm.pauli_masks(Hm.tfim_terms(n, 1.0, 0.7))
    D = d.pauli_sum_device(n, x_m, z_m, c_m)
    N = 1 << n
    L = capi.lib()
    best = ctypes.c_float()  # the plan step: SpMV schedules, the zero-block-skipping one included
    capi.check(L.dq_spmv_tune(D.handle, 3, 0, ctypes.byref(best)))
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    xr = torch.randn(N, dtype=torch.float64, device="cuda", generator=g)
    xi = torch.randn(N, dtype=torch.float64, device="cuda", generator=g)
    yr, yi = torch.empty_like(xr), torch.empty_like(xi)
    wr, wi = torch.empty_like(xr), torch.empty_like(xi)
    sid = torch.cuda.current_stream().cuda_stream
    out = dict(workload=f"spMV_ntimes, TFIM-{n} (J=1, h=0.7) as built on the device, {n_times} products per call, "
                        "complex128, resident vectors")
    res = {}
    try:
        for skip in (1, 0):
            capi.check(L.dq_set_tuning(b"spmv_skip_zero_blocks", skip))
            d.spmv_ntimes_device(D, xr.data_ptr(), xi.data_ptr(), yr.data_ptr(), yi.data_ptr(), n_times,
                                 wr.data_ptr(), wi.data_ptr(), sid)  # warm
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            d.spmv_ntimes_device(D, xr.data_ptr(), xi.data_ptr(), yr.data_ptr(), yi.data_ptr(), n_times,
                                 wr.data_ptr(), wi.data_ptr(), sid)
            e1.record()
            torch.cuda.synchronize()
            res[skip] = (e0.elapsed_time(e1) / n_times, torch.complex(yr, yi))
    finally:
        capi.check(L.dq_set_tuning(b"spmv_skip_zero_blocks", 1))
    rows, cols, offs, vals = Hm.tfim(n, 1.0, 0.7)
    live = live_stored(rows, cols, offs, vals, 256)
    nbytes = 16 * live + 32 * N
    same = bool(torch.equal(res[0][1], res[1][1]))
    out.update(stored_elements=D.stored, live_stored_elements=live, bytes_per_product=nbytes,
               ms_per_product=res[1][0], gbs=nbytes / (res[1][0] * 1e-3) / 1e9,
               ms_per_product_reading_every_block=res[0][0], same_bits_with_and_without_skip=same)
    log(f"[bench] spMV_ntimes TFIM-{n}: {res[1][0]:.3f} ms per product with the zero-block skip, {res[0][0]:.3f} ms "
        f"without; {out['gbs']:.0f} GB/s of live bytes; same bits {same}")
    return out


def structured_spmv_section(args, torch, d, capi, Hm):
    """SpMV on a matrix of the headline's size with the offsets of a 24-qubit transverse-field Ising Hamiltonian
    (0 and +-2^k, k = 0..23; complex-normal values): what the LDS x window gains where diagonals share x."""
    import ctypes
    n = 24
    N = 1 << n
    offs = sorted([0] + [1 << k for k in range(n)] + [-(1 << k) for k in range(n)])
    L = capi.lib()
    A = ctypes.c_void_p()
    o32 = np.asarray(offs, np.int32)
    capi.check(L.dq_devmat_zeros(N, N, len(offs), capi.ptr(o32, ctypes.c_int32), capi.DQ_F64, ctypes.byref(A)))
    try:
        S = Hm.stored_elements(N, N, offs)
        g = torch.Generator(device="cuda")
        g.manual_seed(2)
        re = torch.randn(S, dtype=torch.float64, device="cuda", generator=g)
        im = torch.randn(S, dtype=torch.float64, device="cuda", generator=g)
        capi.check(L.dq_devmat_set_packed_device(A, re.data_ptr(), im.data_ptr(), 0))
        del re, im
        xr = torch.randn(N, dtype=torch.float64, device="cuda", generator=g)
        xi = torch.randn(N, dtype=torch.float64, device="cuda", generator=g)
        yr, yi = torch.empty_like(xr), torch.empty_like(xi)
        stream = torch.cuda.current_stream()
        sid = stream.cuda_stream
        best = ctypes.c_float()
        capi.check(L.dq_spmv_tune(A, 3, sid, ctypes.byref(best)))
        cfg = [ctypes.c_int() for _ in range(5)]
        capi.check(L.dq_spmv_tuned_config(A, *[ctypes.byref(c) for c in cfg]))
        for _ in range(args.warmup):
            capi.check(L.dq_spmv_device(A, xr.data_ptr(), xi.data_ptr(), yr.data_ptr(), yi.data_ptr(), sid))
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            capi.check(L.dq_spmv_device(A, xr.data_ptr(), xi.data_ptr(), yr.data_ptr(), yi.data_ptr(), sid))
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        nbytes = 16 * S + 32 * N
        gbs = nbytes / (ms * 1e-3) / 1e9
        log(f"[bench] structured spmv (TFIM-24 offsets): {ms:.3f} ms  {gbs:.1f} GB/s ({gbs / HBM_PEAK_GBS:.3f} of peak)")
        return dict(workload="DIA SpMV, N=2^24, offsets 0 and +-2^k (k=0..23, a 24-qubit TFIM's), complex-normal "
                             "values (torch seed 2), complex128", N=N, n_diags=len(offs), stored_elements=S,
                    compulsory_bytes=nbytes, ms=ms, gbs=gbs, frac=gbs / HBM_PEAK_GBS,
                    schedule=dict(rows_per_lane=cfg[0].value, nontemporal=cfg[1].value, pipelined=cfg[2].value,
                                  lds_window=cfg[3].value, xcd_tiles=cfg[4].value))
    finally:
        L.dq_devmat_destroy(A)
        torch.cuda.synchronize()


def spgemm_section(args, torch, d, capi, Hm, precisions=("fp64", "fp32")):
    """SpGEMM H*H for the 20-qubit TFIM (h = 0.7, J = 1) on device matrices, FP64 and FP32: time per product (dry-run
    + kernel, the result allocated by the call), compulsory bytes b*(S_A + S_B + S_C), the output offset count and the
    FP32 result's relative Frobenius error against FP64. `seconds` computes the dry-run in every timed call (a product
    of another structure runs before each, untimed); `seconds_dryrun_reused` repeats the same product back to back,
    where the library reuses the previous call's dry-run."""
    n = 20
    r, c, offs, vals = Hm.tfim(n, 1.0, 0.7)
    H = d.to_device(d.diaQMatrix.from_diagonals([r, c], offs, vals))
    out = dict(workload=f"TFIM-{n} H*H (J=1, h=0.7), DeviceMatrix product (no prune)")
    prods = {}
    for prec in precisions:
        A = H if prec == "fp64" else H.astype("fp32")
        other = d.to_device(d.diaQMatrix.from_diagonals([4, 4], [0], [np.ones(4, complex)]))
        other = other if prec == "fp64" else other.astype("fp32")
        C = d.product(A, A)  # warmup (and the result kept for the error)
        torch.cuda.synchronize()

        def timed(fresh_dryrun):
            nonlocal C
            times = []
            for _ in range(3):
                del C
                if fresh_dryrun:
                    d.product(other, other)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                C = d.product(A, A)
                torch.cuda.synchronize()
                times.append(time.perf_counter() - t0)
            return float(np.median(times))

        t = timed(True)
        t_reused = timed(False)
        b = 16 if prec == "fp64" else 8
        nbytes = b * (2 * A.stored + C.stored)
        out[prec] = dict(seconds=t, n_offsets=len(C.diagonals), stored_result=C.stored, compulsory_bytes=nbytes,
                         gbs=nbytes / t / 1e9, seconds_dryrun_reused=t_reused)
        log(f"[bench] spgemm {prec}: {t * 1e3:.2f} ms  {nbytes / t / 1e9:.1f} GB/s  {len(C.diagonals)} offsets; "
            f"{t_reused * 1e3:.2f} ms with the previous call's dry-run reused")
        prods[prec] = C
    if "fp64" in prods and not args.no_cpu:
        try:
            out["cpu_baseline"] = spgemm_cpu_baseline(offs, vals, r, prods["fp64"], out["fp64"]["compulsory_bytes"])
        except Exception as e:
            out["cpu_baseline"] = dict(error=str(e))
            log(f"[bench] spgemm cpu baseline failed: {e}")
    if "fp32" not in prods:
        return out
    out["offsets_identical"] = list(prods["fp64"].diagonals) == list(prods["fp32"].diagonals)
    out["fp32_rel_frobenius_err"] = d.norm_diff(prods["fp32"].astype("fp64"), prods["fp64"]) / d.norm(prods["fp64"])
    log(f"[bench] spgemm fp32 vs fp64: rel Frobenius {out['fp32_rel_frobenius_err']:.2e}, offsets identical "
        f"{out['offsets_identical']}")
    return out


def spgemm_cpu_baseline(offs, vals, n_rows, C_gpu, nbytes):
    """The reference's CPU SpGEMM (restated by the C oracle: OpenMP over tiles of 4 result diagonals x 2048
    positions, pairs in order) on the same TFIM-20 H*H, this host's cores; checked against the GPU result on three
    result diagonals."""
    from oracle import coracle
    from oracle import oracle as O
    nt = cpu_threads()
    A = coracle.packed(offs, vals)
    res, pairs = O.product_dryrun(offs, offs, n_rows, n_rows)
    ps = np.zeros(len(res) + 1, np.int64)
    ps[1:] = np.cumsum([len(p) for p in pairs])
    sa = np.array([a for p in pairs for a, _ in p], np.int32)
    sb = np.array([b for p in pairs for _, b in p], np.int32)
    coracle.product_tasks(n_rows, n_rows, n_rows, A, A, (res, ps, sa, sb), n_threads=nt)  # warmup
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        out = coracle.product_tasks(n_rows, n_rows, n_rows, A, A, (res, ps, sa, sb), n_threads=nt)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    worst = 0.0
    for o in (0, 1, res[0]):
        k = list(out[0]).index(o)
        L = O.diagonal_length(int(o), n_rows, n_rows)
        cpu = out[2][out[1][k]:out[1][k] + L] + 1j * out[3][out[1][k]:out[1][k] + L]
        g = C_gpu.get_diagonal(int(o))
        worst = max(worst, float(np.max(np.abs(np.asarray(g) - cpu)) / max(1e-300, np.max(np.abs(cpu)))))
    log(f"[bench] spgemm cpu baseline: {t:.2f} s ({nbytes / t / 1e9:.1f} GB/s), {nt} threads, GPU-vs-CPU max rel diff "
        f"{worst:.1e}")
    return dict(value=nbytes / t / 1e9, unit="GB/s", seconds=t, cores=nt, kind="port",
                sample="the full TFIM-20 H*H (same values), median of 5 after 1 warmup", gpu_vs_cpu_max_rel_diff=worst)


def spmm_section(args, torch, d, capi, Hm):
    """spmm_dia_device on torch tensors: C = A B with A in row-aligned DIA form (9 diagonals, offsets -4..4) of
    N = 2^22 rows, B and C N x 8, complex128 (interleaved) -- time per call and compulsory bytes
    16 * (n_diags * N + 2 * N * n_right)."""
    N, n_right = 1 << 22, 8
    offs = torch.arange(-4, 5, dtype=torch.int64, device="cuda")
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    vals = torch.randn(len(offs) * N, dtype=torch.complex128, device="cuda", generator=g)
    B = torch.randn(N, n_right, dtype=torch.complex128, device="cuda", generator=g)
    C = torch.empty(N, n_right, dtype=torch.complex128, device="cuda")
    stream = torch.cuda.current_stream()

    def step():
        d.spmm_dia_device(offs.data_ptr(), vals.data_ptr(), len(offs), N, N, B.data_ptr(), n_right, C.data_ptr(),
                          True, stream.cuda_stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    nbytes = 16 * (len(offs) * N + 2 * N * n_right)
    # spot check against a torch reference on a slice of rows
    rows = torch.arange(1000, 1064, device="cuda")
    ref = torch.zeros(len(rows), n_right, dtype=torch.complex128, device="cuda")
    for k, o in enumerate(range(-4, 5)):
        ref += vals[k * N + rows].unsqueeze(1) * B[rows + o]
    err = float((C[rows] - ref).abs().max() / ref.abs().max())
    log(f"[bench] spmm_dia_device: {ms:.3f} ms  {nbytes / (ms * 1e-3) / 1e9:.1f} GB/s  spot-check rel err {err:.1e}")
    return dict(workload="spmm_dia_device, N=2^22 rows, 9 diagonals (-4..4), B and C N x 8, complex128", ms=ms,
                compulsory_bytes=nbytes, gbs=nbytes / (ms * 1e-3) / 1e9, spot_check_rel_err=err)


def conversion_section(args, torch, d, capi, Hm):
    """from_numpy / to_numpy on the device: the 14-qubit TFIM (J = 1, h = 0.7; 29 diagonals) as a dense 16384 x 16384
    complex128 matrix in HBM (4 GiB) -> DIA (ascending nonzero diagonals) and back. The compulsory bytes of each
    direction are the dense matrix once (16 N^2) plus the stored DIA values once (16 S)."""
    r, c, offs, vals = Hm.tfim(14, 1.0, 0.7)
    N = r
    T = torch.zeros((N, N), dtype=torch.complex128, device="cuda")
    for o, v in zip(offs, vals):
        k = torch.arange(len(v), device="cuda")
        rr, cc = (k, k + o) if o >= 0 else (k - o, k)
        T[rr, cc] = torch.from_numpy(np.ascontiguousarray(v)).to("cuda")
    out = torch.empty_like(T)
    A = d.from_dense_device(T.data_ptr(), N, N)  # warmup
    d.to_dense_device(A, out.data_ptr())
    assert A.diagonals == sorted(offs) and torch.equal(out, T)
    S = sum(len(v) for v in vals)
    nbytes = 16 * (N * N + S)
    reps = max(1, min(args.steps, 10))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        A = d.from_dense_device(T.data_ptr(), N, N)
    t_from = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        d.to_dense_device(A, out.data_ptr())
    t_to = (time.perf_counter() - t0) / reps
    log(f"[bench] from_dense_device TFIM-14 dense {N}x{N}: {t_from * 1e3:.3f} ms ({nbytes / t_from / 1e9:.0f} GB/s); "
        f"to_dense_device {t_to * 1e3:.3f} ms ({nbytes / t_to / 1e9:.0f} GB/s)")
    res = dict(workload="TFIM-14 (J=1, h=0.7) as dense 16384 x 16384 complex128 <-> DIA (29 diagonals), on device",
               compulsory_bytes=nbytes, from_dense_ms=t_from * 1e3, from_dense_gbs=nbytes / t_from / 1e9,
               to_dense_ms=t_to * 1e3, to_dense_gbs=nbytes / t_to / 1e9)
    # every diagonal stored: a dense random matrix (the tiled kernels); compulsory bytes 16 N^2 each way on each side
    del A
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    T.copy_(torch.randn((N, N), dtype=torch.complex128, device="cuda", generator=g))
    A = d.from_dense_device(T.data_ptr(), N, N)
    d.to_dense_device(A, out.data_ptr())
    assert A.num_diags == 2 * N - 1 and torch.equal(out, T)
    del A
    nb_full = 16 * 2 * N * N
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        A = d.from_dense_device(T.data_ptr(), N, N)
        del A
    t_from_f = (time.perf_counter() - t0) / reps
    A = d.from_dense_device(T.data_ptr(), N, N)
    t0 = time.perf_counter()
    for _ in range(reps):
        d.to_dense_device(A, out.data_ptr())
    t_to_f = (time.perf_counter() - t0) / reps
    if os.environ.get("DQ_BENCH_CONVERT_TILES"):  # time every tile shape of the tiled kernels
        for tile in range(4):
            capi.check(capi.lib().dq_set_tuning(b"dense_convert_tile", tile))
            B = d.from_dense_device(T.data_ptr(), N, N)
            del B
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                B = d.from_dense_device(T.data_ptr(), N, N)
                del B
            tf = (time.perf_counter() - t0) / reps
            t0 = time.perf_counter()
            for _ in range(reps):
                d.to_dense_device(A, out.data_ptr())
            tt = (time.perf_counter() - t0) / reps
            log(f"[bench]   tile shape {tile}: from_dense {tf * 1e3:.3f} ms, to_dense {tt * 1e3:.3f} ms")
        capi.check(capi.lib().dq_set_tuning(b"dense_convert_tile", 0))
    log(f"[bench] from_dense_device dense random {N}x{N}: {t_from_f * 1e3:.3f} ms "
        f"({(nb_full + 16 * N * N) / t_from_f / 1e9:.0f} GB/s incl. the scan); to_dense_device {t_to_f * 1e3:.3f} ms "
        f"({nb_full / t_to_f / 1e9:.0f} GB/s)")
    res["dense_random"] = dict(workload="dense random 16384 x 16384 complex128 <-> DIA (all 32767 diagonals)",
                               from_dense_ms=t_from_f * 1e3, to_dense_ms=t_to_f * 1e3,
                               from_dense_gbs=(nb_full + 16 * N * N) / t_from_f / 1e9,
                               to_dense_gbs=nb_full / t_to_f / 1e9)
    return res


def builder_section(args, torch, d, capi, Hm):
    """Hamiltonians built on the device (dq_devmat_from_pauli_sum): the 24-qubit XXZ chain with random fields
    (seed 1; 93 Pauli terms, 47 structural diagonals of 2^24 rows) and the 24-qubit TFIM, against the host chain
    builders plus upload (the reference's route: host arrays, then a DeviceMatrix)."""
    res = {}
    for name, terms, host in (("XXZ-24", Hm.xxz_terms(24, seed=1), lambda: Hm.xxz(24, seed=1)),
                              ("TFIM-24", Hm.tfim_terms(24, 1.0, 0.7), lambda: Hm.tfim(24, 1.0, 0.7))):
        x, z, c = Hm.pauli_masks(terms)
        D = d.pauli_sum_device(24, x, z, c)  # warm
        del D
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        D = d.pauli_sum_device(24, x, z, c)
        torch.cuda.synchronize()
        t_dev = time.perf_counter() - t0
        nb = D.stored * 16
        t0 = time.perf_counter()
        r, cc, offs, vals = host()
        H = d.to_device(d.diaQMatrix.from_diagonals([r, cc], offs, vals))
        torch.cuda.synchronize()
        t_host = time.perf_counter() - t0
        log(f"[bench] pauli_sum_device {name}: {len(terms)} terms, {D.num_diags} diagonals, {t_dev * 1e3:.2f} ms "
            f"({nb / t_dev / 1e9:.0f} GB/s of DIA written); host build + upload {t_host * 1e3:.0f} ms")
        res[name] = dict(terms=len(terms), diagonals=D.num_diags, device_ms=t_dev * 1e3, host_build_upload_ms=t_host * 1e3,
                         dia_bytes=nb)
        del D, H
    return res


def trotter_gates(d, n, dt, J=1.0, h=0.7):
    """The gates of one first-order Trotter step of the n-qubit TFIM H = J sum Z_q Z_{q+1} + h sum X_q:
    exp(-i dt J Z_q Z_{q+1}) (a diagonal 4 x 4 gate on qubits q, q+1) for q = 0..n-2, then exp(-i dt h X_q) (2 x 2)
    for every qubit, as (device gate, before, after) with qubit q = bit q of the basis index."""
    zz = np.exp(-1j * dt * J * np.array([1.0, -1.0, -1.0, 1.0]))
    c, s_ = np.cos(dt * h), np.sin(dt * h)
    GZ = d.to_device(d.diaQMatrix.from_diagonals([4, 4], [0], [zz]))
    GX = d.to_device(d.diaQMatrix.from_diagonals([2, 2], [-1, 0, 1], [np.array([-1j * s_]), np.array([c, c]) + 0j,
                                                                       np.array([-1j * s_])]))
    return [(GZ, n - 2 - q, q) for q in range(n - 1)] + [(GX, n - 1 - q, q) for q in range(n)]


def trotter_section(args, torch, d, capi, Hm, n=28):
    """First-order Trotter steps of the 28-qubit TFIM (J = 1, h = 0.7, dt = 0.05) on the device: every step applies
    27 ZZ rotations and 28 X rotations in place to a 2^28-amplitude state (split re/im, FP64; 4.3 GB), once gate by
    gate (dq_kron_apply_device) and once as a fused plan (dq_gate_plan: 4 passes per step). Algorithmic bytes per
    pass over the state: 32 * 2^n (read and written once). CPU baseline: the reference's host loop nest
    (oracle_kron_apply_tasks, y += G x) on a 2^24-amplitude sample of one ZZ and one X gate."""
    N = 1 << n
    gates = trotter_gates(d, n, 0.05)
    stream = torch.cuda.current_stream()
    g = torch.Generator(device="cuda")
    g.manual_seed(2)
    vr = torch.randn(N, dtype=torch.float64, device="cuda", generator=g)
    vi = torch.randn(N, dtype=torch.float64, device="cuda", generator=g)
    nrm = torch.sqrt((vr * vr + vi * vi).sum())
    vr /= nrm
    vi /= nrm

    def step():
        for G, before, after in gates:
            d.kron_apply_device(G, before, after, vr.data_ptr(), vi.data_ptr(), vr.data_ptr(), vi.data_ptr(), False,
                                stream.cuda_stream)

    def timed(fn, K):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(K):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / K

    K = max(1, min(args.steps, 5))
    ms_step = timed(step, K)  # one gate per pass
    per_gate_ms = ms_step / len(gates)
    gbs = 32 * N / (per_gate_ms * 1e-3) / 1e9
    plan = d.GatePlan(n, [(G, after) for G, before, after in gates])
    ms_fused = timed(lambda: plan.run(vr.data_ptr(), vi.data_ptr(), 1, stream.cuda_stream), K)
    per_pass_ms = ms_fused / plan.n_passes
    gbs_fused = 32 * N / (per_pass_ms * 1e-3) / 1e9
    norm_err = abs(float(torch.sqrt((vr * vr + vi * vi).sum())) - 1.0)
    log(f"[bench] trotter TFIM-{n}: gate by gate {ms_step:.2f} ms per step ({len(gates)} gates, {per_gate_ms:.3f} ms "
        f"each, {gbs:.0f} GB/s); fused plan {ms_fused:.2f} ms per step ({plan.n_passes} passes, {per_pass_ms:.3f} ms "
        f"each, {gbs_fused:.0f} GB/s); norm drift {norm_err:.1e}")
    res = dict(workload=f"first-order Trotter step, TFIM-{n} (J=1, h=0.7, dt=0.05): {n - 1} ZZ + {n} X gates in "
                        f"place on a 2^{n} state, FP64", gates_per_step=len(gates),
               steps_per_s=1e3 / ms_fused, ms_per_step=ms_fused, passes_per_step=plan.n_passes,
               ms_per_pass=per_pass_ms, gbs_per_pass=gbs_fused,
               roofline={"bound": "hbm", "achieved": round(gbs_fused, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs_fused / HBM_PEAK_GBS, 4), "kernel": "gates_pass_kernel"},
               gate_by_gate=dict(ms_per_step=ms_step, ms_per_gate=per_gate_ms, gbs=gbs,
                                 amplitudes_per_s=N / (per_gate_ms * 1e-3)),
               norm_drift=norm_err)
    if not args.no_cpu:
        from oracle import coracle as C
        from oracle import oracle as O
        nc = 24
        Nc = 1 << nc
        rng = np.random.default_rng(0)
        xr, xi = rng.standard_normal(Nc), rng.standard_normal(Nc)
        yr, yi = np.zeros(Nc), np.zeros(Nc)
        zz = np.diag(np.exp(-1j * 0.05 * np.array([1.0, -1.0, -1.0, 1.0])))
        c, s_ = np.cos(0.035), np.sin(0.035)
        rx = np.array([[c, -1j * s_], [-1j * s_, c]])
        cases = [(4, O.from_dense(zz), nc - 2 - 11, 11), (2, O.from_dense(rx), nc - 1 - 5, 5)]
        threads = cpu_threads()
        times = []
        for gsz, M, before, after in cases:
            Gp = (gsz,) + C.packed(M.offsets, M.vals)
            C.kron_apply_tasks(before, Gp, after, xr, xi, yr, yi, threads)  # warm
            t0 = time.perf_counter()
            reps = 3
            for _ in range(reps):
                C.kron_apply_tasks(before, Gp, after, xr, xi, yr, yi, threads)
            times.append((time.perf_counter() - t0) / reps)
        t_gate = sum(times) / len(times)
        res["cpu"] = dict(value=round(Nc / t_gate / 1e9, 4), unit="G amplitudes/s per gate", cores=threads,
                          kind="port", sample=f"one ZZ and one X gate on a 2^{nc}-amplitude state, 3 reps each, "
                                              "the reference's host loop nest (y += G x)")
        res["cpu"]["trotter_steps_per_s_at_2^28"] = Nc / t_gate / N / len(gates)
        res["gpu_vs_cpu_steps_per_s"] = (1e3 / ms_fused) / (Nc / t_gate / N / len(gates))
        log(f"[bench] trotter CPU baseline: {Nc / t_gate / 1e9:.3f} G amplitudes/s per gate on {threads} threads")
    return res


def expm_section(args, torch, d, capi, Hm):
    """expm(-i H t) at t = 0.1 for the 12-qubit TFIM (J = 1, h = 0.7) by Taylor with scaling and squaring on the GPU,
    against scipy's dense expm."""
    import scipy.linalg
    r, c, offs, vals = Hm.tfim(12, 1.0, 0.7)
    A = d.diaQMatrix.from_diagonals([r, c], offs, vals)
    U = d.expm_negative(A, 0.1)  # warmup
    t0 = time.perf_counter()
    U = d.expm_negative(A, 0.1)
    t = time.perf_counter() - t0
    # the same on a DeviceMatrix, the result left on the device (no 0.54 GB download into a host matrix)
    D = d.to_device(A)
    d.expm_negative(D, 0.1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    UD = d.expm_negative(D, 0.1)
    torch.cuda.synchronize()
    t_dev = time.perf_counter() - t0
    # the near-dense squaring alone: U (8191 diagonals = a dense 4096 x 4096 complex matrix) times itself through the
    # dense path (DIA -> dense, the complex FP64 MFMA GEMM, dense -> DIA); 8 N^3 flops
    d.pow(UD, 2)
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        P2 = d.pow(UD, 2)
    torch.cuda.synchronize()
    t_sq = (time.perf_counter() - t0) / reps
    del UD, P2
    sq_flops = 8.0 * r ** 3
    dense = np.zeros((r, c), complex)
    for o, v in zip(offs, vals):
        k = np.arange(len(v))
        rows_, cols_ = (k, k + o) if o >= 0 else (k - o, k)
        dense[rows_, cols_] = v
    ref = scipy.linalg.expm(-1j * 0.1 * dense)
    got = d.to_numpy(U)
    err = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
    log(f"[bench] expm_negative TFIM-12 t=0.1: {t:.3f} s, {U.num_diags()} diagonals, max rel error vs scipy {err:.1e}; "
        f"on a DeviceMatrix (result left on the device) {t_dev:.3f} s")
    log(f"[bench] dense squaring of the 4096 x 4096 propagator (pow(U, 2) on the device): {t_sq * 1e3:.2f} ms, "
        f"{sq_flops / t_sq / 1e12:.1f} TFLOP/s including the conversions")
    return dict(workload="expm(-i H t), TFIM-12 (J=1, h=0.7), t=0.1, defaults max_it 100, cutoff 1, tol 1e-12",
                seconds=t, n_diagonals=U.num_diags(), max_rel_err_vs_scipy_expm=err, seconds_device_matrix=t_dev,
                dense_squaring=dict(seconds=t_sq, flops=sq_flops, tflops_incl_conversions=sq_flops / t_sq / 1e12,
                                    peak_fp64_mfma_tflops=78.6))


def batched_section(args, torch, d, capi, Hm, quick=False):
    """256 TFIM Hamiltonians (n ~ U{10..18}, J ~ U[0.5, 1.5], h ~ U[0.2, 2]) x 8 complex64 unit states, each job
    evolved to t ~ log-uniform[0.01, 5]: the plain FP64 path (one fused evolution per state, (m, s) by the plain norm
    rule) against the batched FP32 plan with learned (m, s) (probe, fallback to the plain rule), both against an FP64
    reference (the plain FP64 path at tol 1e-12)."""
    from diaq_amd import evolution as ev
    from diaq_amd import selector as S
    jobs = Hm.batched_tfim_workload(256, 8, seed=0)
    Ds = []
    for j in jobs:
        r, c, offs, vals = Hm.tfim(j["n"], j["J"], j["h"])
        Ds.append(d.to_device(d.diaQMatrix.from_diagonals([r, c], offs, vals)))
    n_ev = 8 * len(jobs)
    tol = 1e-6
    norms = [d.estimate_1norm(D) for D in Ds]
    plain = [ev.plain_rule(nm, j["t"], tol) for nm, j in zip(norms, jobs)]
    sid = torch.cuda.current_stream().cuda_stream

    def fused(sched, tag):
        out = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for q, (j, D, (m, s)) in enumerate(zip(jobs, Ds, sched)):
            res = []
            for st in j["states"]:
                pr = torch.as_tensor(st.real, device="cuda").contiguous()
                pi = torch.as_tensor(st.imag, device="cuda").contiguous()
                d.hamsim_device(D, pr.data_ptr(), pi.data_ptr(), j["t"], 1 << s, m, sid)
                res.append(torch.complex(pr, pi))
            out.append(torch.stack(res))
            if q % 64 == 0:
                log(f"[bench] batched workload, {tag}: job {q}/{len(jobs)} {time.perf_counter() - t0:.1f}s")
        torch.cuda.synchronize()
        return out, time.perf_counter() - t0

    if not quick:
        ref, t_ref = fused([ev.plain_rule(nm, j["t"], 1e-12) for nm, j in zip(norms, jobs)], "FP64 reference")
        plain_out, t_plain = fused(plain, "plain FP64")
        err_plain = max(float(torch.max(torch.linalg.vector_norm(a - b, dim=1))) for a, b in zip(plain_out, ref))
        log(f"[bench] plain FP64 path: {t_plain:.2f}s for {n_ev} evolutions, max state error {err_plain:.2e}")

    sel = S.Selector.load()
    B = ev.BatchedEvolution(Ds, 8, "fp32")

    def load():
        for q, j in enumerate(jobs):
            B.set_states(q, j["states"])

    def learned_run():
        feats = np.array([S.features(nm, j["t"], len(D.diagonals), max(abs(o) for o in D.diagonals), D.shape[0],
                                     tol / 4) for nm, j, D in zip(norms, jobs, Ds)])
        prop = sel.predict(feats)
        B.set_schedule([j["t"] for j in jobs], [m for m, s in prop], [s for m, s in prop])
        est = B.probe()
        final = [p if e <= tol / 4 else q for p, q, e in zip(prop, plain, est)]
        B.set_schedule([j["t"] for j in jobs], [m for m, s in final], [s for m, s in final])
        ms = B.run(use_graph=True, renormalize=True)
        return prop, final, est, ms

    load()
    B.set_schedule([j["t"] for j in jobs], [m for m, s in plain], [s for m, s in plain])
    tuned = B.tune()
    log("[bench] batched tuned (rows per lane, XCD order, form: 0 lane/state, 1 +LDS, 2 four states/lane, 3/4 +2/4 "
        f"diagonals per load batch) per class: {tuned}")
    load()
    learned_run()  # warm-up: builds the graph once
    load()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    prop, final, est, gpu_ms = learned_run()
    torch.cuda.synchronize()
    t_batched = time.perf_counter() - t0
    if quick:
        log(f"[bench] batched FP32 + learned (m,s): {t_batched:.3f}s ({gpu_ms:.1f} ms graph)")
        return dict(seconds=t_batched, graph_ms=gpu_ms)
    out_b = [torch.as_tensor(B.get_states(q, 1 << j["n"]), device="cuda") for q, j in enumerate(jobs)]
    err_b = max(float(torch.max(torch.linalg.vector_norm(a - b, dim=1))) for a, b in zip(out_b, ref))
    fallbacks = int(sum(1 for p, f in zip(prop, final) if p != f))
    spmv_plain = int(sum(m * 2 ** s for m, s in plain)) * 8
    spmv_learned = int(sum(m * 2 ** s for m, s in final)) * 8
    # compulsory bytes of the graph's Horner levels (8 states): a level with k > 0 reads the FP32 values (8 B per
    # stored element) and per state-row the gathered input, psi's FP32 copy, and writes the output (3 x 8 B); the
    # last level of a sub-step also reads and writes psi in FP64 and writes its FP32 copy (+32 B). The values count
    # only in the 32-row blocks of a diagonal that hold a nonzero: the kernel skips the others, and so do these bytes.
    live = {n: live_stored(*Hm.tfim(n, 1.0, 1.0)) for n in sorted({j["n"] for j in jobs})}
    graph_bytes = sum(2 ** s * (m * 8 * live[j["n"]] + 8 * D.shape[0] * (24 * m + 32))
                      for (m, s), D, j in zip(final, Ds, jobs))
    # the same plan with the plain schedule, to separate the batching/precision gain from the selector's
    load()
    B.set_schedule([j["t"] for j in jobs], [m for m, s in plain], [s for m, s in plain])
    B.run(use_graph=True)
    load()
    t1 = time.perf_counter()
    gpu_ms_plain_sched = B.run(use_graph=True)
    t_b_plain = time.perf_counter() - t1
    log(f"[bench] batched graph: {graph_bytes / 1e9:.2f} GB compulsory, {graph_bytes / (gpu_ms * 1e-3) / 1e9:.0f} GB/s")
    log(f"[bench] batched FP32 + learned (m,s): {t_batched:.3f}s ({gpu_ms:.1f} ms graph), max state error "
        f"{err_b:.2e}, fallbacks {fallbacks}; same plan with the plain schedule {t_b_plain:.3f}s")
    return dict(
        workload="256 TFIM (n~U{10..18}, J~U[0.5,1.5], h~U[0.2,2.0]) x 8 complex64 unit states, t~logU[0.01,5], "
                 "seed 0",
        evolutions=n_ev, tol=tol,
        plain_fp64=dict(seconds=t_plain, evolutions_per_s=n_ev / t_plain, max_state_error=err_plain,
                        spmv_per_state_total=spmv_plain),
        batched_fp32_learned=dict(seconds=t_batched, graph_ms=gpu_ms, evolutions_per_s=n_ev / t_batched,
                                  max_state_error=err_b, fallbacks=fallbacks, spmv_total=spmv_learned,
                                  tuned_rows_per_lane_xcd_lds=tuned, graph_compulsory_bytes=graph_bytes,
                                  graph_gbs=graph_bytes / (gpu_ms * 1e-3) / 1e9),
        batched_fp32_plain_schedule=dict(seconds=t_b_plain, graph_ms=gpu_ms_plain_sched,
                                         evolutions_per_s=n_ev / t_b_plain),
        speedup_vs_plain_fp64=t_plain / t_batched,
        reference="plain FP64 fused path at tol 1e-12", reference_seconds=t_ref)


def live_stored(n_rows, n_cols, offs, vals, rows_per_block=32):
    """Stored elements inside the blocks of rows_per_block rows of a diagonal that hold a nonzero value (the blocks
    the batched Horner kernel reads; 32 rows for 8 states per job)."""
    total = 0
    for o, v in zip(offs, vals):
        rlo = max(0, -o)
        rows = np.nonzero(np.asarray(v) != 0)[0] + rlo
        blocks = np.unique(rows // rows_per_block)
        lo = np.maximum(blocks * rows_per_block, rlo)
        hi = np.minimum((blocks + 1) * rows_per_block, rlo + len(v))
        total += int((hi - lo).sum())
    return total


def measured_traffic():
    """HBM bytes per SpMV launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this benchmark's
    kernel (profiles/spmv_fp64_n2e24_traffic.json, counters scaled by a 2 GiB copy), or None if absent."""
    path = os.path.join(ROOT, "profiles", "spmv_fp64_n2e24_traffic.json")
    try:
        with open(path) as f:
            return round(float(json.load(f)["traffic_bytes"]))
    except (OSError, KeyError, ValueError):
        return None


# ----------------------------------------------------------------------------------------------------------------------
# viability-sweep: the reference's 360-case DIA x dense vs CSR x dense sweep (src/dia_viability.hpp:5-18, 36-63) on
# this GPU -- spmm_dia_device against the vendor CSR x dense product (torch's sparse CSR matmul: hipSPARSE / rocSPARSE)
# -- and the fit of the rule's five constants with no case passing the budgets slower in DIA, as many DIA wins kept as
# possible. The library's dq_dia_viability_mi355x applies the fitted constants.
#
#     python -u bench.py viability-sweep [--quick] [--json OUT]
# ----------------------------------------------------------------------------------------------------------------------
import itertools  # noqa: E402

VIA_ROWS = (4096, 65536)
VIA_DIAGS = (1, 4, 16, 64, 256, 1024)
VIA_LAYOUTS = ("band", "scattered")
VIA_FILLS = (0.1, 0.25, 0.5, 0.75, 1.0)
VIA_RIGHT = (1, 16, 128)

# the reference's constants (src/dia_viability.hpp:36-40)
VIA_REFERENCE = dict(max_bytes_over_csr=0.9, min_budget=16, rows_per_diagonal=32, wide_right=32, wide_budget=8)


def via_budgets(n_rows, n_diags, nnz, n_right, p, itemsize=16, index_bytes=8):
    """The reference's two budgets with constants p (src/dia_viability.hpp:42-63), vectorised over cases."""
    stored = n_diags.astype(float) * n_rows
    dia_bytes = stored * itemsize + n_diags * 8.0
    csr_bytes = nnz.astype(float) * (itemsize + index_bytes) + (n_rows + 1.0) * index_bytes
    storage_ok = (n_diags > 0) & (dia_bytes <= p["max_bytes_over_csr"] * csr_bytes)
    budget = np.maximum(p["min_budget"], n_rows // p["rows_per_diagonal"])
    budget = np.where(n_right >= p["wide_right"], np.minimum(budget, p["wide_budget"]), budget)
    return storage_ok & (n_diags <= budget)


def via_fit(cases):
    """The constants with no false 'viable' (a case that passes but runs slower in DIA) and the most DIA wins;
    ties go to the rule nearest the reference's."""
    n_rows = np.array([c["n_rows"] for c in cases])
    n_diags = np.array([c["n_diags"] for c in cases])
    nnz = np.array([c["nnz"] for c in cases])
    n_right = np.array([c["n_right"] for c in cases])
    faster = np.array([c["dia_ms"] < c["csr_ms"] for c in cases])
    best = None
    grid = itertools.product((0.5, 0.6, 0.7, 0.8, 0.9, 1.0, 1.1, 1.25, 1.5, 2.0, 3.0, 4.0, 6.0, 8.0),
                             (4, 8, 16, 32, 64, 128, 256, 1024), (1, 2, 4, 8, 16, 32, 64, 128),
                             (16, 32, 64, 128, 1 << 30), (4, 8, 16, 32, 64, 128, 256, 1024))
    for ratio, mb, rpd, wr, wb in grid:
        p = dict(max_bytes_over_csr=ratio, min_budget=mb, rows_per_diagonal=rpd, wide_right=wr, wide_budget=wb)
        v = via_budgets(n_rows, n_diags, nnz, n_right, p)
        fp = int(np.sum(v & ~faster))
        tp = int(np.sum(v & faster))
        dist = (abs(np.log(ratio / 0.9)) + abs(np.log2(mb / 16)) + abs(np.log2(rpd / 32)) +
                abs(np.log2(min(wr, 1 << 10) / 32)) + abs(np.log2(wb / 8)))
        key = (fp == 0, tp, -dist)
        if best is None or key > best[0]:
            best = (key, p, fp, tp)
    return best[1], best[2], best[3]


def via_score(cases, p):
    n_rows = np.array([c["n_rows"] for c in cases])
    n_diags = np.array([c["n_diags"] for c in cases])
    nnz = np.array([c["nnz"] for c in cases])
    n_right = np.array([c["n_right"] for c in cases])
    faster = np.array([c["dia_ms"] < c["csr_ms"] for c in cases])
    v = via_budgets(n_rows, n_diags, nnz, n_right, p)
    return dict(viable=int(v.sum()), false_viable=int(np.sum(v & ~faster)), dia_wins_kept=int(np.sum(v & faster)),
                dia_wins_total=int(faster.sum()))


def via_run_case(torch, d, n_rows, n_diags, layout, fill, n_right, rng, reps=5):
    n_cols = n_rows
    if layout == "band":
        offs = np.arange(n_diags) - n_diags // 2
    else:
        offs = np.sort(rng.choice(np.arange(-(n_rows - 1), n_cols), size=n_diags, replace=False))
    dev = "cuda"
    o = torch.tensor(offs, dtype=torch.int64, device=dev)
    r = torch.arange(n_rows, device=dev)
    col = r[None, :] + o[:, None]
    inside = (col >= 0) & (col < n_cols)
    g = torch.Generator(device=dev)
    g.manual_seed(int(rng.integers(1 << 30)))
    keep = inside & (torch.rand(n_diags, n_rows, device=dev, generator=g) < fill)
    vals = torch.randn(n_diags, n_rows, dtype=torch.complex128, device=dev, generator=g) * keep
    nnz = int(keep.sum().item())
    B = torch.randn(n_cols, n_right, dtype=torch.complex128, device=dev, generator=g)
    C = torch.empty(n_rows, n_right, dtype=torch.complex128, device=dev)
    s = torch.cuda.current_stream()

    def dia():
        d.spmm_dia_device(o.data_ptr(), vals.data_ptr(), n_diags, n_rows, n_cols, B.data_ptr(), n_right,
                          C.data_ptr(), True, s.cuda_stream)

    kk, rr = torch.nonzero(keep, as_tuple=True)
    cc = rr + o[kk]
    order = torch.argsort(rr * n_cols + cc)
    rows_s, cols_s, v_s = rr[order], cc[order], vals[kk, rr][order]
    crow = torch.zeros(n_rows + 1, dtype=torch.int64, device=dev)
    crow[1:] = torch.cumsum(torch.bincount(rows_s, minlength=n_rows), 0)
    A = torch.sparse_csr_tensor(crow, cols_s, v_s, size=(n_rows, n_cols))

    def csr():
        return A @ B

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(reps):
            fn()
        e1.record(s)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    t_dia = timed(dia)
    t_csr = timed(csr)
    err = float((C - csr()).abs().max() / max(1e-300, float(C.abs().max())))
    return dict(n_rows=n_rows, n_diags=n_diags, layout=layout, fill=fill, n_right=n_right, nnz=nnz,
                dia_ms=t_dia, csr_ms=t_csr, max_rel_diff=err)


def viability_sweep_main(argv):
    ap = argparse.ArgumentParser(prog="bench.py viability-sweep")
    ap.add_argument("--quick", action="store_true", help="4096 rows only, 1 and 16 right-hand columns")
    ap.add_argument("--json", default=None)
    args = ap.parse_args(argv)
    import torch
    import diaq_amd  # noqa: F401
    import diaq_gpu as d
    if not torch.cuda.is_available():
        sys.exit("bench.py viability-sweep needs a GPU")
    rng = np.random.default_rng(0)
    rows = VIA_ROWS[:1] if args.quick else VIA_ROWS
    right = VIA_RIGHT[:2] if args.quick else VIA_RIGHT
    cases = []
    t0 = time.time()
    for n_rows, nd, layout, fill, nr in itertools.product(rows, VIA_DIAGS, VIA_LAYOUTS, VIA_FILLS, right):
        c = via_run_case(torch, d, n_rows, nd, layout, fill, nr, rng)
        cases.append(c)
        print(f"[sweep] rows {n_rows:6d} diags {nd:5d} {layout:9s} fill {fill:4.2f} right {nr:4d}: "
              f"DIA {c['dia_ms']:.3f} ms, CSR {c['csr_ms']:.3f} ms ({c['csr_ms'] / c['dia_ms']:.2f}x), "
              f"diff {c['max_rel_diff']:.1e}  [{time.time() - t0:.0f} s]", flush=True)
    p, fp, tp = via_fit(cases)
    out = dict(cases=cases, reference_rule=dict(constants=VIA_REFERENCE, **via_score(cases, VIA_REFERENCE)),
               fitted_rule=dict(constants=p, **via_score(cases, p)))
    print(json.dumps({k: v for k, v in out.items() if k != "cases"}, indent=1))
    if args.json:
        with open(args.json, "w") as f:
            json.dump(out, f, indent=1)



# ----------------------------------------------------------------------------------------------------------------------
# rocprof-summary: rocprofv3 results (the rocpd SQLite files) as short text -- per-kernel dispatch count and average
# duration, and per-dispatch FETCH_SIZE / WRITE_SIZE in bytes, each also scaled by what the same counter reads for a
# plain copy kernel of known size (the calibration dispatch every `bench.py --profile` run ends with).
#
#     python bench.py rocprof-summary --stats DB [--fetch DB] [--write DB] [--match spmv_rows_kernel] [--json OUT]
# ----------------------------------------------------------------------------------------------------------------------
import sqlite3  # noqa: E402
import statistics  # noqa: E402

PROF_COPY_BYTES = 1 << 31  # bench.py --profile: dst.copy_(src) of 2^28 doubles


def prof_kernel_stats(db, match=None):
    c = sqlite3.connect(db)
    out = []
    for name, calls, total, avg, pct in c.execute("select name, total_calls, total_duration, average, percentage "
                                                  "from top_kernels"):
        if match and match not in name:
            continue
        out.append((name, calls, total, avg, pct))
    return out


def prof_counter(db, cname, match):
    """Per-dispatch values (bytes) of one counter for the kernel matching `match` -- when several kernel names match
    (an autotuner's candidates), the one dispatched most often, i.e. the schedule the timed run used -- and the
    value the counter reads for the calibration copy."""
    c = sqlite3.connect(db)
    rows = c.execute("select kernel_name, value, duration from counters_collection where counter_name=? order by id",
                     (cname,)).fetchall()
    by_name = {}
    for k, v, d in rows:
        if match in k:
            by_name.setdefault(k, []).append((v * 1024.0, d))
    name = max(by_name, key=lambda k: len(by_name[k])) if by_name else None
    hit = by_name.get(name, [])
    # calibration: the largest copy-like dispatch (the 2 GiB copy that the profiling run ends with)
    cal = [v * 1024.0 for k, v, d in rows if ("copyBuffer" in k or "elementwise" in k) and v * 1024.0 > 1e9]
    return hit, (max(cal) if cal else None), name


def rocprof_summary_main(argv):
    ap = argparse.ArgumentParser(prog="bench.py rocprof-summary")
    ap.add_argument("--stats", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--match", default="spmv_rows_kernel")
    ap.add_argument("--top", type=int, default=6)
    ap.add_argument("--json", help="also write {kernel, fetch, write, factors, traffic_bytes} to this file")
    a = ap.parse_args(argv)
    summary = {}
    print("kernel stats (rocprofv3 --kernel-trace --stats), durations in microseconds")
    for name, calls, total, avg, pct in prof_kernel_stats(a.stats)[: a.top]:
        print(f"  {avg:10.1f} us avg  {calls:4d} calls  {pct:5.1f}%  {name[:110]}")
    for label, db, cname in (("FETCH_SIZE", a.fetch, "FETCH_SIZE"), ("WRITE_SIZE", a.write, "WRITE_SIZE")):
        if not db:
            continue
        hit, cal, name = prof_counter(db, cname, a.match)
        if not hit:
            print(f"{label}: no dispatch matching {a.match}")
            continue
        med = statistics.median(v for v, _ in hit)
        short = name.split("(")[0].replace("void ", "")
        line = f"{label} per dispatch of {short}: median {med / 1e9:.3f} GB over {len(hit)} dispatches"
        factor = PROF_COPY_BYTES / cal if cal else None
        if cal:
            line += (f"; the 2 GiB calibration copy reads {cal / 1e9:.3f} GB, factor {factor:.3f} -> "
                     f"{med * factor / 1e9:.3f} GB")
        print(line)
        summary[label] = dict(raw_bytes=med, factor=factor, bytes=med * (factor or 1.0), kernel=short)
    if a.json and "FETCH_SIZE" in summary and "WRITE_SIZE" in summary:
        import json
        summary["kernel"] = summary["FETCH_SIZE"]["kernel"]
        summary["traffic_bytes"] = summary["FETCH_SIZE"]["bytes"] + summary["WRITE_SIZE"]["bytes"]
        with open(a.json, "w") as f:
            json.dump(summary, f, indent=1)



def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-evolution", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--profile", nargs="?", const="spmv", default=None,
                    choices=["spmv", "spmv_structured", "spmv_ntimes", "spgemm", "spmm_dia", "expm", "evolution",
                             "batched", "conversion", "trotter", "builders"],
                    help="profiling run of one kernel's workload (default: the FP64 SpMV steps), then one plain copy "
                         "of 2 GiB (a kernel of known size to calibrate the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    args.vendor = args.profile is None  # the vendor CSR comparison runs in the full bench only
    if args.gpus != 1:
        sys.exit("bench.py measures one GPU: this build has no multi-GPU path (every box here has one GPU); "
                 "run with --gpus 1")
    import torch
    import diaq_amd  # noqa: F401
    import diaq_gpu as d
    from diaq_amd import capi
    from diaq_amd import hamiltonians as Hm
    if not torch.cuda.is_available():
        sys.exit("bench.py needs a GPU")
    torch.cuda.set_device(0)
    if args.profile:
        args.no_cpu = True
        if args.profile == "spmv":
            sp = spmv_section(args, torch, d, capi, Hm, fp32=False)
            log(f"[bench] profile: spmv {sp['ms']:.3f} ms")
        elif args.profile == "spmv_structured":
            sp = structured_spmv_section(args, torch, d, capi, Hm)
            log(f"[bench] profile: structured spmv {sp['ms']:.3f} ms")
        elif args.profile == "spmv_ntimes":
            sp = ntimes_section(args, torch, d, capi, Hm)
            log(f"[bench] profile: spMV_ntimes {sp['ms_per_product']:.3f} ms per product")
        elif args.profile == "spgemm":
            sg = spgemm_section(args, torch, d, capi, Hm, precisions=("fp64",))
            log(f"[bench] profile: spgemm {sg['fp64']['seconds'] * 1e3:.2f} ms")
        elif args.profile == "spmm_dia":
            sm = spmm_section(args, torch, d, capi, Hm)
            log(f"[bench] profile: spmm_dia {sm['ms']:.3f} ms")
        elif args.profile == "expm":
            ex = expm_section(args, torch, d, capi, Hm)
            log(f"[bench] profile: expm {ex['seconds']:.3f} s (host matrices), {ex['seconds_device_matrix']:.3f} s "
                "(DeviceMatrix)")
        elif args.profile == "conversion":
            cv = conversion_section(args, torch, d, capi, Hm)
            log(f"[bench] profile: from_dense {cv['from_dense_ms']:.3f} ms, to_dense {cv['to_dense_ms']:.3f} ms")
        elif args.profile == "trotter":
            tr = trotter_section(args, torch, d, capi, Hm)
            log(f"[bench] profile: trotter {tr['ms_per_step']:.2f} ms per step (fused plan)")
        elif args.profile == "builders":
            builder_section(args, torch, d, capi, Hm)
        elif args.profile == "batched":
            bt = batched_section(args, torch, d, capi, Hm, quick=True)
            log(f"[bench] profile: batched graph {bt['graph_ms']:.1f} ms")
        else:
            ev = evolution_section(args, torch, d, capi, Hm, steps=10, fp32=False, m100=False)
            log(f"[bench] profile: evolution {ev['plain_fp64']['seconds']:.3f} s")
        src = torch.ones(1 << 28, dtype=torch.float64, device="cuda")  # 2 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] profile: calibration copy of {src.numel() * 8} bytes done")
        return
    sp = spmv_section(args, torch, d, capi, Hm)
    evo = None
    if not args.no_evolution:
        try:
            evo = evolution_section(args, torch, d, capi, Hm)
        except Exception as e:
            evo = dict(error=str(e))
            log(f"[bench] evolution failed: {e}")
    extra = {}
    if not args.no_evolution:
        for key, fn in (("spmv_structured", structured_spmv_section), ("spmv_ntimes", ntimes_section),
                        ("spgemm", spgemm_section),
                        ("spmm_dia", spmm_section), ("expm", expm_section), ("conversion", conversion_section),
                        ("trotter", trotter_section), ("builders", builder_section)):
            try:
                extra[key] = fn(args, torch, d, capi, Hm)
            except Exception as e:
                extra[key] = dict(error=str(e))
                log(f"[bench] {key} failed: {e}")
    bat = None
    if not args.no_batched:
        try:
            bat = batched_section(args, torch, d, capi, Hm)
        except Exception as e:
            bat = dict(error=str(e))
            log(f"[bench] batched section failed: {e}")
    cpu = sp.get("cpu", {})
    line = {
        "metric": METRIC,
        "value": round(sp["gbs"], 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sp["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "DIA SpMV y=Ax, N=2^24, 64 diagonals at uniform-random offsets in [-N/2, N/2] "
                               "(numpy default_rng seed 0), complex128 split re/im, values and x torch.randn seed 0",
                   "N": sp["N"], "n_diags": 64, "stored_elements": sp["S"], "compulsory_bytes": sp["bytes"],
                   "spmv_schedule": sp.get("tuned")},
        "roofline": {"bound": "hbm", "achieved": round(sp["gbs"], 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(sp["gbs"] / HBM_PEAK_GBS, 4), "traffic": measured_traffic(),
                     "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE of this kernel and workload, calibrated by "
                                       "a 2 GiB copy (profiles/spmv_fp64_n2e24_traffic.json), not measured in this run",
                     "launch_ms": {"mean": round(sp["ms"], 4), "median": round(sp["ms_median"], 4),
                                   "min": round(sp["ms_min"], 4), "max": round(sp["ms_max"], 4), "n": args.steps},
                     "achieved_at_median_launch": round(sp["bytes"] / (sp["ms_median"] * 1e-3) / 1e9, 2),
                     "binding": "L2-miss (fill) rate: 1.912e8 misses of 128 B per launch at 6.0e10/s with a "
                                "~1,400-cycle loaded latency and <1 % DRAM credit stalls; x is gathered once per "
                                "diagonal (9 % L2 hits), 1.43x the misses of the structured-offset matrix "
                                "(profiles/spmv_fp64_n2e24_counters.txt)"},
        "cpu_baseline": ({k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample")} if "value" in cpu
                         else None),
        "spmv_fp32": sp.get("fp32"),
        "spmv_host_vectors": sp.get("host_vectors"),
        "spmv_vendor_csr": sp.get("vendor_csr"),
        "cpu_detail": cpu or None,
        "spmv_structured": extra.get("spmv_structured"),
        "spmv_ntimes": extra.get("spmv_ntimes"),
        "evolution": evo,
        "spgemm": extra.get("spgemm"),
        "spmm_dia": extra.get("spmm_dia"),
        "expm": extra.get("expm"),
        "conversion": extra.get("conversion"),
        "trotter": extra.get("trotter"),
        "builders": extra.get("builders"),
        "batched_evolution": bat,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "viability-sweep":
        viability_sweep_main(sys.argv[2:])
    elif len(sys.argv) > 1 and sys.argv[1] == "rocprof-summary":
        rocprof_summary_main(sys.argv[2:])
    else:
        main()
