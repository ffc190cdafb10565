"""GPU tests of the batched evolution plan (FP32 Taylor terms, FP64 states, HIP graph) against the FP64 oracle."""
import numpy as np
import pytest
import scipy.linalg

pytestmark = pytest.mark.gpu

from oracle import oracle as O  # noqa: E402
from diaq_amd import hamiltonians as Hm  # noqa: E402


@pytest.fixture(scope="module")
def mods():
    import diaq_amd  # noqa: F401
    import diaq_gpu
    from diaq_amd import evolution
    return diaq_gpu, evolution


def small_batch(d, seed=0, n_states=4):
    rng = np.random.default_rng(seed)
    jobs = []
    for n in (6, 7, 8, 6, 8, 9):
        J, h = rng.uniform(0.5, 1.5), rng.uniform(0.2, 2.0)
        t = float(np.exp(rng.uniform(np.log(0.05), np.log(2.0))))
        r, c, offs, vals = Hm.tfim(n, J, h)
        D = d.to_device(d.diaQMatrix.from_diagonals([r, c], offs, vals))
        st = rng.standard_normal((n_states, r)) + 1j * rng.standard_normal((n_states, r))
        st /= np.linalg.norm(st, axis=1, keepdims=True)
        jobs.append(dict(n=n, t=t, D=D, dense=O.to_dense(O.Dia(r, c, offs, vals)), states=st))
    return jobs


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_batched_vs_dense_expm_and_graph_replays(mods, prec):
    d, ev = mods
    jobs = small_batch(d)
    tol = 1e-8
    sched = [ev.plain_rule(d.estimate_1norm(j["D"]), j["t"], tol) for j in jobs]
    B = ev.BatchedEvolution([j["D"] for j in jobs], n_states=4, value_precision=prec)
    assert B.n_classes == 4  # sizes 6, 7, 8, 9
    B.set_schedule([j["t"] for j in jobs], [m for m, s in sched], [s for m, s in sched])

    def load():
        for q, j in enumerate(jobs):
            B.set_states(q, j["states"])

    load()
    B.run(use_graph=False, renormalize=True)
    eager = [B.get_states(q, 1 << j["n"]) for q, j in enumerate(jobs)]
    worst = 0.0
    for j, got in zip(jobs, eager):
        U = scipy.linalg.expm(-1j * j["t"] * j["dense"])
        for st, g in zip(j["states"], got):
            worst = max(worst, np.linalg.norm(g - U @ st))
    assert worst < 1e-6, worst
    for _ in range(3):  # graph replays, each from the initial states, each equal to the eager run
        load()
        B.run(use_graph=True, renormalize=True)
        for q, j in enumerate(jobs):
            np.testing.assert_array_equal(B.get_states(q, 1 << j["n"]), eager[q])


def test_graph_follows_new_schedules(mods):
    """A new (t, m, s) that keeps every class's number of Horner levels (max over its jobs of m 2^s) replays the
    captured graph, which must then run the new schedule; one that changes a level count is captured again. Every
    graph run equals the eager run of the same schedule bit for bit."""
    d, ev = mods
    jobs = small_batch(d, seed=3, n_states=2)  # sizes 6, 7, 8, 6, 8, 9: classes {0, 3}, {1}, {2, 4}, {5}
    B = ev.BatchedEvolution([j["D"] for j in jobs], n_states=2, value_precision="fp32")

    def run(t, ms, graph):
        B.set_schedule(t, [m for m, s in ms], [s for m, s in ms])
        for q, j in enumerate(jobs):
            B.set_states(q, j["states"])
        B.run(use_graph=graph, renormalize=True)
        return [B.get_states(q, 1 << j["n"]) for q, j in enumerate(jobs)]

    t0 = [j["t"] for j in jobs]
    t1 = [0.6 * t for t in t0]
    levels = [(6, 1), (5, 2), (8, 0), (4, 1), (7, 1), (3, 2)]  # per class: 12, 20, 14, 12 levels
    same = [(3, 2), (10, 1), (7, 1), (12, 0), (2, 2), (6, 1)]  # other jobs' schedules, the same class maxima
    more = [(3, 2), (6, 2), (7, 1), (12, 0), (2, 2), (6, 1)]   # class {1}: 24 levels
    for t, ms in ((t0, levels), (t1, same), (t0, same), (t1, more), (t0, levels)):
        graph = run(t, ms, True)
        eager = run(t, ms, False)
        for g, e in zip(graph, eager):
            np.testing.assert_array_equal(g, e)
    # the runs above differ from each other: a stale schedule could not pass
    assert not np.array_equal(run(t0, levels, True)[0], run(t1, same, True)[0])


def test_batched_matches_taylor_action_oracle_for_any_schedule(mods):
    d, ev = mods
    jobs = small_batch(d, seed=1, n_states=2)
    ms = [(3, 1), (5, 0), (2, 3), (7, 2), (1, 4), (4, 2)]
    B = ev.BatchedEvolution([j["D"] for j in jobs], n_states=2, value_precision="fp64")
    B.set_schedule([j["t"] for j in jobs], [m for m, s in ms], [s for m, s in ms])
    for q, j in enumerate(jobs):
        B.set_states(q, j["states"])
    B.run(use_graph=True, renormalize=False)
    for q, (j, (m, s)) in enumerate(zip(jobs, ms)):
        got = B.get_states(q, 1 << j["n"])
        for st, g in zip(j["states"], got):
            ref = O.taylor_action_ms(j["dense"], st, j["t"], m, s)
            assert np.linalg.norm(g - ref) / np.linalg.norm(ref) < 2e-6


def test_probe_flags_too_small_degree(mods):
    d, ev = mods
    jobs = small_batch(d, seed=2)
    B = ev.BatchedEvolution([j["D"] for j in jobs], n_states=4)
    good = [ev.plain_rule(d.estimate_1norm(j["D"]), j["t"], 1e-8) for j in jobs]
    B.set_schedule([j["t"] for j in jobs], [m for m, s in good], [s for m, s in good])
    for q, j in enumerate(jobs):
        B.set_states(q, j["states"])
    est_good = B.probe()
    assert np.all(est_good < 1e-6), est_good
    B.set_schedule([j["t"] for j in jobs], [2] * len(jobs), [0] * len(jobs))
    est_bad = B.probe()
    assert np.all(est_bad > 1e-6), est_bad


@pytest.mark.parametrize("n_states", [4, 8])
def test_every_horner_schedule_gives_the_same_bits(mods, n_states):
    """Rows per lane, the row-tile order, the LDS window, the kernel form (a lane per state, or four states per lane
    with 16-byte gathers, with 1, 2 or 4 diagonals per load batch) and the streams the classes run on only change which
    lane computes a (row, state), where its inputs come from and when, not its arithmetic: every schedule, forced
    through the process-wide knobs, reproduces the default run bit for bit, directly and through the captured graph,
    and so does the tuned plan."""
    from diaq_amd import capi
    d, ev = mods
    jobs = small_batch(d, seed=3, n_states=n_states)
    sched = [ev.plain_rule(d.estimate_1norm(j["D"]), j["t"], 1e-8) for j in jobs]
    B = ev.BatchedEvolution([j["D"] for j in jobs], n_states=n_states, value_precision="fp32")
    B.set_schedule([j["t"] for j in jobs], [m for m, s in sched], [s for m, s in sched])
    L = capi.lib()

    def run():
        for q, j in enumerate(jobs):
            B.set_states(q, j["states"])
        B.run(use_graph=False, renormalize=True)
        return [B.get_states(q, 1 << j["n"]) for q, j in enumerate(jobs)]

    capi.check(L.dq_set_tuning(b"evo_vec", 0))
    base = run()
    try:
        for vec in (0, 1):
            for rpt in (1, 2, 4, 8):
                for xcd in (0, 1):
                    for lds in ((0, 1) if vec == 0 else (0,)):
                        for du in ((1, 2, 4) if vec == 1 and rpt <= 2 else (0,)):
                            capi.check(L.dq_set_tuning(b"evo_vec", vec))
                            capi.check(L.dq_set_tuning(b"evo_rows_per_lane", rpt))
                            capi.check(L.dq_set_tuning(b"evo_xcd", xcd))
                            capi.check(L.dq_set_tuning(b"evo_lds", lds))
                            capi.check(L.dq_set_tuning(b"evo_diag_batch", du))
                            for a, b in zip(run(), base):
                                np.testing.assert_array_equal(
                                    a, b, err_msg=f"vec={vec} rpt={rpt} xcd={xcd} lds={lds} du={du}")
        capi.check(L.dq_set_tuning(b"evo_vec", 1))
        capi.check(L.dq_set_tuning(b"evo_rows_per_lane", 1))
        capi.check(L.dq_set_tuning(b"evo_diag_batch", 2))
        for lanes in (0, 1, 2, 5):
            capi.check(L.dq_set_tuning(b"evo_lanes", lanes))
            for q, j in enumerate(jobs):
                B.set_states(q, j["states"])
            B.run(use_graph=True, renormalize=True)
            for q, j in enumerate(jobs):
                np.testing.assert_array_equal(B.get_states(q, 1 << j["n"]), base[q], err_msg=f"lanes={lanes}")
    finally:
        capi.check(L.dq_set_tuning(b"evo_rows_per_lane", 0))
        capi.check(L.dq_set_tuning(b"evo_xcd", -1))
        capi.check(L.dq_set_tuning(b"evo_lds", -1))
        capi.check(L.dq_set_tuning(b"evo_vec", -1))
        capi.check(L.dq_set_tuning(b"evo_diag_batch", 0))
        capi.check(L.dq_set_tuning(b"evo_lanes", 4))
    chosen = B.tune()
    assert len(chosen) == B.n_classes and all(r in (1, 2, 4, 8) and x in (0, 1) and w in (0, 1, 2, 3, 4)
                                              for r, x, w in chosen)
    for a, b in zip(run(), base):
        np.testing.assert_array_equal(a, b)


def test_zero_block_skip_matches_full_reads(mods):
    """Blocks of a diagonal whose values are all zero (a TFIM's flip diagonals are zero on alternate runs of 2^k rows)
    add nothing: skipping them agrees with reading them to FP32 rounding, stays schedule-independent bit for bit, and
    a Hamiltonian with a whole zero diagonal (h = 0 on one site) evolves as the dense expm says."""
    from diaq_amd import capi
    d, ev = mods
    L = capi.lib()
    jobs = small_batch(d, seed=4, n_states=8)
    rng = np.random.default_rng(5)
    r, c, offs, vals = Hm.tfim(10, 1.1, 0.6)
    vals = [np.array(v, dtype=complex) for v in vals]
    for o in (1 << 7, -(1 << 7)):  # no field on qubit 7: both of its flip diagonals zero everywhere
        vals[offs.index(o)][:] = 0.0
    D = d.to_device(d.diaQMatrix.from_diagonals([r, c], offs, vals))
    st = rng.standard_normal((8, r)) + 1j * rng.standard_normal((8, r))
    st /= np.linalg.norm(st, axis=1, keepdims=True)
    jobs.append(dict(n=10, t=0.7, D=D, dense=O.to_dense(O.Dia(r, c, offs, vals)), states=st))
    sched = [ev.plain_rule(d.estimate_1norm(j["D"]), j["t"], 1e-8) for j in jobs]
    B = ev.BatchedEvolution([j["D"] for j in jobs], n_states=8, value_precision="fp32")
    B.set_schedule([j["t"] for j in jobs], [m for m, s in sched], [s for m, s in sched])

    def run():
        for q, j in enumerate(jobs):
            B.set_states(q, j["states"])
        B.run(use_graph=False, renormalize=True)
        return [B.get_states(q, 1 << j["n"]) for q, j in enumerate(jobs)]

    skip = run()
    try:
        capi.check(L.dq_set_tuning(b"evo_skip_zero_blocks", 0))
        full = run()
        capi.check(L.dq_set_tuning(b"evo_skip_zero_blocks", 1))
        for vec in (0, 1):
            for rpt in (1, 8):
                capi.check(L.dq_set_tuning(b"evo_vec", vec))
                capi.check(L.dq_set_tuning(b"evo_rows_per_lane", rpt))
                for a, b in zip(run(), skip):
                    np.testing.assert_array_equal(a, b)
    finally:
        capi.check(L.dq_set_tuning(b"evo_skip_zero_blocks", 1))
        capi.check(L.dq_set_tuning(b"evo_rows_per_lane", 0))
        capi.check(L.dq_set_tuning(b"evo_vec", -1))
    for a, b in zip(skip, full):
        assert np.max(np.abs(a - b)) < 1e-6
    for j, got in zip(jobs, skip):
        U = scipy.linalg.expm(-1j * j["t"] * j["dense"])
        for s_, g in zip(j["states"], got):
            assert np.linalg.norm(g - U @ s_) < 1e-6
    with pytest.raises(RuntimeError, match="evo_skip_zero_blocks"):
        capi.check(L.dq_set_tuning(b"evo_skip_zero_blocks", 2))
