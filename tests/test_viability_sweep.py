"""CPU checks of the viability calibration helpers: the vectorised budget rule with the reference's constants is the
oracle's dia_viability (src/dia_viability.hpp:42-63), and the fit keeps no slower-in-DIA case viable."""
import numpy as np

from diaq_amd import viability_sweep as V
from oracle import oracle as O


def test_budgets_with_reference_constants_match_the_oracle():
    rng = np.random.default_rng(3)
    n_rows = rng.choice([4096, 65536, 1000, 77], 200)
    n_diags = rng.integers(1, 2000, 200)
    nnz = (n_diags * n_rows * rng.random(200)).astype(np.int64)
    n_right = rng.choice([1, 16, 31, 32, 128], 200)
    got = V.budgets(n_rows, n_diags, nnz, n_right, V.REFERENCE)
    want = [O.dia_viability(int(a), int(a), int(b), int(c), int(e))["viable"]
            for a, b, c, e in zip(n_rows, n_diags, nnz, n_right)]
    assert list(got) == want


def test_fit_has_no_false_viable_cases():
    rng = np.random.default_rng(4)
    cases = []
    for n_rows in (4096, 65536):
        for nd in (1, 4, 16, 64, 256, 1024):
            for fill in (0.1, 0.5, 1.0):
                for nr in (1, 16, 128):
                    nnz = int(nd * n_rows * fill)
                    dia = nd * n_rows * 16 * (1 + nr / 8) / 5e9 * (1 + (nd > n_rows / 64) * 5)
                    csr = nnz * 24 * (1 + nr / 4) / 4e9
                    cases.append(dict(n_rows=n_rows, n_diags=nd, nnz=nnz, n_right=nr, dia_ms=dia * rng.uniform(0.9, 1.1),
                                      csr_ms=csr))
    p, fp, tp = V.fit(cases)
    assert fp == 0 and tp >= 1
    s = V.score(cases, p)
    assert s["false_viable"] == 0 and s["dia_wins_kept"] == tp


def test_library_mi355x_rule_is_the_fitted_constants():
    """dq_dia_viability_mi355x applies the constants the sweep fitted on MI355X (profiles/viability_sweep_*.txt)."""
    import ctypes
    from diaq_amd import capi
    fitted = dict(max_bytes_over_csr=6.0, min_budget=64, rows_per_diagonal=128, wide_right=16, wide_budget=64)
    rng = np.random.default_rng(5)
    for _ in range(300):
        n_rows = int(rng.choice([4096, 65536, 1000, 77]))
        nd = int(rng.integers(1, 2000))
        nnz = int(nd * n_rows * rng.random())
        nr = int(rng.choice([1, 15, 16, 128]))
        v = capi.DiaViability()
        capi.check(capi.lib().dq_dia_viability_mi355x(n_rows, n_rows, nd, nnz, nr, 16, 8, ctypes.byref(v)))
        want = V.budgets(np.array([n_rows]), np.array([nd]), np.array([nnz]), np.array([nr]), fitted)[0]
        assert bool(v.viable) == bool(want)
