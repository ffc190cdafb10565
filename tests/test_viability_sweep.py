"""The DIA x dense viability rule in the library (host logic, no GPU needed): dq_dia_viability keeps the reference's
constants and equals the oracle's rule (src/dia_viability.hpp:42-63); dq_dia_viability_mi355x applies the constants
that the MI355X sweep fitted, as the committed sweep table states them."""
import ast
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import oracle as O

TABLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "profiles",
                     "viability_sweep_spmm_dia_vs_csr.txt")


def sweep_rule(label):
    """(constants, viable, slower in DIA, DIA wins kept, DIA wins) of the reference or fitted rule in the table."""
    text = open(TABLE).read()
    m = re.search(label + r" rule: constants (\{[^}]*\}); viable (\d+), of them slower in DIA (\d+); DIA wins kept "
                  r"(\d+) of (\d+)", text)
    assert m, label
    return ast.literal_eval(m.group(1)), int(m.group(2)), int(m.group(3)), int(m.group(4)), int(m.group(5))


def library(fn, n_rows, nd, nnz, nr):
    from diaq_amd import capi
    v = capi.DiaViability()
    capi.check(getattr(capi.lib(), fn)(n_rows, n_rows, nd, nnz, nr, 16, 8, ctypes.byref(v)))
    return v


def random_cases(seed, n=300):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        n_rows = int(rng.choice([4096, 65536, 1000, 77]))
        nd = int(rng.integers(1, 2000))
        yield n_rows, nd, int(nd * n_rows * rng.random()), int(rng.choice([1, 15, 16, 31, 32, 128]))


def test_sweep_table_states_the_reference_constants_and_no_false_viable_case():
    ref, viable, slower, kept, wins = sweep_rule("reference")
    assert ref == O.VIABILITY_CONSTANTS
    assert slower == 0 and viable == kept
    fit, viable, slower, kept, wins = sweep_rule("fitted")
    assert slower == 0 and viable == kept and kept <= wins


@pytest.mark.parametrize("fn,label", [("dq_dia_viability", "reference"), ("dq_dia_viability_mi355x", "fitted")])
def test_library_rule_matches_the_oracle_with_the_table_constants(fn, label):
    constants = sweep_rule(label)[0]
    for n_rows, nd, nnz, nr in random_cases(5 if label == "fitted" else 3):
        v = library(fn, n_rows, nd, nnz, nr)
        want = O.dia_viability(n_rows, n_rows, nd, nnz, nr, constants=constants)
        assert bool(v.viable) == want["viable"], (n_rows, nd, nnz, nr)
        assert bool(v.storage_ok) == want["storage_ok"] and bool(v.diagonals_ok) == want["diagonals_ok"]
        assert v.diagonal_budget == want["diagonal_budget"]
        assert v.dia_bytes == want["dia_bytes"] and v.csr_bytes == want["csr_bytes"]
