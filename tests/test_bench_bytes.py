"""The algorithmic bytes the benchmark reports for kernels that skip all-zero row blocks: bench.live_stored against a
brute-force count and against the closed form for the transverse-field Ising chain (CPU only)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import bench  # noqa: E402
from diaq_amd import hamiltonians as Hm  # noqa: E402


def brute_live(rows, offs, vals, block):
    """Rows r of each diagonal, in [rlo, rlo + len), whose block r // block holds a nonzero value of that diagonal."""
    total = 0
    for o, v in zip(offs, vals):
        rlo = max(0, -o)
        r = np.arange(len(v)) + rlo
        live_blocks = set((r[np.asarray(v) != 0] // block).tolist())
        total += sum(1 for x in r if x // block in live_blocks)
    return total


@pytest.mark.parametrize("n,b", [(8, 5), (10, 5), (10, 8), (12, 5), (12, 8), (7, 3)])
def test_tfim_live_elements_closed_form(n, b):
    """Flip diagonal +-2^k of the TFIM is nonzero exactly where bit k of the row is 0 (+2^k) or 1 (-2^k): blocks of
    2^b rows hold both values when k < b and one when k >= b, so the live elements are
    N (main) + sum_{k<b} 2 (N - 2^k) + sum_{k>=b} N."""
    rows, cols, offs, vals = Hm.tfim(n, 1.0, 0.7)
    N = 1 << n
    closed = N + 2 * b * N - 2 * ((1 << b) - 1) + (n - b) * N
    got = bench.live_stored(rows, cols, offs, vals, 1 << b)
    assert got == closed
    assert got == brute_live(rows, offs, vals, 1 << b)
    assert got < sum(len(v) for v in vals)


def test_live_elements_xxz_and_random_bands():
    rows, cols, offs, vals = Hm.xxz(9, seed=3)
    for block in (4, 32, 256):
        assert bench.live_stored(rows, cols, offs, vals, block) == brute_live(rows, offs, vals, block)
    rng = np.random.default_rng(0)
    n = 1000
    offs = [-700, -3, 0, 5, 999]
    vals = []
    for o in offs:
        v = rng.standard_normal(n - abs(o)) + 0j
        v[rng.random(len(v)) < 0.9] = 0  # mostly zeros, so some blocks are dead
        vals.append(v)
    for block in (8, 32, 256):
        assert bench.live_stored(n, n, offs, vals, block) == brute_live(n, offs, vals, block)
    full = [np.ones(n - abs(o), complex) for o in offs]  # nothing to skip
    assert bench.live_stored(n, n, offs, full, 32) == sum(len(v) for v in full)
