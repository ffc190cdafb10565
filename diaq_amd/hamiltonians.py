"""Synthetic Hamiltonians and states built directly in diagonal (DIA) form.

The reference has no bulk constructor: its Python API builds matrices from dense arrays (from_numpy,
src/dense_to_diagonals.hpp:20-58) or element by element (set_at_diag, src/diaQMatrix.hpp:666-674), which cannot
reach 2^20-2^24 rows. These builders produce the diagonals themselves, in ascending offset order (the order
from_numpy would give), as ``(rows, cols, offsets, vals)`` with ``vals[k]`` the complex128 storage of diagonal
``offsets[k]`` (position p holds A[p, p+d] for d >= 0 and A[p-d, p] for d < 0).

Conventions: qubit i is bit i of the basis index; Z|0> = |0>, so z_i = 1 - 2*bit_i; open boundary.
"""
from __future__ import annotations

import numpy as np


def _bits(n_qubits: int, N: int, i: int, idx=None) -> np.ndarray:
    idx = np.arange(N, dtype=np.int64) if idx is None else idx
    return (idx >> i) & 1


def tfim(n_qubits: int, J: float = 1.0, h: float = 0.7):
    """H = -J sum_{i<n-1} Z_i Z_{i+1} - h sum_i X_i: main diagonal plus offsets +-2^k (2n + 1 diagonals).

    X_k maps |r> to |r xor 2^k>, so H[r, r + 2^k] = -h where bit k of r is 0 (offset +2^k at position p = r) and
    H[r, r - 2^k] = -h where bit k of r is 1 (offset -2^k at position p = r - 2^k, whose bit k is 0): on both
    diagonals the value at position p is -h exactly when bit k of p is 0.
    """
    N = 1 << n_qubits
    idx = np.arange(N, dtype=np.int64)
    z = 1.0 - 2.0 * np.stack([_bits(n_qubits, N, i, idx) for i in range(n_qubits)]).astype(np.float64)
    diag = np.zeros(N)
    for i in range(n_qubits - 1):
        diag += -J * z[i] * z[i + 1]
    diags = {0: diag.astype(np.complex128)}
    for k in range(n_qubits):
        L = N - (1 << k)
        p = np.arange(L, dtype=np.int64)
        v = np.where(((p >> k) & 1) == 0, -h, 0.0).astype(np.complex128)
        diags[1 << k] = v
        diags[-(1 << k)] = v.copy()
    offs = sorted(diags)
    return N, N, offs, [diags[d] for d in offs]


def xxz(n_qubits: int, Jxy: float = 1.0, Jz: float = 0.8, fields=None, seed: int = 0):
    """H = sum_{i<n-1} [Jxy (X_i X_{i+1} + Y_i Y_{i+1}) + Jz Z_i Z_{i+1}] + sum_i h_i Z_i, h_i ~ U[-1, 1] (seeded).

    XX + YY = 2 (s+ s- + s- s+) swaps the bit pair (i, i+1) when the bits differ, with amplitude 2 Jxy; the 00 <-> 11
    parts of XX and YY cancel, so only offsets +-2^i (i = 0..n-2) appear: H[r, r + 2^i] = 2 Jxy where bits (i, i+1)
    of r are (1, 0), and H[r, r - 2^i] = 2 Jxy where they are (0, 1); at position p both diagonals hold 2 Jxy
    exactly when bits (i, i+1) of p are (1, 0). 2(n-1) + 1 diagonals.
    """
    N = 1 << n_qubits
    if fields is None:
        fields = np.random.default_rng(seed).uniform(-1.0, 1.0, n_qubits)
    fields = np.asarray(fields, np.float64)
    idx = np.arange(N, dtype=np.int64)
    z = 1.0 - 2.0 * np.stack([_bits(n_qubits, N, i, idx) for i in range(n_qubits)]).astype(np.float64)
    diag = np.zeros(N)
    for i in range(n_qubits - 1):
        diag += Jz * z[i] * z[i + 1]
    for i in range(n_qubits):
        diag += fields[i] * z[i]
    diags = {0: diag.astype(np.complex128)}
    for i in range(n_qubits - 1):
        L = N - (1 << i)
        p = np.arange(L, dtype=np.int64)
        v = np.where((((p >> i) & 1) == 1) & (((p >> (i + 1)) & 1) == 0), 2.0 * Jxy, 0.0).astype(np.complex128)
        diags[1 << i] = v
        diags[-(1 << i)] = v.copy()
    offs = sorted(diags)
    return N, N, offs, [diags[d] for d in offs]


def random_offsets(N: int, n_diags: int, seed: int = 0) -> list:
    """n_diags distinct integer offsets drawn uniformly without replacement from [-N/2, N/2] (numpy
    default_rng(seed).choice), ascending. N = 2^24, 64 diagonals, seed 0 stores 798,254,616 elements."""
    rng = np.random.default_rng(seed)
    return sorted(int(o) for o in rng.choice(np.arange(-(N // 2), N // 2 + 1), n_diags, replace=False))


def product_state(n_qubits: int, seed: int = 0) -> np.ndarray:
    """(x)_i (cos t_i |0> + e^{i p_i} sin t_i |1>) with t_i ~ U[0, pi), p_i ~ U[0, 2 pi) (seeded): unit norm."""
    rng = np.random.default_rng(seed)
    th = rng.uniform(0.0, np.pi, n_qubits)
    ph = rng.uniform(0.0, 2.0 * np.pi, n_qubits)
    psi = np.ones(1, np.complex128)
    for i in range(n_qubits):  # qubit i is bit i: the later kron factor is the lower bit
        q = np.array([np.cos(th[i]), np.exp(1j * ph[i]) * np.sin(th[i])])
        psi = np.kron(q, psi)
    return psi


def random_state(N: int, seed: int = 0) -> np.ndarray:
    """Complex-normal vector divided by its 2-norm (seeded numpy default_rng)."""
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    return v / np.linalg.norm(v)


def stored_elements(rows: int, cols: int, offsets) -> int:
    """Sum of the diagonal lengths (src/diaQMatrix.hpp:390-409)."""
    s = 0
    for d in offsets:
        n = rows + d if d < 0 else (cols - d if d > 0 else min(rows, cols))
        s += max(0, min(n, rows, cols))
    return s


def batched_tfim_workload(n_ham: int = 256, n_states: int = 8, seed: int = 0, n_min: int = 10, n_max: int = 18):
    """The batched-evolution workload: n_ham TFIM Hamiltonians with n ~ U{n_min..n_max} qubits, J ~ U[0.5, 1.5],
    h ~ U[0.2, 2.0], each with n_states random complex64 unit states, evolved to t ~ log-uniform[0.01, 5] (all
    seeded). Returns a list of dicts (n, J, h, t, states as complex128 holding complex64-rounded values)."""
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(n_ham):
        n = int(rng.integers(n_min, n_max, endpoint=True))
        J = float(rng.uniform(0.5, 1.5))
        h = float(rng.uniform(0.2, 2.0))
        t = float(np.exp(rng.uniform(np.log(0.01), np.log(5.0))))
        N = 1 << n
        st = rng.standard_normal((n_states, N)) + 1j * rng.standard_normal((n_states, N))
        st /= np.linalg.norm(st, axis=1, keepdims=True)
        st = st.astype(np.complex64).astype(np.complex128)
        jobs.append(dict(n=n, J=J, h=h, t=t, states=st))
    return jobs


def pauli_sum(n_qubits: int, n_strings: int, rng, min_weight: int = 2, max_weight: int = 4):
    """Random Hermitian Pauli sum H = sum_p c_p P_p with c_p ~ N(0, 1) and each P_p acting as X, Y or Z on
    min_weight..max_weight distinct qubits (identity elsewhere), built directly in DIA.

    P|r> = phase(r) |r xor f> with f the mask of X/Y qubits and phase(r) = i^{#Y} (-1)^{popcount(r & (Z|Y mask))}
    (Y|0> = i|1>, Y|1> = -i|0>), so H[r xor f, r] += c phase(r): the entry sits on diagonal d = r - (r xor f),
    one of up to 2^{|f|} offsets. Returns (N, N, offsets ascending, vals, strings) with strings as (coef, ops dict).
    """
    N = 1 << n_qubits
    r = np.arange(N, dtype=np.int64)
    acc = {}
    strings = []
    for _ in range(n_strings):
        w = int(rng.integers(min_weight, max_weight, endpoint=True))
        qs = rng.choice(n_qubits, size=w, replace=False)
        ops = {int(q): "XYZ"[int(rng.integers(0, 3))] for q in qs}
        c = float(rng.standard_normal())
        strings.append((c, ops))
        f = sum(1 << q for q, o in ops.items() if o in "XY")
        zy = sum(1 << q for q, o in ops.items() if o in "YZ")
        n_y = sum(1 for o in ops.values() if o == "Y")
        par = np.bitwise_count(r & zy) & 1
        phase = (1j ** n_y) * (1.0 - 2.0 * par)
        rows = r ^ f
        dvec = r - rows
        order = np.argsort(dvec, kind="stable")
        ds = dvec[order]
        cuts = np.flatnonzero(np.diff(ds)) + 1
        for grp in np.split(order, cuts):
            dv = int(dvec[grp[0]])
            pos = rows[grp] if dv >= 0 else r[grp]
            if dv not in acc:
                acc[dv] = np.zeros(N - abs(dv), np.complex128)
            acc[dv][pos] += c * phase[grp]  # r -> r xor f is one-to-one: no repeated positions
    offs = sorted(d for d, a in acc.items() if np.any(a != 0))
    return N, N, offs, [acc[d] for d in offs], strings


def pauli_masks(strings):
    """(x_masks, z_masks, coefs) of Pauli strings given as (coef, {qubit: 'X' | 'Y' | 'Z'}) pairs (as returned by
    pauli_sum): x holds the qubits acted on by X or Y, z those acted on by Z or Y."""
    xs, zs, cs = [], [], []
    for c, ops in strings:
        xs.append(sum(1 << q for q, o in ops.items() if o in "XY"))
        zs.append(sum(1 << q for q, o in ops.items() if o in "YZ"))
        cs.append(complex(c))
    return xs, zs, cs


def tfim_terms(n_qubits: int, J: float = 1.0, h: float = 0.7):
    """The transverse-field Ising chain of tfim() as Pauli terms: -J Z_i Z_{i+1} (i < n-1) and -h X_i."""
    strings = [(-J, {i: "Z", i + 1: "Z"}) for i in range(n_qubits - 1)]
    strings += [(-h, {i: "X"}) for i in range(n_qubits)]
    return strings


def xxz_terms(n_qubits: int, Jxy: float = 1.0, Jz: float = 0.8, fields=None, seed: int = 0):
    """The XXZ chain of xxz() as Pauli terms: Jxy (X_i X_{i+1} + Y_i Y_{i+1}) + Jz Z_i Z_{i+1} and h_i Z_i, with the
    same seeded fields."""
    if fields is None:
        fields = np.random.default_rng(seed).uniform(-1.0, 1.0, n_qubits)
    strings = []
    for i in range(n_qubits - 1):
        strings += [(Jxy, {i: "X", i + 1: "X"}), (Jxy, {i: "Y", i + 1: "Y"}), (Jz, {i: "Z", i + 1: "Z"})]
    strings += [(float(fields[i]), {i: "Z"}) for i in range(n_qubits)]
    return strings
