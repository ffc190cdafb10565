"""CPU oracle for the DiaQ GPU hot path -- TEST INFRASTRUCTURE ONLY.

This module restates, in numpy, what the reference computes on its GPU path, so that the HIP implementation in
``diaq_amd`` can be checked against it. It is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import it.

Every function cites the reference file:line it follows (paths relative to the reference checkout).

Storage convention (``src/diaQMatrix.hpp:43-63``, ``getMatrixPosition`` :433-453): diagonal ``d`` of a
``rows x cols`` matrix holds ``A[p, p + d]`` at position ``p`` when ``d >= 0`` and ``A[p - d, p]`` at position ``p``
when ``d < 0``; its length is ``diagonal_length(d, rows, cols)``.

How each function is pinned (checked against something other than itself) is listed in ``PINS`` at the bottom and
exercised by ``tests/test_oracle.py``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

VEC_WIDTH = 8  # AVX-512 doubles per vector: the reference pads each diagonal start to it (src/diaQMatrix.hpp:313-327)


# ----------------------------------------------------------------------------------------------------------------------
# Container
# ----------------------------------------------------------------------------------------------------------------------
def diagonal_length(d: int, rows: int, cols: int) -> int:
    """Length of diagonal d (src/deviceMatrix_gpu.cu:16-28, src/diaQMatrix.hpp:390-409)."""
    if d < 0:
        n = rows + d
    elif d > 0:
        n = cols - d
    else:
        n = min(rows, cols)
    n = min(n, rows, cols)
    return max(n, 0)


def padded_offsets(offsets, rows: int, cols: int, vec_width: int = VEC_WIDTH):
    """Host storage starts and total length with each start rounded up to vec_width (src/diaQMatrix.hpp:290-357)."""
    starts, total = [], 0
    for d in offsets:
        if vec_width > 1 and total % vec_width:
            total += vec_width - total % vec_width
        starts.append(total)
        total += diagonal_length(d, rows, cols)
    return starts, total


def symmetric_order(n: int):
    """Default diagonal order 0, -1, +1, -2, +2, ... (src/diaQMatrix.hpp:25-39)."""
    out = []
    for i in range(n):
        if i == 0:
            out.append(0)
        elif i % 2 == 1:
            out.append(-((i + 1) // 2))
        else:
            out.append(i // 2)
    return out


@dataclass
class Dia:
    """A diagonal-format complex matrix: ``vals[k]`` is the storage of diagonal ``offsets[k]`` (slot order kept)."""

    rows: int
    cols: int
    offsets: list = field(default_factory=list)
    vals: list = field(default_factory=list)

    @staticmethod
    def zeros(rows, cols, offsets):
        offsets = [int(d) for d in offsets]
        return Dia(rows, cols, offsets,
                   [np.zeros(diagonal_length(d, rows, cols), np.complex128) for d in offsets])

    def copy(self):
        return Dia(self.rows, self.cols, list(self.offsets), [v.copy() for v in self.vals])

    @property
    def stored(self) -> int:
        return int(sum(len(v) for v in self.vals))


def position_to_rc(d: int, p):
    """Matrix (row, col) of position p on diagonal d (src/diaQMatrix.hpp:433-453)."""
    p = np.asarray(p, dtype=np.int64)
    if d >= 0:
        return p, p + d
    return p - d, p


def to_dense(M: Dia) -> np.ndarray:
    """Dense values of a DIA matrix (src/conversions.cpp:70-101)."""
    out = np.zeros((M.rows, M.cols), np.complex128)
    for d, v in zip(M.offsets, M.vals):
        r, c = position_to_rc(d, np.arange(len(v)))
        out[r, c] = v
    return out


def from_dense(A: np.ndarray) -> Dia:
    """Dense to DIA keeping the diagonals with a nonzero, ascending (src/dense_to_diagonals.hpp:20-58)."""
    A = np.asarray(A, dtype=np.complex128)
    rows, cols = A.shape
    offsets, vals = [], []
    for d in range(-(rows - 1), cols):
        n = diagonal_length(d, rows, cols)
        if n <= 0:
            continue
        r, c = position_to_rc(d, np.arange(n))
        v = A[r, c]
        if np.any(v != 0):
            offsets.append(d)
            vals.append(v.copy())
    return Dia(rows, cols, offsets, vals)


def identity(n: int) -> Dia:
    """identityMatrix(s): one main diagonal of ones (src/specialMatrices.hpp:9-24)."""
    return Dia(n, n, [0], [np.ones(n, np.complex128)])


# ----------------------------------------------------------------------------------------------------------------------
# SpMV
# ----------------------------------------------------------------------------------------------------------------------
def spmv(M: Dia, x: np.ndarray) -> np.ndarray:
    """y = A x, y written (not accumulated), diagonals summed in slot order per row.

    Row kernel: src/spMV_gpu.cu:74-136 (pos = row if d >= 0 else row + d, col = row + d, bounds-checked against the
    diagonal length and the column count); dispatcher src/spMV_gpu.cu:142-168 (zero y, then the one-diagonal kernel
    :17-68 or the row kernel). Both give the same values: the one-diagonal kernel touches the same (row, col) pairs.
    """
    x = np.asarray(x, dtype=np.complex128)
    if x.shape[0] != M.cols:
        raise RuntimeError("spmv_cuda: dimension mismatch (x length vs A.cols)")
    y = np.zeros(M.rows, np.complex128)
    for d, v in zip(M.offsets, M.vals):
        p = np.arange(len(v))
        r, c = position_to_rc(d, p)
        ok = (c >= 0) & (c < M.cols) & (r >= 0) & (r < M.rows)
        y[r[ok]] += v[ok] * x[c[ok]]
    return y


def spmv_ntimes(M: Dia, x: np.ndarray, n: int) -> np.ndarray:
    """y = A^n x by repeated SpMV (src/spMV_ntimes_gpu.cu:61-263): n = 0 copies x (square only), no diagonals gives
    zeros, n > 1 needs a square matrix."""
    if n < 0:
        raise RuntimeError("spmv_ntimes_cuda: num_times must be >= 0")
    if n == 0:
        if M.rows != M.cols:
            raise RuntimeError("spmv_ntimes_cuda: num_times==0 requires square matrix (M==N) to copy x -> y")
        return np.array(x, dtype=np.complex128, copy=True)
    if not M.offsets:
        return np.zeros(M.rows, np.complex128)
    if n > 1 and M.rows != M.cols:
        raise RuntimeError("spmv_ntimes_cuda: repeated SpMV requires square matrix (M==N) for num_times > 1")
    y = np.asarray(x, dtype=np.complex128)
    for _ in range(n):
        y = spmv(M, y)
    return y


# ----------------------------------------------------------------------------------------------------------------------
# SpGEMM
# ----------------------------------------------------------------------------------------------------------------------
def product_dryrun(offsA, offsB, rows: int, cols: int):
    """Structural dry-run of C = A B (src/spGEMM_gpu.cu:90-125).

    Every pair (slot a, slot b) contributes to dC = dA + dB when -(rows-1) <= dC <= cols-1; result offsets ascend;
    the pairs of one dC stay in A-slot order (then B-slot order). Returns (offsets, pairs) with pairs[k] the list of
    (slotA, slotB) of result diagonal k. Nothing is value-dependent.
    """
    lo, hi = -int(rows) + 1, int(cols) - 1
    by = {}
    for sa, da in enumerate(offsA):
        for sb, db in enumerate(offsB):
            dc = int(da) + int(db)
            if lo <= dc <= hi:
                by.setdefault(dc, []).append((sa, sb))
    offs = sorted(by)
    return offs, [by[d] for d in offs]


def product(A: Dia, B: Dia) -> Dia:
    """C = A B without pruning (product_device, src/spGEMM_gpu.cu:84-151; kernel :25-78).

    For result position t: i = t (dC >= 0) or t - dC, j = t + dC or t; pA = i (dA >= 0) or i + dA; pB = i + dA
    (dB >= 0) or j; pairs with pA or pB out of range are skipped; pairs are summed in A-slot order.
    """
    if A.cols != B.rows:
        raise RuntimeError("diaQMatrix product_cuda: dimension mismatch.")
    rows, cols = A.rows, B.cols
    offs, pairs = product_dryrun(A.offsets, B.offsets, rows, cols)
    C = Dia.zeros(rows, cols, offs)
    for k, dc in enumerate(offs):
        n = len(C.vals[k])
        t = np.arange(n, dtype=np.int64)
        i = t if dc >= 0 else t - dc
        j = t + dc if dc >= 0 else t
        acc = np.zeros(n, np.complex128)
        for sa, sb in pairs[k]:
            da, db = A.offsets[sa], B.offsets[sb]
            pa = i if da >= 0 else i + da
            pb = i + da if db >= 0 else j
            ok = (pa >= 0) & (pb >= 0) & (pa < len(A.vals[sa])) & (pb < len(B.vals[sb]))
            contrib = np.zeros(n, np.complex128)
            contrib[ok] = A.vals[sa][pa[ok]] * B.vals[sb][pb[ok]]
            acc += contrib
        C.vals[k] = acc
    return C


def prune(M: Dia) -> Dia:
    """Drop the diagonals whose values are all exactly zero, keeping slot order (src/diaQMatrix.hpp:1543-1604)."""
    keep = [k for k, v in enumerate(M.vals) if np.any((v.real != 0) | (v.imag != 0))]
    return Dia(M.rows, M.cols, [M.offsets[k] for k in keep], [M.vals[k].copy() for k in keep])


def product_host(A: Dia, B: Dia) -> Dia:
    """diaq.product on host matrices in the GPU build: product then prune (src/spGEMM.cpp:303-307, :516-534;
    AVOID_ZERO_DIAGS default ON, CMakeLists.txt:85)."""
    return prune(product(A, B))


# ----------------------------------------------------------------------------------------------------------------------
# Host helpers used between products in SpME
# ----------------------------------------------------------------------------------------------------------------------
def add_union(A: Dia, B: Dia, sign: float = 1.0) -> Dia:
    """C = A + sign*B over the ascending union of the diagonals (src/diaQMatrix.hpp:867-916)."""
    if A.rows != B.rows or A.cols != B.cols:
        raise RuntimeError("diaQMatrix::add: shape mismatch")
    offs = sorted(set(A.offsets) | set(B.offsets))
    sa = {d: k for k, d in enumerate(A.offsets)}
    sb = {d: k for k, d in enumerate(B.offsets)}
    vals = []
    for d in offs:
        if d in sa and d in sb:
            vals.append(A.vals[sa[d]] + sign * B.vals[sb[d]])
        elif d in sa:
            vals.append(A.vals[sa[d]].copy())
        else:
            vals.append(sign * B.vals[sb[d]])
    return Dia(A.rows, A.cols, offs, vals)


def frobenius(M: Dia) -> float:
    """sqrt(sum re^2 + im^2) over the stored values (src/diaQMatrix.hpp:1255-1284)."""
    s = 0.0
    for v in M.vals:
        s += float(np.sum(v.real * v.real + v.imag * v.imag))
    return math.sqrt(s)


def estimate_1norm(M: Dia) -> float:
    """max over columns of sum(|re| + |im|) -- |re|+|im|, not |z| (src/diaQMatrix.hpp:1287-1358)."""
    col = np.zeros(M.cols, np.float64)
    for d, v in zip(M.offsets, M.vals):
        p = np.arange(len(v))
        c = p + d if d >= 0 else p
        ok = (c >= 0) & (c < M.cols)
        np.add.at(col, c[ok], np.abs(v.real[ok]) + np.abs(v.imag[ok]))
    return float(col.max()) if M.cols else 0.0


def trace(M: Dia) -> complex:
    """Sum of the main diagonal (src/matrixKernels.cpp:9-26)."""
    for d, v in zip(M.offsets, M.vals):
        if d == 0:
            return complex(v.sum())
    return 0j


def scale_by_imag_copy(M: Dia, imag: float) -> Dia:
    """C = (i*imag) * A: (re, im) -> (-im*imag, re*imag) (src/diaQMatrix.hpp:1181-1211)."""
    if imag == 0.0:
        return Dia.zeros(M.rows, M.cols, M.offsets)
    return Dia(M.rows, M.cols, list(M.offsets),
               [(-v.imag * imag) + 1j * (v.real * imag) for v in M.vals])


def scaling_exponent(norm: float, t: float, cutoff: float):
    """s = ceil(log2((|t| ||A||_1 + 1e-12) / cutoff)) when that exceeds the cutoff, else 0 (src/spME_gpu.cu:251-261;
    positive form :32-42 with t = 1)."""
    scaled = abs(t) * norm + 1e-12
    return int(math.ceil(math.log2(scaled / cutoff))) if scaled > cutoff else 0


# ----------------------------------------------------------------------------------------------------------------------
# pow and SpME
# ----------------------------------------------------------------------------------------------------------------------
def pow_gpu(A: Dia, n: int) -> Dia:
    """A^n as pow_cuda (src/spPow_gpu.cu:21-56): n = 0 gives I, n = 1 a copy, 2^s by s squarings, otherwise binary
    exponentiation starting from the first odd power (not from I)."""
    if n < 0:
        raise RuntimeError("pow_cuda: negative exponent not supported")
    if n == 0:
        return identity(A.rows)
    if n == 1:
        return A.copy()
    base = A.copy()
    if n & (n - 1) == 0:
        for _ in range(n.bit_length() - 1):
            base = product(base, base)
        return base
    result = None
    e = n
    while e > 0:
        if e & 1:
            result = product(result, base) if result is not None else base.copy()
        e >>= 1
        if e:
            base = product(base, base)
    return result


def _single_diag_exp(A: Dia, t: float, negative: bool) -> Dia:
    """One-diagonal fast path (src/spME_gpu.cu:99-124 and :170-197): the identity with its first len(A's diagonal)
    main-diagonal entries replaced by exp(a + ib) (expm) or e^{tb}(cos ta - i sin ta) (expm_negative). The lone
    diagonal is written onto the main diagonal whatever its offset (a reference quirk kept here)."""
    R = identity(A.rows)
    v = A.vals[0]
    a, b = v.real, v.imag
    if negative:
        e = np.exp(t * b)
        R.vals[0][: len(v)] = e * np.cos(t * a) - 1j * (e * np.sin(t * a))
    else:
        e = np.exp(a)
        R.vals[0][: len(v)] = e * np.cos(b) + 1j * (e * np.sin(b))
    return R


def taylor_loop(X: Dia, n: int, max_iterations: int, tolerance: float):
    """result = I + sum_i X^i / i!, stopping after the first term whose Frobenius norm is below tolerance (checked
    after adding it), at most max_iterations terms; products are not pruned (src/spME_gpu.cu:126-135 and :199-209).
    Returns (result, number of terms added)."""
    result = identity(n)
    term = identity(n)
    i = 0
    for i in range(1, max_iterations + 1):
        term = product(term, X)
        term = Dia(term.rows, term.cols, term.offsets, [v / i for v in term.vals])
        result = add_union(result, term)
        if frobenius(term) < tolerance:
            break
    return result, i


def expm_negative(A: Dia, t: float, max_iterations: int = 100, scale_cutoff: float = 1.0,
                  tolerance: float = 1e-12) -> Dia:
    """exp(-i t A) by Taylor with scaling and squaring (expm_negative_taylor_scaled_cuda, src/spME_gpu.cu:229-294,
    Taylor part :156-222, squaring via pow_cuda(U, 1 << s))."""
    if len(A.offsets) == 1:
        return _single_diag_exp(A, t, negative=True)
    s = scaling_exponent(estimate_1norm(A), t, scale_cutoff)
    ts = math.ldexp(t, -s)
    tol = max(1e-15, tolerance * math.ldexp(1.0, -s))
    X = scale_by_imag_copy(A, -1.0 * ts)
    U, _ = taylor_loop(X, A.rows, max_iterations, tol)
    return pow_gpu(U, 1 << s)


def expm(A: Dia, max_iterations: int = 100, scale_cutoff: float = 1.0, tolerance: float = 1e-12) -> Dia:
    """exp(A) by Taylor with scaling and squaring (expm_taylor_scaled_cuda, src/spME_gpu.cu:10-78, Taylor part
    :86-148): A is scaled by 2^-s, the series of A/2^s is summed, then squared s times."""
    if len(A.offsets) == 1:
        return _single_diag_exp(A, 1.0, negative=False)
    s = scaling_exponent(estimate_1norm(A), 1.0, scale_cutoff)
    inv = math.ldexp(1.0, -s)
    As = Dia(A.rows, A.cols, list(A.offsets), [v * inv for v in A.vals])
    tol = max(1e-15, tolerance * inv)
    U, _ = taylor_loop(As, A.rows, max_iterations, tol)
    return pow_gpu(U, 1 << s)


def _offsets_union(a, b):
    return sorted(set(a) | set(b))


def expm_negative_dense(H: np.ndarray, offsets, t: float, max_iterations: int = 100, scale_cutoff: float = 1.0,
                        tolerance: float = 1e-12):
    """expm_negative with dense values: the same steps as ``expm_negative`` (src/spME_gpu.cu:229-294, :156-222,
    src/spPow_gpu.cu:21-56) with the stored-diagonal sets carried alongside through the structural dry-run, for sizes
    where DIA products in numpy are too slow. ``H`` is dense, ``offsets`` its stored diagonals (slot order).
    Returns (U dense, result offsets, terms summed, s). Matrices with more than one diagonal only."""
    n = H.shape[0]
    D = np.abs(H.real) + np.abs(H.imag)
    s = scaling_exponent(float(D.sum(axis=0).max()), t, scale_cutoff)
    ts = math.ldexp(t, -s)
    tol = max(1e-15, tolerance * math.ldexp(1.0, -s))
    X = 1j * (-1.0 * ts) * H
    result = np.eye(n, dtype=np.complex128)
    term = np.eye(n, dtype=np.complex128)
    r_offs, t_offs = [0], [0]
    i = 0
    for i in range(1, max_iterations + 1):
        term = (term @ X) / i
        t_offs, _ = product_dryrun(t_offs, offsets, n, n)
        result = result + term
        r_offs = _offsets_union(r_offs, t_offs)
        if np.linalg.norm(term) < tol:
            break
    for _ in range(s):
        result = result @ result
        r_offs, _ = product_dryrun(r_offs, r_offs, n, n)
    return result, r_offs, min(i, max_iterations), s


# ----------------------------------------------------------------------------------------------------------------------
# Fused time evolution
# ----------------------------------------------------------------------------------------------------------------------
def taylor_action_step(H: Dia, psi: np.ndarray, dt: float, m: int) -> np.ndarray:
    """out = sum_{k=0}^{m} (-i dt H)^k psi / k! with the fixed m (src/spHamSim_gpu_all_fused.cu:235-277): tmp = H term;
    term = (-i dt/k) tmp, i.e. (b*s, -a*s) with s = dt/k; out += term."""
    out = np.array(psi, dtype=np.complex128, copy=True)
    term = out.copy()
    for k in range(1, m + 1):
        tmp = spmv(H, term)
        s = dt / k
        term = (tmp.imag * s) + 1j * (-tmp.real * s)
        out += term
    return out


def hamiltonian_simulation(H: Dia, psi0: np.ndarray, t: float, r: int, max_iterations: int = 100) -> np.ndarray:
    """psi(t) with r uniform steps of dt = t / r, each a fixed-m Taylor action; r = 0 returns psi0; H must be square
    (src/spHamSim_gpu_all_fused.cu:287-376). The Python keyword ``dt`` of the binding is this total t."""
    if H.rows != H.cols:
        raise RuntimeError("hamiltonian_simulation_cuda: H must be square")
    psi = np.asarray(psi0, dtype=np.complex128)
    if psi.shape[0] != H.cols:
        raise RuntimeError("hamiltonian_simulation_cuda: psi0 dimension mismatch")
    if r == 0:
        return psi.copy()
    dt = t / r
    for _ in range(r):
        psi = taylor_action_step(H, psi, dt, max_iterations)
    return psi


def taylor_action_ms(H: np.ndarray, psi: np.ndarray, t: float, m: int, s: int) -> np.ndarray:
    """[T_m(Y)]^(2^s) psi with Y = -i (t / 2^s) H and T_m the degree-m Taylor polynomial of exp: the approximation
    a scheduled (m, s) evolution computes (the per-step action of src/spHamSim_gpu_all_fused.cu:235-277 with
    2^s steps of t / 2^s and degree m). Dense FP64 (H may be a dense or scipy sparse matrix)."""
    tau = math.ldexp(t, -s)
    out = np.array(psi, dtype=np.complex128, copy=True)
    for _ in range(1 << s):
        term = out.copy()
        acc = out.copy()
        for k in range(1, m + 1):
            term = (-1j * tau / k) * (H @ term)
            acc = acc + term
        out = acc
    return out


# ----------------------------------------------------------------------------------------------------------------------
# DIA x dense and the viability rule
# ----------------------------------------------------------------------------------------------------------------------
def spmm_dia(offsets, vals, n_rows: int, n_cols: int, B: np.ndarray) -> np.ndarray:
    """C = A B for row-aligned DIA A (vals[k, r] = A[r, r + offsets[k]]) and dense row-major B, C overwritten
    (src/spMM_device_gpu.cu:44-79): entries with r + offset outside [0, n_cols) are skipped."""
    vals = np.asarray(vals).reshape(len(offsets), n_rows)
    B = np.asarray(B)
    C = np.zeros((n_rows, B.shape[1]), dtype=np.result_type(vals, B))
    r = np.arange(n_rows)
    for k, o in enumerate(offsets):
        c = r + int(o)
        ok = (c >= 0) & (c < n_cols)
        C[ok] += vals[k, ok][:, None] * B[c[ok]]
    return C


# ----------------------------------------------------------------------------------------------------------------------
# Gates on a state vector without forming the n-qubit matrix
# ----------------------------------------------------------------------------------------------------------------------
def kron_apply(before: int, G: Dia, after: int, x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """y + (I_{2^before} (x) G (x) I_{2^after}) x, following identity_kron_matrix_kron_identity_times_vector
    (src/spKron.cpp:433-496): for every block rep < 2^before and column j < 2^after, each diagonal slot in order adds
    G's values into y at the reference's offsets (skip_places = rep * 2^after * (len + |d|); x_offset and y_offset
    shifted by 2^after * |d| on the side the diagonal leaves). G is square. Returns the new y (y is not modified)."""
    if G.rows != G.cols:
        raise ValueError("gate must be square")
    B, A = 1 << before, 1 << after
    x = np.asarray(x, np.complex128)
    y = np.array(y, np.complex128)
    for rep in range(B):
        for a, d in enumerate(G.offsets):
            ln = diagonal_length(d, G.rows, G.cols)
            skip = rep * (A * (ln + abs(d)))
            x_off = skip if d < 0 else A * d + skip
            y_off = skip - A * d if d < 0 else skip
            i = np.arange(ln)[:, None] * A + np.arange(A)[None, :]  # (i, j)
            y[y_off + i] += G.vals[a][:, None] * x[x_off + i]
    return y


def identity_kron_matrix_times_vector(n: int, G: Dia, x, y):
    """y + (I_{2^n} (x) G) x (src/spKron.cpp:286-335): the case after = 0."""
    return kron_apply(n, G, 0, x, y)


def matrix_kron_identity_times_vector(G: Dia, n: int, x, y):
    """y + (G (x) I_{2^n}) x (src/spKron.cpp:498-571; zero values are skipped there, which adds nothing): the case
    before = 0."""
    return kron_apply(0, G, n, x, y)


VIABILITY_CONSTANTS = dict(max_bytes_over_csr=0.9, min_budget=16, rows_per_diagonal=32, wide_right=32, wide_budget=8)


def dia_viability(n_rows, n_cols, n_diags, nnz, n_right, itemsize=16, index_bytes=8, constants=None):
    """Storage and diagonal budgets for DIA x dense (src/dia_viability.hpp:42-63). `constants` replaces the reference's
    five (src/dia_viability.hpp:36-40, VIABILITY_CONSTANTS) in the same rule."""
    c = VIABILITY_CONSTANTS if constants is None else constants
    stored = float(n_diags) * float(n_rows)
    dia_bytes = stored * itemsize + float(n_diags) * 8.0
    csr_bytes = float(nnz) * (itemsize + index_bytes) + float(n_rows + 1) * index_bytes
    density = float(nnz) / stored if stored > 0 else 0.0
    storage_ok = n_diags > 0 and dia_bytes <= c["max_bytes_over_csr"] * csr_bytes
    budget = max(c["min_budget"], int(n_rows) // c["rows_per_diagonal"])
    if n_right >= c["wide_right"]:
        budget = min(budget, c["wide_budget"])
    diagonals_ok = n_diags <= budget
    viable = storage_ok and diagonals_ok
    if not storage_ok:
        reason = "storage: DIA would take more than 90% of CSR's bytes (diagonals too sparsely filled)"
    elif not diagonals_ok:
        reason = f"diagonals: {n_diags} over the budget of {budget}"
    else:
        reason = "viable"
    return dict(viable=viable, storage_ok=storage_ok, diagonals_ok=diagonals_ok, dia_bytes=dia_bytes,
                csr_bytes=csr_bytes, density=density, diagonal_budget=budget, reason=reason)


# ----------------------------------------------------------------------------------------------------------------------
# tensordot
# ----------------------------------------------------------------------------------------------------------------------
def tensordot(A: np.ndarray, B: np.ndarray, axes) -> Dia:
    """np.tensordot(A, B, axes) as the reference computes it (src/tensordotKernels.cpp:21-74): A's axes permuted to
    (free, contracted in the order given) and read as a free_A x sum_A matrix of its nonzeros, B's to (contracted,
    free) as sum_B x free_B (transpose, src/transposeKernels.cpp:19-110; matrix_reshape, src/reshapeKernels.cpp:21-73),
    then product() with its zero-diagonal prune (src/spGEMM.cpp:516-534). A and B are dense tensors; returns the
    free_A x free_B DIA result (its tensor shape is A's free axes then B's)."""
    ax_a, ax_b = [list(a) for a in axes]
    if len(ax_a) != len(ax_b):
        raise ValueError("tensordot: axes must be two lists of the same length")
    free_a = [i for i in range(A.ndim) if i not in ax_a]
    free_b = [i for i in range(B.ndim) if i not in ax_b]
    fa = int(np.prod([A.shape[i] for i in free_a])) if free_a else 1
    fb = int(np.prod([B.shape[i] for i in free_b])) if free_b else 1
    sa = int(np.prod([A.shape[i] for i in ax_a])) if ax_a else 1
    Am = np.transpose(A, free_a + ax_a).reshape(fa, sa)
    Bm = np.transpose(B, ax_b + free_b).reshape(sa, fb)
    return prune(product(from_dense(Am), from_dense(Bm)))


# ----------------------------------------------------------------------------------------------------------------------
# The (m, s) selector's network (no reference counterpart): numpy twin of the HIP kernels, for gradient checks
# ----------------------------------------------------------------------------------------------------------------------
def mlp_split(params, dims):
    """Views W1, b1, W2, b2, W3, b3 of a flat parameter vector laid out as the C ABI states."""
    n_in, h1, h2, n_out = dims
    sizes = [h1 * n_in, h1, h2 * h1, h2, n_out * h2, n_out]
    out, o = [], 0
    for sz in sizes:
        out.append(params[o:o + sz])
        o += sz
    W1, b1, W2, b2, W3, b3 = out
    return W1.reshape(h1, n_in), b1, W2.reshape(h2, h1), b2, W3.reshape(n_out, h2), b3


def mlp_forward(params, dims, x):
    """logits and hidden activations: a1 = tanh(W1 x + b1), a2 = tanh(W2 a1 + b2), z = W3 a2 + b3 (float64)."""
    W1, b1, W2, b2, W3, b3 = mlp_split(np.asarray(params, np.float64), dims)
    x = np.asarray(x, np.float64)
    a1 = np.tanh(x @ W1.T + b1)
    a2 = np.tanh(a1 @ W2.T + b2)
    return a2 @ W3.T + b3, a1, a2


def mlp_loss_grad(params, dims, x, y):
    """Mean softmax cross-entropy over the batch and its gradient, flat in the parameter layout."""
    x = np.asarray(x, np.float64)
    z, a1, a2 = mlp_forward(params, dims, x)
    B = x.shape[0]
    zm = z - z.max(axis=1, keepdims=True)
    p = np.exp(zm) / np.exp(zm).sum(axis=1, keepdims=True)
    loss = float(np.mean(-np.log(p[np.arange(B), y])))
    d3 = p.copy()
    d3[np.arange(B), y] -= 1.0
    d3 /= B
    W1, b1, W2, b2, W3, b3 = mlp_split(np.asarray(params, np.float64), dims)
    gW3, gb3 = d3.T @ a2, d3.sum(0)
    d2 = (d3 @ W3) * (1 - a2 ** 2)
    gW2, gb2 = d2.T @ a1, d2.sum(0)
    d1 = (d2 @ W2) * (1 - a1 ** 2)
    gW1, gb1 = d1.T @ x, d1.sum(0)
    return loss, np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])


def adam_step(params, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """Adam with bias correction (Kingma & Ba 2015), in place on numpy arrays."""
    m[:] = b1 * m + (1 - b1) * grad
    v[:] = b2 * v + (1 - b2) * grad * grad
    params -= lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)


PINS = {
    "diagonal_length": "brute force over dense masks (tests/test_oracle.py::test_diagonal_length_brute)",
    "padded_offsets": "closed form: starts are multiples of 8 and cover every diagonal",
    "from_dense/to_dense": "round trip on random dense matrices; ascending nonzero offsets",
    "spmv": "dense A @ x; spmv_2 known answer (cpp_tests/test_main.cpp:1782-1794); identity gives x",
    "product_dryrun": "nonzero diagonals of dense A @ B before pruning; TFIM-20 H.H closed form 727 offsets",
    "product": "51 SpGEMM known-answer pairs (cpp_tests/test_main.cpp), dense A @ B",
    "prune": "constructed cancellation (I + E01)(I - E01) = I",
    "add_union/frobenius/trace": "dense A + B, numpy norm, numpy trace",
    "estimate_1norm": "dense max column sum of |re| + |im|; TFIM-12 closed form 19.4",
    "pow_gpu": "np.linalg.matrix_power",
    "expm/expm_negative": "scipy.linalg.expm at <= 12 qubits; TFIM-12 s = 1",
    "taylor_action/hamsim": "scipy.linalg.expm_multiply / dense expm; 2-norm preservation",
    "spmm_dia": "dense matmul",
    "kron_apply": "np.kron(np.kron(I, G), I) @ x on random gates (the reference's test_api_vs_numpy.py check); "
                  "the spKron.cpp:2231-2262 prime-gate case against the kron matrix",
    "dia_viability": "hand-computed budget cases",
    "mlp_loss_grad": "central finite differences of mlp_forward's loss",
    "tensordot": "np.tensordot values; offsets = the nonzero diagonals of the dense result "
                 "(tests/test_oracle.py::test_tensordot_vs_numpy)",
    "adam_step": "closed forms worked by hand: step 1 from zero moments, step 2 with a zero gradient, a constant "
                 "gradient (tests/test_oracle.py::test_adam_step_closed_forms)",
}
