/*
 * diaq_hip.h -- C ABI of the MI355X (gfx950) DiaQ GPU library, libdiaq_hip.so.
 *
 * Plain C: pointers, sizes and status codes; no C++ or torch types cross this boundary. Every entry point returns
 * DQ_OK (0) or an error code, and dq_last_error() gives the message of the calling thread's last failure (the texts
 * follow the reference's exception messages, so a binding can rethrow them unchanged).
 *
 * Each declaration names the reference interface it replaces (paths relative to the reference checkout). Pointers
 * named *_dev are device pointers the caller owns (borrowed for the call); streams travel as uintptr_t
 * (hipStream_t; 0 is the null stream). Entry points without a stream argument run on the null stream and return
 * after their results are complete.
 */
#ifndef DIAQ_HIP_H
#define DIAQ_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    DQ_OK = 0,
    DQ_ERR_ARG = 1,    /* bad argument: shape mismatch, bad size, null pointer ... */
    DQ_ERR_HIP = 2,    /* a HIP runtime call failed */
    DQ_ERR_ALLOC = 3,  /* device or host allocation failed */
    DQ_ERR_STATE = 4   /* anything else */
};

enum { DQ_F64 = 0, DQ_F32 = 1 }; /* value precision of a device matrix */

/* Opaque device-resident DIA matrix: deviceDiaQMatrix (src/deviceMatrix_gpu.hpp:37-68). */
typedef struct dq_devmat dq_devmat;

/* Message of this thread's last failed call ("" if none). No reference counterpart: the reference throws C++
 * exceptions, which cannot cross a C ABI. */
const char *dq_last_error(void);

/* Library version string. Replaces version() (include/pythonWrappers.cpp:15-17). */
const char *dq_version(void);

/* ------------------------------------------------------------------------------------------------------------------
 * Device matrices
 * ---------------------------------------------------------------------------------------------------------------- */

/* Upload a host DIA matrix: n_diags diagonals with offsets[k]; diagonal k's values start at element starts[k] of the
 * host arrays re / im (so the reference's SIMD-padded host layout and a packed one both work). precision DQ_F64
 * keeps the values; DQ_F32 rounds them to float. Replaces deviceDiaQMatrix(const shared_ptr<diaQMatrix>&)
 * (src/deviceMatrix_gpu.hpp:46, src/deviceMatrix_gpu.cu:62-76). */
int dq_devmat_upload(uint32_t rows, uint32_t cols, int32_t n_diags, const int32_t *offsets, const int64_t *starts,
                     const double *re, const double *im, int precision, dq_devmat **out);

/* A matrix with these diagonals, all values zero. Replaces deviceDiaQMatrix(shape, diagonals)
 * (src/deviceMatrix_gpu.hpp:48, src/deviceMatrix_gpu.cu:56-60). */
int dq_devmat_zeros(uint32_t rows, uint32_t cols, int32_t n_diags, const int32_t *offsets, int precision,
                    dq_devmat **out);

/* A Pauli-sum Hamiltonian H = sum_j c_j P_j on n_qubits qubits (N = 2^n_qubits), built on the device. Term j is
 * given by x_masks[j] (qubits acted on by X or Y; qubit k = bit k of the basis index), z_masks[j] (qubits acted on by
 * Z or Y) and the complex coefficient (coef_re[j], coef_im[j]); P|c> = i^{popcount(x & z)} (-1)^{popcount(c & z)}
 * |c xor x>. The diagonals are the structural set (every +-2^k combination of each term's flipped qubits,
 * ascending); dq_prune drops the ones whose terms cancel. No reference counterpart: the reference builds matrices on
 * the host from dense arrays or element by element (src/diaQMatrix.hpp:666-674, include/pythonWrappers.cpp:64). */
int dq_devmat_from_pauli_sum(int n_qubits, int n_terms, const uint64_t *x_masks, const uint64_t *z_masks,
                             const double *coef_re, const double *coef_im, int precision, uintptr_t stream,
                             dq_devmat **out);

/* from_numpy on the device: the nonzero diagonals of a dense rows x cols matrix that is already in device memory,
 * in ascending order, as a device matrix. dense is row-major complex128 (re, im interleaved), 16-byte aligned, with
 * row stride ld >= cols complex elements. A diagonal is kept when one of its elements has a nonzero real or
 * imaginary part (-0.0 counts as zero). Synchronises the stream once, to read the diagonal set. Replaces from_numpy
 * for matrices (src/conversions.cpp:103-181, src/conversions_gpu.cu:13-40, src/dense_to_diagonals.hpp:20-58), which
 * needs the dense array on the host. */
int dq_devmat_from_dense_device(uint32_t rows, uint32_t cols, const double *dense, int64_t ld, int precision,
                                uintptr_t stream, dq_devmat **out);

/* to_numpy on the device: write A as a dense row-major complex128 matrix (layout as above; every element of the
 * rows x cols block is written, zeros off A's diagonals). Replaces to_numpy / convertToDenseMatrix
 * (src/conversions.cpp:11-48, 70-101), which build the dense array on the host. Returns after the copy is done. */
int dq_devmat_to_dense_device(const dq_devmat *A, double *dense, int64_t ld, uintptr_t stream);

/* Fill all values from packed device arrays (diagonal k's len_k values back to back, in slot order) of the
 * matrix precision, on `stream`. No reference counterpart: the bulk constructor for device-generated matrices. */
int dq_devmat_set_packed_device(dq_devmat *A, const void *re_dev, const void *im_dev, uintptr_t stream);

/* Device-to-device copy. Replaces deviceDiaQMatrix::copy (src/deviceMatrix_gpu.hpp:53, .cu:87-95). */
int dq_devmat_copy(const dq_devmat *A, dq_devmat **out);

/* A copy in another precision (FP64 <-> FP32). No reference counterpart (precision form). */
int dq_devmat_convert(const dq_devmat *A, int precision, dq_devmat **out);

/* Download into host arrays laid out by starts[] (as dq_devmat_upload); values are widened to double, and the
 * elements between the first start and the last diagonal's end that belong to no diagonal (the layout's padding)
 * are written as zeros. Repacked on the device, then one copy per array. Replaces deviceDiaQMatrix::to_host
 * (src/deviceMatrix_gpu.hpp:55, .cu:97-113). */
int dq_devmat_download(const dq_devmat *A, const int64_t *starts, double *re, double *im);

/* Download one diagonal (by offset) into host arrays of its length, widened to double. No reference counterpart (the
 * reference downloads whole matrices only). */
int dq_devmat_get_diagonal(const dq_devmat *A, int32_t offset, double *re, double *im);

/* Free the matrix (NULL is ignored). Replaces ~deviceDiaQMatrix (src/deviceMatrix_gpu.cu:78-85). */
void dq_devmat_destroy(dq_devmat *A);

/* Shape, number of diagonals, precision, stored elements (sum of diagonal lengths). Replaces the matrix_shape /
 * num_diags members (src/deviceMatrix_gpu.hpp:39, :60). */
int dq_devmat_info(const dq_devmat *A, uint32_t *rows, uint32_t *cols, int32_t *n_diags, int *precision,
                   int64_t *stored);

/* The diagonal offsets in slot order (n_diags entries). Replaces the dIndices member (src/deviceMatrix_gpu.hpp:41). */
int dq_devmat_offsets(const dq_devmat *A, int32_t *out);

/* ------------------------------------------------------------------------------------------------------------------
 * SpMV
 * ---------------------------------------------------------------------------------------------------------------- */

/* y = A x on the device, y overwritten, asynchronous on `stream`. x / y hold cols / rows values of the matrix
 * precision (double for DQ_F64, float for DQ_F32) as split real and imaginary arrays.
 * Replaces spmv_device (src/deviceMatrix_gpu.hpp:71-72, src/spMV_gpu.cu:142-168). */
int dq_spmv_device(const dq_devmat *A, const void *xr_dev, const void *xi_dev, void *yr_dev, void *yi_dev,
                   uintptr_t stream);

/* y = A x for host vectors (copied in and out). Replaces spmv_host_vectors / spmv_cuda(DeviceMatrix, x)
 * (src/spMV_gpu.cu:171-183, :200-204). */
int dq_spmv_host(const dq_devmat *A, const double *xr, const double *xi, double *yr, double *yi);

/* y = A^n x for host vectors, by n SpMVs on the device. Replaces spMV_ntimes_cuda (src/spMV_ntimes_gpu.cu:61-263).
 * For n >= 2 the products skip A's all-zero 256-row blocks (a map made once per call; the same bits in FP64). */
int dq_spmv_ntimes_host(const dq_devmat *A, const double *xr, const double *xi, double *yr, double *yi, int n);

/* y = A^num_times x on device vectors of A's precision, asynchronous on `stream`: the reference's spMV_ntimes
 * (src/spMV_ntimes_gpu.cu:61-263, same errors and num_times == 0 / no-diagonal rules) on data that stays resident.
 * work_r / work_i are one more vector of A.rows elements, needed when num_times > 1 or x == y (in place: the result
 * overwrites x); the products alternate between y and the work vectors. For num_times >= 2 the products skip A's
 * all-zero 256-row blocks through a map made once per call in stream-ordered memory (hipMallocAsync), with the
 * schedule dq_spmv_tune chose for that; while `stream` is being captured into a graph no map is made. The FP64 result
 * has the same bits either way. */
int dq_spmv_ntimes_device(const dq_devmat *A, const void *xr, const void *xi, void *yr, void *yi, int num_times,
                          void *work_r, void *work_i, uintptr_t stream);

/* dq_spmv_ntimes_device without any device allocation: the same products, results and errors, but the zero-block map
 * a sequence of two or more products makes for an untuned matrix is written to a static 16 MiB device array of the
 * library (one per GPU, held until the process ends) instead of stream-ordered memory, and it is made even while
 * `stream` is being captured into a graph. A map that would not fit (more than 2^22 words: diagonals times
 * ceil(rows / 8192)) is an error; a matrix tuned by dq_spmv_tune uses the map kept in its handle and needs no scratch.
 * All calls on one GPU share the array: calls that overlap on one GPU (other streams, other threads) race, and a
 * captured graph that made its map there must not be replayed while another such call runs. No reference
 * counterpart: for callers that cannot allocate device memory (code under graph capture, serving loops). */
int dq_spmv_ntimes_device_noalloc(const dq_devmat *A, const void *xr, const void *xi, void *yr, void *yi,
                                  int num_times, void *work_r, void *work_i, uintptr_t stream);

/* y = A x with the column range split into n_bands passes (n_bands = 0 picks the library's choice for the size);
 * same result as dq_spmv_device up to summation order. No reference counterpart (tuning entry point). */
int dq_spmv_device_banded(const dq_devmat *A, const void *xr_dev, const void *xi_dev, void *yr_dev, void *yi_dev,
                          int n_bands, uintptr_t stream);

/* SpMV autotuner: times every (load schedule, rows per lane, non-temporal loads, LDS x window) candidate `repeats`
 * times on scratch vectors and keeps the fastest in the matrix handle, used by every later SpMV with A. When A has
 * all-zero 256-row blocks it also keeps their map in the handle and the fastest schedule that skips them: later FP64
 * SpMVs with A and spMV_ntimes skip those blocks (the same bits; dq_devmat_set_packed_device recomputes the map for
 * the new values, in place and in stream order, so graphs captured earlier stay correct). best_ms (may be NULL)
 * receives the time of the schedule without the skip. The candidates are passed to the kernels directly: SpMVs of
 * other matrices meanwhile keep their schedules, but calls with A itself must not overlap its tuning. No reference
 * counterpart. */
int dq_spmv_tune(dq_devmat *A, int repeats, uintptr_t stream, float *best_ms);

/* The SpMV schedule A currently uses (after tuning or the defaults); any output pointer may be NULL. No reference
 * counterpart. */
int dq_spmv_tuned_config(const dq_devmat *A, int *rows_per_lane, int *nontemporal, int *pipelined, int *lds_window,
                         int *xcd_tiles);

/* Return every cached device block to the runtime. The library keeps freed device blocks (up to a budget, by
 * default an eighth of the device memory; dq_set_tuning "alloc_cache_mb") to serve later allocations without
 * hipMalloc. Also drops the product dry-runs kept for reuse (the calling thread's at once, other threads' at their
 * next product). No reference counterpart. */
int dq_release_cached_memory(void);

/* Process-wide tuning knobs for the autotuner and benchmarks: "spmv_rows_per_lane" (0 = automatic, or 1, 2, 4, 8,
 * 16), "spmv_bands" (0 = automatic, or the number of column bands), "spmv_nontemporal" (-1 = automatic, 0, 1:
 * stream the matrix values with non-temporal loads), "spmv_pipelined" (-1 = automatic, 0 = per-row predicated
 * loads, 1 = batched loads), "spmv_lds" (-1 = automatic, 0, 1: stage the x window of each row tile in LDS for the
 * diagonals within half a tile) and "spmv_xcd" (-1 = automatic, 0 = row tiles round-robin over the XCDs, 1 = each
 * XCD a contiguous eighth of the tiles) override a matrix's tuned SpMV schedule; "product_interleave" (0, 1 = default: row
 * chunks of a product interleaved over the XCDs) and "product_nt_store" (0, 1 = default: non-temporal stores of the
 * product's values), "product_rows_per_lane" (0 = default: 4 below 2^16 rows, else 8; or 4, 8, 16: row chunks of 256 x that) and
 * "product_skip_zero_blocks" (1 = default: for operands of at least 2^16 rows, map their all-zero 256-row blocks and
 * drop the pairs whose rows in a chunk are all zero, same bits; 0 = read every pair) set the SpGEMM kernel's schedule; "evo_rows_per_lane" (0 = tuned, or 1, 2, 4, 8), "evo_xcd"
 * (-1 = tuned, 0, 1), "evo_lds" (-1 = tuned, 0, 1) and "evo_vec" (-1 = tuned, 0 = a lane per state, 1 = four states
 * per lane for 4, 8 or 16 states) and "evo_diag_batch" (0 = tuned, or 1, 2, 4: the diagonals whose loads the
 * four-states-per-lane kernel issues together at 1 or 2 rows per lane) override every batched-evolution class's
 * Horner schedule; "evo_lanes" (4 = default, 1 to 8, or 0 = a stream per class) sets over how many streams dq_evo_run
 * spreads the classes (longest first, by an estimate of each class's time),
 * and "evo_skip_zero_blocks" (1 = default: skip the blocks of a diagonal whose values are all zero; 0 = read them);
 * "spmv_skip_zero_blocks" (1 = default, 0) and "hamsim_skip_zero_blocks" (1 = default, 0) do the same for the
 * SpMV (the map dq_spmv_tune keeps, and spMV_ntimes's own) and for the fused evolution's Taylor terms, whose schedule
 * "hamsim_rows_per_lane" (0 = timed per matrix, or 2, 4, 8) and "hamsim_lds" (-1 = timed, 0, 1) override;
 * "spmm_tiled" (0 = not tiled, the default; 1, 2, 4 = LDS-tiled with that many elements per lane) and
 * "spmm_cols_per_lane" (0 = automatic: 2 for an even n_right, else 1; 1, 2, 4, 8; -1 = one C element per lane),
 * "spmm_rows_per_lane" (-1 = automatic: 2 for >= 32768 rows and an even n_right >= 64, else 0; 0 = off; 2, 4, 8: the
 * row-block form, whose lane keeps a window of B rows that a band reuses)
 * pick the spmm_dia_device kernel; "product_dense" (0 = never, 1 = default: a near-dense product inside pow or the matrix
 * exponential goes through a dense library GEMM); "dense_convert" (0 = automatic, 1 = per-row kernels, 2 = tiled
 * kernels) picks the kernels of dq_devmat_from_dense_device / dq_devmat_to_dense_device; "dense_convert_tile" (0 =
 * 32 x 64 tiles, the default, 1 = 64 x 32, 2 = 64 x 64, 3 = 128 x 32) their tiled kernels' tile; "gates_tile_qubits"
 * (10, 11 = default, 12) the tile of later gate plans, "gates_low_qubits" (4, 5 = default, 6) the low qubits that
 * every tile of theirs keeps (runs of 2^low contiguous amplitudes) and "gates_kernel" (0 = LDS tiles, the default;
 * 1 = amplitudes in registers with cross-lane shuffles) their pass kernel; "alloc_cache_mb" (-1 = an
 * eighth of the device memory, 0 = no cache) bounds the cache of freed device blocks. No reference counterpart. */
int dq_set_tuning(const char *key, int value);

/* ------------------------------------------------------------------------------------------------------------------
 * Gates on a state vector (Trotter steps) without forming the n-qubit matrix
 * ---------------------------------------------------------------------------------------------------------------- */

/* y (+)= (I_{2^before} (x) G (x) I_{2^after}) x on device vectors of G's precision (split re/im, length
 * 2^before * g * 2^after for a square g x g gate G). accumulate = 1 adds into y, as the reference does; 0 overwrites
 * it. Each output adds G's diagonals in slot order. x and y must not overlap, except that they may be the same
 * vectors (in place) when g <= 16 and G has at most 64 diagonals. Asynchronous on `stream`. Replaces
 * identity_kron_matrix_kron_identity_times_vector (src/spKron.cpp:433-496) and its cases before = 0
 * (matrix_kron_identity_times_vector, src/spKron.cpp:498-571) and after = 0 (identity_kron_matrix_times_vector,
 * src/spKron.cpp:286-335), which run on the host. */
int dq_kron_apply_device(const dq_devmat *G, uint32_t before, uint32_t after, const void *xr, const void *xi,
                         void *yr, void *yi, int accumulate, uintptr_t stream);

/* The same on host FP64 vectors (copied to the device and back; y += ...): the reference's host-vector form
 * (include/pythonWrappers.cpp:409-455). G must be FP64. */
int dq_kron_apply_host(const dq_devmat *G, uint32_t before, uint32_t after, const double *xr, const double *xi,
                       double *yr, double *yi);

/* A planned sequence of gates (a Trotter step): gate j (a 2^k x 2^k device matrix, 1 <= k <= 4) acts on qubits
 * first_qubit[j] .. first_qubit[j] + k - 1 of an n_qubits state (qubit q = bit q of the amplitude index), i.e. it is
 * I_{2^(n - q - k)} (x) G_j (x) I_{2^q}. The plan cuts the sequence into passes that each stream the state through
 * HBM once and apply a stretch of gates to tiles of 2^11 amplitudes held in LDS (dq_set_tuning
 * "gates_tile_qubits": 10, 11 or 12). Running it is bitwise equal to applying the gates in order with
 * dq_kron_apply_device(G_j, n - q - k, q, state, state, accumulate 0). All gates share one precision; the plan copies
 * their values (at most 512 nonzero entries per gate). n_qubits must be at least the tile qubits and at most 40.
 * No reference counterpart: the reference applies each gate in its own host loop (src/spKron.cpp:433-496), the step
 * that the diagonal-budgeted Trotterization workflow of the reference's README (README.md:340-352) builds on. */
typedef struct dq_gate_plan dq_gate_plan;
int dq_gate_plan_create(int n_qubits, int n_gates, const dq_devmat *const *gates, const uint32_t *first_qubit,
                        dq_gate_plan **out);
/* Number of passes over the state per run of the plan. */
int dq_gate_plan_passes(const dq_gate_plan *plan, int *n_passes);
/* state <- (G_{m-1} ... G_0)^repetitions state, in place on the split re/im device vectors (2^n_qubits amplitudes,
 * the gates' precision); asynchronous on `stream`. */
int dq_gate_plan_run(const dq_gate_plan *plan, void *re, void *im, int repetitions, uintptr_t stream);
void dq_gate_plan_destroy(dq_gate_plan *plan);

/* ------------------------------------------------------------------------------------------------------------------
 * SpGEMM, prune, norms
 * ---------------------------------------------------------------------------------------------------------------- */

/* Structural dry-run of C = A B: the ascending result offsets and, per result diagonal, its (slot in A, slot in B)
 * pairs in A-slot order. Call with res_offsets = NULL to get the counts, then with arrays of those sizes.
 * Replaces the host pairing in product_device (src/spGEMM_gpu.cu:90-125). */
int dq_product_dryrun(uint32_t rows, uint32_t cols, int32_t n_a, const int32_t *offs_a, int32_t n_b,
                      const int32_t *offs_b, int32_t *n_res, int64_t *n_pairs, int32_t *res_offsets,
                      int64_t *res_pair_start, int32_t *pair_slot_a, int32_t *pair_slot_b);

/* C = A B on the device, without pruning. Replaces product_device (src/deviceMatrix_gpu.hpp:78,
 * src/spGEMM_gpu.cu:84-151). */
int dq_product_device(const dq_devmat *A, const dq_devmat *B, dq_devmat **out);

/* Drop the diagonals whose values are all exactly zero. Replaces diaQMatrix::pruneMatrix on the product's result
 * (src/diaQMatrix.hpp:1543-1604, src/spGEMM.cpp:516-534). */
int dq_prune(const dq_devmat *A, dq_devmat **out);

/* max over columns of sum(|re| + |im|). Replaces diaQMatrix::estimate_1norm (src/diaQMatrix.hpp:1287-1358). */
int dq_estimate_1norm(const dq_devmat *A, double *out);

/* sqrt(sum re^2 + im^2). Replaces diaQMatrix::norm (src/diaQMatrix.hpp:1255-1284). */
int dq_frobenius_norm(const dq_devmat *A, double *out);

/* ||A - B||_F for two matrices with the same shape and diagonals (precisions may differ). No reference counterpart:
 * the error report of the precision form. */
int dq_frobenius_diff(const dq_devmat *A, const dq_devmat *B, double *out);

/* Sum of the main diagonal. Replaces trace (src/matrixKernels.cpp:9-26). */
int dq_trace(const dq_devmat *A, double *re, double *im);

/* np.tensordot(A, B, axes) for DIA-stored tensors, on the device. A holds a tensor of rank rank_a and shape shape_a
 * (rank_a = 0: the matrix itself, rank 2) in row-major order over its matrix positions; likewise B. The n_axes axes
 * axes_a[k] of A are contracted with axes_b[k] of B. The result's axes are A's free axes, then B's, each in their
 * order (out_shape, with room for rank_a + rank_b entries, and out_rank; a full contraction gives shape {1, 1}), as a
 * free_A x free_B matrix whose exactly-zero diagonals are pruned. A and B are permuted into free x contracted and
 * contracted x free matrices of their nonzeros by index-map kernels, then multiplied by dq_product_device. Replaces
 * tensordot (src/tensordotKernels.cpp:21-74) with its transpose / matrix_reshape / product chain; errors as the
 * reference's (axes of different lengths, an axis out of range, contracted sizes that differ). */
int dq_tensordot(const dq_devmat *A, int rank_a, const uint32_t *shape_a, const dq_devmat *B, int rank_b,
                 const uint32_t *shape_b, int n_axes, const uint32_t *axes_a, const uint32_t *axes_b, dq_devmat **out,
                 uint32_t *out_shape, int *out_rank);

/* ------------------------------------------------------------------------------------------------------------------
 * pow and SpME
 * ---------------------------------------------------------------------------------------------------------------- */

/* A^n (n >= 0). Replaces pow_cuda (src/spPow_gpu.cu:21-56). */
int dq_pow(const dq_devmat *A, int64_t n, dq_devmat **out);

/* exp(A) by Taylor series with scaling and squaring. Replaces expm_taylor_scaled_cuda (src/spME_gpu.cu:10-78). */
int dq_expm_taylor_scaled(const dq_devmat *A, int max_iterations, double scale_cutoff, double tolerance,
                          dq_devmat **out);

/* exp(-i t A) by Taylor series with scaling and squaring. Replaces expm_negative_taylor_scaled_cuda
 * (src/spME_gpu.cu:229-294). terms_out (may be NULL) receives the number of Taylor terms summed and s_out (may be
 * NULL) the number of squarings. */
int dq_expm_negative_taylor_scaled(const dq_devmat *A, double t, int max_iterations, double scale_cutoff,
                                   double tolerance, dq_devmat **out, int *terms_out, int *s_out);

/* ------------------------------------------------------------------------------------------------------------------
 * Fused time evolution
 * ---------------------------------------------------------------------------------------------------------------- */

/* psi(t) = prod_r exp(-i H t/r) psi0 with a fixed-m Taylor action per step (m = max_iterations); host vectors in and
 * out (FP64). H must be square. Replaces hamiltonian_simulation_cuda (src/spHamSim_gpu_all_fused.cu:287-376). An FP32
 * H is the precision form: Taylor terms in FP32 (compensated row sums), each step's output summed in FP64, and the
 * final state renormalised in FP64 to psi0's norm. */
int dq_hamsim_host(const dq_devmat *H, const double *psi_r, const double *psi_i, double t, uint64_t r,
                   int max_iterations, double *out_r, double *out_i);

/* The same evolution in place on device vectors of H's precision, asynchronous on `stream`. For an FP32 H the state
 * is widened to FP64 for the evolution (as above) and rounded to FP32 once at the end. No reference counterpart (the
 * resident form of the fused evolution). */
int dq_hamsim_device(const dq_devmat *H, void *psi_r_dev, void *psi_i_dev, double t, uint64_t r, int m,
                     uintptr_t stream);

/* The resident evolution with an FP64 state whatever H's precision: for an FP64 H the same as dq_hamsim_device; for
 * an FP32 H the precision form of dq_hamsim_host, in place on the device state. No reference counterpart (the
 * precision form's entry point). */
int dq_hamsim_device_f64state(const dq_devmat *H, double *psi_r_dev, double *psi_i_dev, double t, uint64_t r, int m,
                              uintptr_t stream);

/* ------------------------------------------------------------------------------------------------------------------
 * DIA x dense on device pointers
 * ---------------------------------------------------------------------------------------------------------------- */

/* C = A B, A row-aligned DIA (offsets int64[n_diags], vals[n_diags * n_rows]: vals[k*n_rows + r] = A[r, r+off_k]),
 * B n_cols x n_right and C n_rows x n_right row-major, double or interleaved complex double; asynchronous on
 * `stream`. Replaces spmm_dia_device (src/spMM_device_gpu.hpp:19-20, src/spMM_device_gpu.cu:116-127). */
int dq_spmm_dia_device(uintptr_t offsets, uintptr_t vals, int n_diags, int64_t n_rows, int64_t n_cols, uintptr_t B,
                       int64_t n_right, uintptr_t C, int is_complex, uintptr_t stream);

/* Storage and diagonal budgets for DIA x dense. Replaces dia_viability (src/dia_viability.hpp:42-63). */
typedef struct {
    int viable, storage_ok, diagonals_ok;
    double dia_bytes, csr_bytes, density;
    int64_t diagonal_budget;
    char reason[160];
} dq_dia_viability_t;

int dq_dia_viability(int64_t n_rows, int64_t n_cols, int64_t n_diags, int64_t nnz, int64_t n_right, int itemsize,
                     int index_bytes, dq_dia_viability_t *out);

/* The same two budgets with constants fitted on MI355X over the reference's 360-case grid (spmm_dia_device against
 * hipSPARSE's CSR x dense, complex double): DIA bytes <= 6 x CSR bytes, n_diags <= max(64, n_rows / 128), capped at 64
 * when n_right >= 16. No case passing them ran slower in DIA; they keep 228 of the 305 DIA wins (the reference's
 * constants keep 63). No reference counterpart (the reference's constants come from an A100). */
int dq_dia_viability_mi355x(int64_t n_rows, int64_t n_cols, int64_t n_diags, int64_t nnz, int64_t n_right,
                            int itemsize, int index_bytes, dq_dia_viability_t *out);

/* ------------------------------------------------------------------------------------------------------------------
 * Batched time evolution (no reference counterpart: the reference evolves one state of one Hamiltonian per call,
 * hamiltonian_simulation_cuda, src/spHamSim_gpu_all_fused.cu:287-376)
 * ---------------------------------------------------------------------------------------------------------------- */

/* Opaque plan: n_jobs square Hamiltonians (jobs of the same size and diagonal set share one class), n_states states
 * per job (1, 2, 4, 8 or 16), Hamiltonian values stored in value_precision (DQ_F32 or DQ_F64). States are FP64; the
 * Taylor terms run in FP32 with compensated row sums. The plan copies the matrices: they may be freed afterwards. */
typedef struct dq_evo dq_evo;
int dq_evo_create(int n_jobs, dq_devmat *const *H, int n_states, int value_precision, dq_evo **out);
void dq_evo_destroy(dq_evo *E);
int dq_evo_info(const dq_evo *E, int *n_classes, int *n_jobs, int *n_states);

/* States of one job, state-major host arrays: state s, row r at [s*N + r]. */
int dq_evo_set_states(dq_evo *E, int job, const double *re, const double *im);
int dq_evo_get_states(const dq_evo *E, int job, double *re, double *im);

/* Per job: evolve to t[j] by 2^s[j] sub-steps of a degree-m[j] Taylor polynomial (1 <= m <= 64, 0 <= s <= 24). */
int dq_evo_set_schedule(dq_evo *E, const double *t, const int32_t *m, const int32_t *s);

/* Rows per lane, row-tile order (0 = round-robin over the XCDs, 1 = each XCD a contiguous eighth of the rows) and
 * kernel form of every class by timing the candidates (after the schedule is set); chosen_rpt / chosen_xcd /
 * chosen_lds (may be NULL) receive one value per class. chosen_lds: 0 = a lane per state, 1 = a lane per state with
 * the LDS window (each tile stages its input rows +-T/2 in LDS), 2 = four states per lane (16-byte gathers; plans
 * of 4, 8 or 16 states), 3 and 4 = four states per lane with the loads of 2 or 4 diagonals issued together. Every
 * form gives the same bits. */
int dq_evo_tune(dq_evo *E, int *chosen_rpt, int *chosen_xcd, int *chosen_lds);

/* A-posteriori truncation estimate per job from its current (initial) states: 2^s ||t_{m+1}|| / (1 - ratio) of the
 * first sub-step's Taylor terms, the largest over the job's states. Leaves the states unchanged. */
int dq_evo_probe(dq_evo *E, double *est);

/* Evolve every job by its schedule; the classes run on evo_lanes streams (see dq_set_tuning), forked from and joined
 * to `stream`. use_graph != 0 captures the whole batch into a HIP graph on the first call and replays it;
 * renormalize != 0 ends with an FP64 renormalisation of every state. ms_out (may be NULL) receives the GPU time. The
 * graph is captured again only when what its launches depend on changed: a class's number of Horner levels (max over
 * its jobs of m 2^s), the tuned schedule, the evo_* tuning keys, the classes' assignment to streams or renormalize.
 * A new schedule that keeps those replays the same graph (the kernels read t, m and s from device memory). */
int dq_evo_run(dq_evo *E, int use_graph, int renormalize, uintptr_t stream, float *ms_out);

/* ------------------------------------------------------------------------------------------------------------------
 * The learned (m, s) selector's network: an FP32 MLP n_in -> h1 -> h2 -> n_out (tanh hidden layers, softmax
 * output) trained and run by hand-written HIP kernels. No reference counterpart.
 * Parameter layout (flat, FP32): W1 [h1 x n_in], b1 [h1], W2 [h2 x h1], b2 [h2], W3 [n_out x h2], b3 [n_out].
 * ---------------------------------------------------------------------------------------------------------------- */
typedef struct dq_mlp dq_mlp;
int dq_mlp_create(int n_in, int h1, int h2, int n_out, uint64_t seed, dq_mlp **out); /* Glorot-uniform init */
void dq_mlp_destroy(dq_mlp *M);
int64_t dq_mlp_num_params(const dq_mlp *M);
int dq_mlp_get_params(const dq_mlp *M, float *host);
int dq_mlp_set_params(dq_mlp *M, const float *host); /* also resets the optimiser state */
/* x_dev: B x n_in features on the device; logits_host (may be NULL) receives B x n_out logits. */
int dq_mlp_forward(dq_mlp *M, int B, const float *x_dev, float *logits_host);
/* Mean softmax cross-entropy of labels y_dev (B int32 classes) and its gradient (grad_host: n_params, may be NULL). */
int dq_mlp_loss_grad(dq_mlp *M, int B, const float *x_dev, const int32_t *y_dev, float *loss, float *grad_host);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) at learning rate lr; loss (may be NULL): the batch loss. */
int dq_mlp_train_step(dq_mlp *M, int B, const float *x_dev, const int32_t *y_dev, float lr, float *loss);
/* Argmax class per sample over the classes with mask_dev[o] != 0 (mask_dev may be NULL), into cls_dev (device). */
int dq_mlp_predict(dq_mlp *M, int B, const float *x_dev, const int32_t *mask_dev, int32_t *cls_dev, uintptr_t stream);

#ifdef __cplusplus
}
#endif

#endif /* DIAQ_HIP_H */
