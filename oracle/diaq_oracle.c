/*
 * CPU oracle for the DiaQ GPU hot path, plain C + OpenMP -- TEST INFRASTRUCTURE ONLY.
 *
 * This is the checker and the CPU baseline instrument, never the product: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it (through ctypes). Paths in citations are relative to the reference checkout.
 *
 * Matrix layout taken by every function: rows x cols, n_diags diagonals with offsets[k] (int32) whose storage starts at
 * element starts[k] (int64) of the split real/imaginary arrays re/im; position p of diagonal d holds A[p, p+d] for
 * d >= 0 and A[p-d, p] for d < 0 (src/diaQMatrix.hpp:433-453).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#include <omp.h>

static int64_t diag_len(int d, int64_t rows, int64_t cols) {
    /* src/diaQMatrix.hpp:390-409 */
    int64_t n = d < 0 ? rows + d : (d > 0 ? cols - d : (rows < cols ? rows : cols));
    if (n > rows) n = rows;
    if (n > cols) n = cols;
    return n < 0 ? 0 : n;
}

/* ---------------------------------------------------------------------------------------------------------------
 * SpMV with the GPU row-kernel semantics (src/spMV_gpu.cu:74-136): for every row, the diagonals are summed in slot
 * order, each product a_r*x_r - a_i*x_i / a_r*x_i + a_i*x_r, and y is written once. Rows are independent, so the
 * OpenMP split over rows does not change any sum.
 * ------------------------------------------------------------------------------------------------------------- */
void oracle_spmv_rows(int64_t rows, int64_t cols, int n_diags, const int32_t *offsets, const int64_t *starts,
                      const double *re, const double *im, const double *xr, const double *xi, double *yr,
                      double *yi) {
#pragma omp parallel for schedule(static)
    for (int64_t row = 0; row < rows; ++row) {
        double ar = 0.0, ai = 0.0;
        for (int k = 0; k < n_diags; ++k) {
            const int d = offsets[k];
            const int64_t len = diag_len(d, rows, cols);
            const int64_t pos = d >= 0 ? row : row + d;
            const int64_t col = row + d;
            if (pos < 0 || pos >= len || col < 0 || col >= cols) continue;
            const double a_r = re[starts[k] + pos], a_i = im[starts[k] + pos];
            const double x_r = xr[col], x_i = xi[col];
            ar += a_r * x_r - a_i * x_i;
            ai += a_r * x_i + a_i * x_r;
        }
        yr[row] = ar;
        yi[row] = ai;
    }
}

/* ---------------------------------------------------------------------------------------------------------------
 * The reference's CPU SpMV, restated (src/spMV.cpp:46-320 with the AVX-512 tile kernels
 * src/simd_avx512_kernels.cpp:241-441):
 *   - tasks are (diagonal, tile of block_size positions) pairs, built diagonal by diagonal (:72-103);
 *   - every thread owns a zeroed copy of y, allocated per call (:150-157), and runs its static share of the tasks,
 *     y_t[y_idx] += A[i] * x[x_idx] with x_idx = i, y_idx = i - d for d < 0 and x_idx = i + d, y_idx = i for d > 0;
 *   - y is the sum of the thread copies (:291-307).
 * The tile loops are written as plain loops; target_clones lets gcc vectorise them for AVX-512 where the host has it
 * (AVX2 otherwise), the vector widths of the reference's own kernels.
 * block_size <= 0 selects the reference rule floor(floor(L1d * 0.95 / 48) / 8) * 8 (src/config.hpp:109-116).
 * ------------------------------------------------------------------------------------------------------------- */
int oracle_block_size(void) {
    long l1 = sysconf(_SC_LEVEL1_DCACHE_SIZE);
    if (l1 <= 0) l1 = 32768;
    int raw = (int)((double)l1 * 0.95 / (6.0 * 8.0));
    int b = (raw / 8) * 8;
    return b > 0 ? b : 8;
}

__attribute__((target_clones("avx512f", "avx2", "default")))
static void tile_kernel(int64_t start, int64_t end, int64_t xoff, int64_t yoff, const double *restrict ar,
                        const double *restrict ai, const double *restrict xr, const double *restrict xi,
                        double *restrict yr, double *restrict yi) {
    for (int64_t i = start; i < end; ++i) {
        const double a_r = ar[i], a_i = ai[i], x_r = xr[i + xoff], x_i = xi[i + xoff];
        yr[i + yoff] += a_r * x_r - a_i * x_i;
        yi[i + yoff] += a_r * x_i + a_i * x_r;
    }
}

__attribute__((target_clones("avx512f", "avx2", "default")))
static void merge_kernel(int64_t lo, int64_t hi, int nt, double *const *yl_r, double *const *yl_i, double *yr,
                         double *yi) {
    for (int64_t i = lo; i < hi; ++i) {
        double sr = 0.0, si = 0.0;
        for (int t = 0; t < nt; ++t) {
            sr += yl_r[t][i];
            si += yl_i[t][i];
        }
        yr[i] = sr;
        yi[i] = si;
    }
}

int oracle_spmv_tasks(int64_t rows, int64_t cols, int n_diags, const int32_t *offsets, const int64_t *starts,
                      const double *re, const double *im, const double *xr, const double *xi, double *yr,
                      double *yi, int block_size, int n_threads) {
    if (block_size <= 0) block_size = oracle_block_size();
    if (n_threads <= 0) n_threads = omp_get_max_threads();
    if (n_diags == 0 || rows == 0) return 0;
    /* task list, diagonal by diagonal */
    int64_t n_tasks = 0;
    for (int k = 0; k < n_diags; ++k) {
        const int64_t len = diag_len(offsets[k], rows, cols);
        n_tasks += (len + block_size - 1) / block_size;
    }
    int32_t *t_diag = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n_tasks ? n_tasks : 1));
    int64_t *t_lo = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n_tasks ? n_tasks : 1));
    if (!t_diag || !t_lo) { free(t_diag); free(t_lo); return -1; }
    int64_t t = 0;
    for (int k = 0; k < n_diags; ++k) {
        const int64_t len = diag_len(offsets[k], rows, cols);
        for (int64_t s = 0; s < len; s += block_size) { t_diag[t] = k; t_lo[t] = s; ++t; }
    }
    /* thread-local zeroed copies of y */
    double **yl_r = (double **)calloc((size_t)n_threads, sizeof(double *));
    double **yl_i = (double **)calloc((size_t)n_threads, sizeof(double *));
    int fail = 0;
    for (int th = 0; th < n_threads; ++th) {
        yl_r[th] = (double *)aligned_alloc(64, ((size_t)rows * 8 + 63) / 64 * 64);
        yl_i[th] = (double *)aligned_alloc(64, ((size_t)rows * 8 + 63) / 64 * 64);
        if (!yl_r[th] || !yl_i[th]) fail = 1;
    }
    if (!fail) {
#pragma omp parallel num_threads(n_threads)
        {
            const int th = omp_get_thread_num();
            double *lr = yl_r[th], *li = yl_i[th];
            memset(lr, 0, (size_t)rows * sizeof(double));  /* each thread zeroes its own copy */
            memset(li, 0, (size_t)rows * sizeof(double));
#pragma omp barrier
#pragma omp for schedule(static)
            for (int64_t ti = 0; ti < n_tasks; ++ti) {
                const int k = t_diag[ti];
                const int d = offsets[k];
                const int64_t len = diag_len(d, rows, cols);
                const int64_t lo = t_lo[ti];
                const int64_t hi = lo + block_size < len ? lo + block_size : len;
                const int64_t xoff = d > 0 ? d : 0;
                const int64_t yoff = d < 0 ? -(int64_t)d : 0;
                tile_kernel(lo, hi, xoff, yoff, re + starts[k], im + starts[k], xr, xi, lr, li);
            }
#pragma omp barrier
#pragma omp for schedule(static)
            for (int64_t c = 0; c < (rows + 4095) / 4096; ++c) {
                const int64_t lo = c * 4096, hi = lo + 4096 < rows ? lo + 4096 : rows;
                merge_kernel(lo, hi, n_threads, yl_r, yl_i, yr, yi);
            }
        }
    }
    for (int th = 0; th < n_threads; ++th) { free(yl_r[th]); free(yl_i[th]); }
    free(yl_r);
    free(yl_i);
    free(t_diag);
    free(t_lo);
    return fail ? -1 : 0;
}

/* ---------------------------------------------------------------------------------------------------------------
 * SpGEMM values for one result diagonal from a pair list, with the GPU kernel's index rules and pair order
 * (src/spGEMM_gpu.cu:25-78). The pair list (slot_a, slot_b) is the structural dry-run (:90-125), built by the caller.
 * ------------------------------------------------------------------------------------------------------------- */
void oracle_product_diag(int64_t rows_a, int64_t cols_a, int64_t cols_b, int n_a, const int32_t *off_a,
                         const int64_t *st_a, const double *a_re, const double *a_im, int n_b, const int32_t *off_b,
                         const int64_t *st_b, const double *b_re, const double *b_im, int dc, int n_pairs,
                         const int32_t *slot_a, const int32_t *slot_b, double *c_re, double *c_im) {
    (void)n_a;
    (void)n_b;
    const int64_t len_c = diag_len(dc, rows_a, cols_b);
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < len_c; ++t) {
        const int64_t i = dc >= 0 ? t : t - dc;
        const int64_t j = dc >= 0 ? t + dc : t;
        double cr = 0.0, ci = 0.0;
        for (int p = 0; p < n_pairs; ++p) {
            const int sa = slot_a[p], sb = slot_b[p];
            const int da = off_a[sa], db = off_b[sb];
            const int64_t pa = da >= 0 ? i : i + da;
            const int64_t pb = db >= 0 ? i + da : j;
            if (pa < 0 || pb < 0) continue;
            if (pa >= diag_len(da, rows_a, cols_a) || pb >= diag_len(db, cols_a, cols_b)) continue;
            const double ar = a_re[st_a[sa] + pa], ai = a_im[st_a[sa] + pa];
            const double br = b_re[st_b[sb] + pb], bi = b_im[st_b[sb] + pb];
            cr += ar * br - ai * bi;
            ci += ar * bi + ai * br;
        }
        c_re[t] = cr;
        c_im[t] = ci;
    }
}

/* ---------------------------------------------------------------------------------------------------------------
 * The reference's CPU SpGEMM restated (src/spGEMM.cpp:303-514 with run_avx512<double>, src/simd_avx512_kernels.cpp:
 * 20-128): OpenMP over tiles of res_block result diagonals (CPU_RES_BLOCK = 4, src/spGEMM.cpp:19-24); inside a tile,
 * element blocks of elem_block positions (CPU_ELEM_BLOCK = 2048); per result diagonal its pairs in order, each adding
 * A_a * B_b over the positions both diagonals cover -- a contiguous, vectorised complex multiply-add. C (packed at
 * st_c) must be zeroed by the caller, as the reference's result is. The pair list is the structural dry-run.
 * In position space the pair maps result position t (row i = t + max(-dc, 0)) to A position t + ca and B position
 * t + cb, ca = i0 + min(da, 0), cb = i0 + (db >= 0 ? da : dc), i0 = max(-dc, 0): the reference's case analysis of
 * multiplyDiagonalsSegment (src/spGEMM.cpp:53-300) computes the same alignment.
 * ------------------------------------------------------------------------------------------------------------- */
__attribute__((target_clones("avx512f", "avx2", "default")))
static void cmadd_kernel(int64_t lo, int64_t hi, int64_t ca, int64_t cb, const double *restrict ar,
                         const double *restrict ai, const double *restrict br, const double *restrict bi,
                         double *restrict cr, double *restrict ci) {
    for (int64_t t = lo; t < hi; ++t) {
        const double a_r = ar[t + ca], a_i = ai[t + ca], b_r = br[t + cb], b_i = bi[t + cb];
        cr[t] += a_r * b_r - a_i * b_i;
        ci[t] += a_r * b_i + a_i * b_r;
    }
}

void oracle_product_tasks(int64_t rows_a, int64_t cols_a, int64_t cols_b, const int32_t *off_a, const int64_t *st_a,
                          const double *a_re, const double *a_im, const int32_t *off_b, const int64_t *st_b,
                          const double *b_re, const double *b_im, int n_res, const int32_t *res_off,
                          const int64_t *pair_start, const int32_t *slot_a, const int32_t *slot_b,
                          const int64_t *st_c, double *c_re, double *c_im, int res_block, int elem_block,
                          int n_threads) {
    if (res_block <= 0) res_block = 4;
    if (elem_block <= 0) elem_block = 2048;
    if (n_threads <= 0) n_threads = omp_get_max_threads();
    const int n_tiles = (n_res + res_block - 1) / res_block;
#pragma omp parallel for schedule(static) num_threads(n_threads)
    for (int tile = 0; tile < n_tiles; ++tile) {
        const int r0 = tile * res_block, r1 = r0 + res_block < n_res ? r0 + res_block : n_res;
        int64_t max_len = 0;
        for (int rs = r0; rs < r1; ++rs) {
            const int64_t l = diag_len(res_off[rs], rows_a, cols_b);
            if (l > max_len) max_len = l;
        }
        for (int64_t j0 = 0; j0 < max_len; j0 += elem_block) {
            const int64_t j1 = j0 + elem_block < max_len ? j0 + elem_block : max_len;
            for (int rs = r0; rs < r1; ++rs) {
                const int dc = res_off[rs];
                const int64_t len_c = diag_len(dc, rows_a, cols_b);
                if (j0 >= len_c) continue;
                const int64_t je = j1 < len_c ? j1 : len_c;
                const int64_t i0 = dc < 0 ? -(int64_t)dc : 0;
                for (int64_t p = pair_start[rs]; p < pair_start[rs + 1]; ++p) {
                    const int sa = slot_a[p], sb = slot_b[p];
                    const int da = off_a[sa], db = off_b[sb];
                    const int64_t ca = i0 + (da < 0 ? da : 0);
                    const int64_t cb = i0 + (db >= 0 ? da : dc);
                    const int64_t len_a = diag_len(da, rows_a, cols_a), len_b = diag_len(db, cols_a, cols_b);
                    int64_t lo = j0, hi = je;
                    if (lo < -ca) lo = -ca;
                    if (lo < -cb) lo = -cb;
                    if (hi > len_a - ca) hi = len_a - ca;
                    if (hi > len_b - cb) hi = len_b - cb;
                    if (hi <= lo) continue;
                    cmadd_kernel(lo, hi, ca, cb, a_re + st_a[sa], a_im + st_a[sa], b_re + st_b[sb], b_im + st_b[sb],
                                 c_re + st_c[rs], c_im + st_c[rs]);
                }
            }
        }
    }
}

/* ---------------------------------------------------------------------------------------------------------------
 * One fixed-degree Taylor step on the CPU with the reference's CPU SpMV (oracle_spmv_tasks above):
 * out = sum_{k=0}^{m} (-i dt H)^k psi / k!, term_k = (-i dt/k) H term_{k-1} (the arithmetic of
 * src/spHamSim_gpu_all_fused.cu:235-277, each SpMV the reference's CPU kernel). Scratch: 4*rows doubles.
 * ------------------------------------------------------------------------------------------------------------- */
int oracle_taylor_step_tasks(int64_t rows, int n_diags, const int32_t *offsets, const int64_t *starts,
                             const double *re, const double *im, const double *pr, const double *pi, double dt, int m,
                             double *outr, double *outi, double *scratch, int n_threads) {
    double *tr = scratch, *ti = scratch + rows, *hr = scratch + 2 * rows, *hi = scratch + 3 * rows;
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(n_threads)
    for (int64_t i = 0; i < rows; ++i) {
        tr[i] = pr[i];
        ti[i] = pi[i];
        outr[i] = pr[i];
        outi[i] = pi[i];
    }
    for (int k = 1; k <= m; ++k) {
        if (oracle_spmv_tasks(rows, rows, n_diags, offsets, starts, re, im, tr, ti, hr, hi, 0, n_threads)) return -1;
        const double s = dt / (double)k;
#pragma omp parallel for schedule(static) num_threads(n_threads)
        for (int64_t i = 0; i < rows; ++i) {
            const double vr = hi[i] * s, vi = -hr[i] * s; /* (-i s)(a + ib) */
            tr[i] = vr;
            ti[i] = vi;
            outr[i] += vr;
            outi[i] += vi;
        }
    }
    return 0;
}

/* ---------------------------------------------------------------------------------------------------------------
 * y += (I_{2^before} (x) G (x) I_{2^after}) x with the reference's CPU loop nest
 * (src/spKron.cpp:433-496, identity_kron_matrix_kron_identity_times_vector): parallel over (block rep, column j)
 * pairs (omp collapse(2)), then per diagonal slot of the square g x g gate a strided run of len elements.
 * G's diagonals are packed: slot k holds len_k = g - |d_k| values from starts[k].
 * ------------------------------------------------------------------------------------------------------------- */
void oracle_kron_apply_tasks(int before, int g, int n_diags, const int32_t *offsets, const int64_t *starts,
                             const double *gre, const double *gim, int after, const double *xr, const double *xi,
                             double *yr, double *yi, int n_threads) {
    const int64_t B = (int64_t)1 << before, A = (int64_t)1 << after;
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel for collapse(2) schedule(static) num_threads(n_threads)
    for (int64_t rep = 0; rep < B; ++rep) {
        for (int64_t j = 0; j < A; ++j) {
            for (int k = 0; k < n_diags; ++k) {
                const int64_t d = offsets[k], ad = d < 0 ? -d : d, len = g - ad;
                const double *vr = gre + starts[k], *vi = gim + starts[k];
                const int64_t skip = rep * (A * (len + ad));
                const int64_t xo = d < 0 ? skip : A * d + skip, yo = d < 0 ? skip + A * ad : skip;
                for (int64_t i = 0; i < len; ++i) {
                    const int64_t xa = xo + i * A + j, ya = yo + i * A + j;
                    yr[ya] += vr[i] * xr[xa] - vi[i] * xi[xa];
                    yi[ya] += vr[i] * xi[xa] + vi[i] * xr[xa];
                }
            }
        }
    }
}
