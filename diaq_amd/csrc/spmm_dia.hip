// DIA x dense on device pointers (torch / cupy data pointers): C = A B with A in row-aligned DIA form.
//
// Reference: spmm_dia_kernel / launch_rows / launch / spmm_dia_device (src/spMM_device_gpu.cu:44-127) and the
// viability rule dia_viability (src/dia_viability.hpp:42-63).
//
// Layout (as the reference): vals[k * n_rows + r] = A[r, r + offsets[k]]; B is n_cols x n_right and C n_rows x
// n_right, both row-major; double or interleaved complex double.
//
// Default mapping on gfx950 (spmm_dia_cols_kernel): a lane owns two adjacent columns of one row of C when n_right is
// even, one element otherwise; many rows with wide right-hand sides use the row-block form (spmm_dia_rows_kernel).
// One element per lane (spmm_dia_kernel): lane e owns C's flat element e = r * n_right + j. Its B operand for
// diagonal k is B[(r + o_k) * n_right + j] = B[e + o_k * n_right]: for every diagonal a wave reads one contiguous run
// of B and writes one contiguous run of C, whatever n_right is (the reference assigns a column per thread and tiles
// rows, which coalesces only for wide B). The diagonal values vals[k * n_rows + r] are shared by the n_right lanes of
// a row.
// Each element sums its diagonals in order, as the reference does; the loads of 8 diagonals are issued before their
// products are added, so a lane keeps 16 loads in flight, and a diagonal's offset and base pointers are uniform
// (scalar) values.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "dq_common.hpp"

namespace dq {

constexpr int SPMM_UNROLL = 8;

struct cplx {
    double x, y;
};

// acc + a b with one fixed rounding sequence (explicit fmas), so every kernel variant gives the same bits.
__device__ inline double fma_val(double acc, double a, double b) { return fma(a, b, acc); }
__device__ inline cplx fma_val(cplx acc, cplx a, cplx b) {
    return cplx{fma(-a.y, b.y, fma(a.x, b.x, acc.x)), fma(a.y, b.x, fma(a.x, b.y, acc.y))};
}
// The diagonal values are read once and C is written once: non-temporal, so B's rows keep the caches.
typedef double d2v __attribute__((ext_vector_type(2)));  // one 16-byte access for a complex value
__device__ inline double ld_stream(const double *p) { return __builtin_nontemporal_load(p); }
__device__ inline cplx ld_stream(const cplx *p) {
    const d2v v = __builtin_nontemporal_load(reinterpret_cast<const d2v *>(p));
    return cplx{v.x, v.y};
}
__device__ inline void st_stream(double *p, double v) { __builtin_nontemporal_store(v, p); }
__device__ inline void st_stream(cplx *p, cplx v) {
    d2v w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, reinterpret_cast<d2v *>(p));
}
template <typename T> __device__ inline T zero_val();
template <> __device__ inline double zero_val<double>() { return 0.0; }
template <> __device__ inline cplx zero_val<cplx>() { return cplx{0.0, 0.0}; }

template <typename T, typename I>
__global__ __launch_bounds__(256) void spmm_dia_kernel(I n_rows, I n_cols, I n_right, int n_diags,
                                                       const int64_t *__restrict__ offsets, const T *__restrict__ vals,
                                                       const T *__restrict__ B, T *__restrict__ C) {
    const I total = n_rows * n_right;
    for (I e = static_cast<I>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
         e += static_cast<I>(gridDim.x) * blockDim.x) {
        const I r = e / n_right;
        T acc = zero_val<T>();
        // SPMM_UNROLL diagonals at a time: their loads are issued together, then added in diagonal order. The
        // diagonal's offset and base pointers are uniform (scalar loads and arithmetic); a lane adds only its row.
        for (int k0 = 0; k0 < n_diags; k0 += SPMM_UNROLL) {
            T a[SPMM_UNROLL], b[SPMM_UNROLL];
            bool ok[SPMM_UNROLL];
            int64_t o[SPMM_UNROLL];
#pragma unroll
            for (int u = 0; u < SPMM_UNROLL; ++u)  // all offsets first: independent scalar loads, one wait
                o[u] = offsets[k0 + u < n_diags ? k0 + u : n_diags - 1];
#pragma unroll
            for (int u = 0; u < SPMM_UNROLL; ++u) {
                const int k = k0 + u;
                const int64_t c = static_cast<int64_t>(r) + o[u];
                ok[u] = k < n_diags && c >= 0 && c < static_cast<int64_t>(n_cols);
                a[u] = zero_val<T>();
                b[u] = zero_val<T>();
                if (ok[u]) {
                    a[u] = ld_stream(vals + static_cast<int64_t>(k) * n_rows + r);
                    b[u] = B[static_cast<int64_t>(e) + o[u] * static_cast<int64_t>(n_right)];
                }
            }
#pragma unroll
            for (int u = 0; u < SPMM_UNROLL; ++u)
                if (ok[u]) acc = fma_val(acc, a[u], b[u]);
        }
        st_stream(C + e, acc);
    }
}

// Row-chunk form: lane w owns row r = w / (n_right / J) and the J consecutive columns [jc J, jc J + J) of C, with
// jc = w mod (n_right / J) (J divides n_right). Per diagonal it reads the row's value once and J consecutive B values
// of one row of B (contiguous 16-byte loads), instead of one value and one B element per lane: a wave moves the same
// B bytes with J times fewer value loads. U diagonals' loads are issued before their products are added in order.
template <typename T, typename I, int J, int U>
__global__ __launch_bounds__(256) void spmm_dia_cols_kernel(I n_rows, I n_cols, I n_right, int n_diags,
                                                            const int64_t *__restrict__ offsets,
                                                            const T *__restrict__ vals, const T *__restrict__ B,
                                                            T *__restrict__ C) {
    const I nchunk = n_right / J;
    const I total = n_rows * nchunk;
    for (I w = static_cast<I>(blockIdx.x) * blockDim.x + threadIdx.x; w < total;
         w += static_cast<I>(gridDim.x) * blockDim.x) {
        const I r = w / nchunk;
        const int64_t e0 = static_cast<int64_t>(r) * n_right + static_cast<int64_t>(w - r * nchunk) * J;
        T acc[J];
#pragma unroll
        for (int j = 0; j < J; ++j) acc[j] = zero_val<T>();
        for (int k0 = 0; k0 < n_diags; k0 += U) {
            T a[U], b[U][J];
            bool ok[U];
            int64_t o[U];
#pragma unroll
            for (int u = 0; u < U; ++u) o[u] = offsets[k0 + u < n_diags ? k0 + u : n_diags - 1];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = k0 + u;
                const int64_t c = static_cast<int64_t>(r) + o[u];
                ok[u] = k < n_diags && c >= 0 && c < static_cast<int64_t>(n_cols);
                a[u] = zero_val<T>();
#pragma unroll
                for (int j = 0; j < J; ++j) b[u][j] = zero_val<T>();
                if (ok[u]) {
                    a[u] = ld_stream(vals + static_cast<int64_t>(k) * n_rows + r);
                    const T *bp = B + e0 + o[u] * static_cast<int64_t>(n_right);
#pragma unroll
                    for (int j = 0; j < J; ++j) b[u][j] = bp[j];
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (ok[u]) {
#pragma unroll
                    for (int j = 0; j < J; ++j) acc[j] = fma_val(acc[j], a[u], b[u][j]);
                }
        }
#pragma unroll
        for (int j = 0; j < J; ++j) st_stream(C + e0 + j, acc[j]);
    }
}

// Row-block form for wide right-hand sides: a lane owns J adjacent columns of R consecutive rows of C. It keeps the
// B rows its R rows need for the current diagonal in registers; when the next diagonal's offset is one more (a
// band), its rows are the window shifted by one, so one new B row is loaded instead of R. Other offsets reload the
// window. Each element adds its diagonals in slot order with the same fma_val sequence as the other forms.
template <typename T, typename I, int J, int R>
__global__ __launch_bounds__(256) void spmm_dia_rows_kernel(I n_rows, I n_cols, I n_right, int n_diags,
                                                            const int64_t *__restrict__ offsets,
                                                            const T *__restrict__ vals, const T *__restrict__ B,
                                                            T *__restrict__ C) {
    const I nchunk = n_right / J;
    const I nblk = (n_rows + R - 1) / R;
    const int64_t total = static_cast<int64_t>(nblk) * nchunk;
    for (int64_t w = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; w < total;
         w += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t blk = w / nchunk;
        const int64_t col = (w - blk * nchunk) * J;
        const int64_t r0 = blk * R;
        T acc[R][J], win[R][J];
#pragma unroll
        for (int i = 0; i < R; ++i)
#pragma unroll
            for (int j = 0; j < J; ++j) acc[i][j] = zero_val<T>();
        auto load_row = [&](int64_t b, T (&dst)[J]) {  // B row b (zero outside B)
            const bool in = b >= 0 && b < static_cast<int64_t>(n_cols);
            const T *bp = B + (in ? b : 0) * static_cast<int64_t>(n_right) + col;
#pragma unroll
            for (int j = 0; j < J; ++j) dst[j] = in ? bp[j] : zero_val<T>();
        };
        int64_t prev = 0;
        for (int d = 0; d < n_diags; ++d) {
            const int64_t o = offsets[d];  // uniform
            if (d > 0 && o == prev + 1) {  // uniform: a band step, the window moves down one row
#pragma unroll
                for (int i = 0; i + 1 < R; ++i)
#pragma unroll
                    for (int j = 0; j < J; ++j) win[i][j] = win[i + 1][j];
                load_row(r0 + R - 1 + o, win[R - 1]);
            } else {
#pragma unroll
                for (int i = 0; i < R; ++i) load_row(r0 + i + o, win[i]);
            }
            prev = o;
#pragma unroll
            for (int i = 0; i < R; ++i) {
                const int64_t r = r0 + i;
                const int64_t c = r + o;
                if (r < static_cast<int64_t>(n_rows) && c >= 0 && c < static_cast<int64_t>(n_cols)) {
                    const T a = ld_stream(vals + static_cast<int64_t>(d) * n_rows + r);
#pragma unroll
                    for (int j = 0; j < J; ++j) acc[i][j] = fma_val(acc[i][j], a, win[i][j]);
                }
            }
        }
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const int64_t r = r0 + i;
            if (r < static_cast<int64_t>(n_rows)) {
#pragma unroll
                for (int j = 0; j < J; ++j) st_stream(C + r * static_cast<int64_t>(n_right) + col + j, acc[i][j]);
            }
        }
    }
}

// Tiled form: workgroup b owns rows [r0, r0 + TR) with TR = SPMM_TILE / n_right (SPMM_EPT elements per lane). When
// two or more diagonals have |offset| <= TR/2 it first stages B's rows [r0 - TR/2, r0 + 3 TR/2) (zero outside B) in
// LDS and those diagonals read B from there; the others read it from memory. Each element still sums its diagonals
// in order.
template <typename T, typename I, int SPMM_EPT>
__global__ __launch_bounds__(256) void spmm_dia_tiled_kernel(I n_rows, I n_cols, I n_right, int n_diags,
                                                             const int64_t *__restrict__ offsets,
                                                             const T *__restrict__ vals, const T *__restrict__ B,
                                                             T *__restrict__ C, I rows_per_tile) {
    __shared__ T sb[2 * 256 * SPMM_EPT];
    const int64_t TR = rows_per_tile;
    // the window pays only when two or more diagonals fall inside it (uniform: every lane reads the same offsets)
    int near = 0;
    for (int k = 0; k < n_diags; ++k) {
        const int64_t o = offsets[k];
        near += (o >= -(TR / 2) && o <= TR / 2) ? 1 : 0;
    }
    const int64_t H = near >= 2 ? TR / 2 : -1;  // H < 0: no window, every diagonal reads B from memory
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * TR;
    if (H >= 0) {
        const int64_t W = (TR + 2 * H) * static_cast<int64_t>(n_right);  // <= 2 * 256 * SPMM_EPT
        const int64_t b0 = (r0 - H) * static_cast<int64_t>(n_right);
        const int64_t nb = static_cast<int64_t>(n_cols) * n_right;
        for (int64_t e = threadIdx.x; e < W; e += 256) {
            const int64_t gi = b0 + e;
            T v = zero_val<T>();
            if (gi >= 0 && gi < nb) v = B[gi];
            sb[e] = v;
        }
        __syncthreads();
    }
    const int64_t rows_here = min(TR, static_cast<int64_t>(n_rows) - r0);
    const I elems = static_cast<I>(rows_here * n_right);
#pragma unroll
    for (int u = 0; u < SPMM_EPT; ++u) {
        const I local = static_cast<I>(threadIdx.x + 256 * u);
        if (local >= elems) break;
        const I lr = local / n_right;
        const I r = static_cast<I>(r0) + lr;
        const int64_t e = static_cast<int64_t>(r0) * n_right + local;
        T acc = zero_val<T>();
        for (int k0 = 0; k0 < n_diags; k0 += SPMM_UNROLL) {
            T a[SPMM_UNROLL], b[SPMM_UNROLL];
            bool ok[SPMM_UNROLL];
            int64_t o[SPMM_UNROLL];
#pragma unroll
            for (int w = 0; w < SPMM_UNROLL; ++w) o[w] = offsets[k0 + w < n_diags ? k0 + w : n_diags - 1];
#pragma unroll
            for (int w = 0; w < SPMM_UNROLL; ++w) {
                const int k = k0 + w;
                const int64_t c = static_cast<int64_t>(r) + o[w];
                ok[w] = k < n_diags && c >= 0 && c < static_cast<int64_t>(n_cols);
                a[w] = zero_val<T>();
                b[w] = zero_val<T>();
                if (ok[w]) {
                    a[w] = ld_stream(vals + static_cast<int64_t>(k) * n_rows + r);
                    if (o[w] >= -H && o[w] <= H)  // uniform
                        b[w] = sb[local + (H + o[w]) * static_cast<int64_t>(n_right)];
                    else
                        b[w] = B[e + o[w] * static_cast<int64_t>(n_right)];
                }
            }
#pragma unroll
            for (int w = 0; w < SPMM_UNROLL; ++w)
                if (ok[w]) acc = fma_val(acc, a[w], b[w]);
        }
        st_stream(C + e, acc);
    }
}

template <typename T, int J>
static void launch_cols(int64_t n_rows, int64_t n_cols, int64_t n_right, int n_diags, const int64_t *off,
                        const T *vals, const T *B, T *C, bool small, hipStream_t s) {
    constexpr int U = J >= 8 ? 4 : 8;
    const unsigned grid = blocks_for(n_rows * (n_right / J), 256, 1 << 20);
    if (small)
        hipLaunchKernelGGL((spmm_dia_cols_kernel<T, uint32_t, J, U>), dim3(grid), dim3(256), 0, s, (uint32_t)n_rows,
                           (uint32_t)n_cols, (uint32_t)n_right, n_diags, off, vals, B, C);
    else
        hipLaunchKernelGGL((spmm_dia_cols_kernel<T, int64_t, J, U>), dim3(grid), dim3(256), 0, s, n_rows, n_cols,
                           n_right, n_diags, off, vals, B, C);
}

// dq_set_tuning "spmm_cols_per_lane": 0 = automatic (2 when n_right is even, else 1), 1, 2, 4, 8 = that many columns
// per lane when it divides n_right, -1 = one C element per lane (spmm_dia_kernel). Two columns per lane measured
// fastest for every even n_right tried (8, 16, 32, 128: 10-22 % over one element per lane); 4 and 8 were slower (a
// wave's loads then spread over 4-8 times more cache lines per instruction).
static int g_spmm_cols = 0;
// dq_set_tuning "spmm_rows_per_lane": rows per lane of the row-block form (spmm_dia_rows_kernel); -1 = automatic
// (2 rows per lane for at least 32768 rows and an even n_right of at least 64, else off), 0 = off, 2, 4, 8. Measured
// on 65536 rows x 128 columns: 1.34x for a 256-band, 1.11x for 256 scattered diagonals; with 16 columns, or 4096
// rows (too few lanes), the flat forms are faster.
static int g_spmm_rows = -1;

template <typename T, int J, int R>
static void launch_rows_block(int64_t n_rows, int64_t n_cols, int64_t n_right, int n_diags, const int64_t *off,
                              const T *vals, const T *B, T *C, bool small, hipStream_t s) {
    const int64_t work = ((n_rows + R - 1) / R) * (n_right / J);
    const unsigned grid = blocks_for(work, 256, 1 << 20);
    if (small)
        hipLaunchKernelGGL((spmm_dia_rows_kernel<T, uint32_t, J, R>), dim3(grid), dim3(256), 0, s, (uint32_t)n_rows,
                           (uint32_t)n_cols, (uint32_t)n_right, n_diags, off, vals, B, C);
    else
        hipLaunchKernelGGL((spmm_dia_rows_kernel<T, int64_t, J, R>), dim3(grid), dim3(256), 0, s, n_rows, n_cols,
                           n_right, n_diags, off, vals, B, C);
}

template <typename T>
static void launch(int64_t n_rows, int64_t n_cols, int64_t n_right, int n_diags, const int64_t *off, const T *vals,
                   const T *B, T *C, int ept, hipStream_t s) {
    const int64_t total = n_rows * n_right;
    if (total == 0) return;
    const bool small = total < (int64_t(1) << 31) && n_cols < (int64_t(1) << 31);
    const int rows_per_lane =
        g_spmm_rows >= 0 ? g_spmm_rows : ((n_rows >= 32768 && n_right >= 64 && n_right % 2 == 0) ? 2 : 0);
    if (ept == 0 && rows_per_lane > 0) {
        const int J = n_right % 2 == 0 ? 2 : 1;
        if (rows_per_lane == 8) {
            if (J == 2) launch_rows_block<T, 2, 8>(n_rows, n_cols, n_right, n_diags, off, vals, B, C, small, s);
            else launch_rows_block<T, 1, 8>(n_rows, n_cols, n_right, n_diags, off, vals, B, C, small, s);
        } else if (rows_per_lane == 4) {
            if (J == 2) launch_rows_block<T, 2, 4>(n_rows, n_cols, n_right, n_diags, off, vals, B, C, small, s);
            else launch_rows_block<T, 1, 4>(n_rows, n_cols, n_right, n_diags, off, vals, B, C, small, s);
        } else {
            if (J == 2) launch_rows_block<T, 2, 2>(n_rows, n_cols, n_right, n_diags, off, vals, B, C, small, s);
            else launch_rows_block<T, 1, 2>(n_rows, n_cols, n_right, n_diags, off, vals, B, C, small, s);
        }
        hipError_t err = hipGetLastError();
        if (err != hipSuccess) throw HipError(std::string("spmm_dia_device: ") + hipGetErrorString(err));
        return;
    }
    if (ept == 0 && g_spmm_cols >= 0) {
        int J = 1;
        if (g_spmm_cols == 0) J = n_right % 2 == 0 ? 2 : 1;
        else if (n_right % g_spmm_cols == 0) J = g_spmm_cols;
        if (J == 8) launch_cols<T, 8>(n_rows, n_cols, n_right, n_diags, off, vals, B, C, small, s);
        else if (J == 4) launch_cols<T, 4>(n_rows, n_cols, n_right, n_diags, off, vals, B, C, small, s);
        else if (J == 2) launch_cols<T, 2>(n_rows, n_cols, n_right, n_diags, off, vals, B, C, small, s);
        else launch_cols<T, 1>(n_rows, n_cols, n_right, n_diags, off, vals, B, C, small, s);
        hipError_t err = hipGetLastError();
        if (err != hipSuccess) throw HipError(std::string("spmm_dia_device: ") + hipGetErrorString(err));
        return;
    }
    if (ept > 0 && n_right <= 256 * ept) {
        const int64_t tr = 256 * ept / n_right;
        const int64_t tiles = (n_rows + tr - 1) / tr;
        if (tiles > 0x7fffffffLL) throw ArgError("spmm_dia_device: too many rows");
#define DQ_SPMM_TILED(E)                                                                                          \
    if (small)                                                                                                    \
        hipLaunchKernelGGL((spmm_dia_tiled_kernel<T, uint32_t, E>), dim3(static_cast<unsigned>(tiles)), dim3(256),  \
                           0, s, (uint32_t)n_rows, (uint32_t)n_cols, (uint32_t)n_right, n_diags, off, vals, B, C,   \
                           (uint32_t)tr);                                                                         \
    else                                                                                                          \
        hipLaunchKernelGGL((spmm_dia_tiled_kernel<T, int64_t, E>), dim3(static_cast<unsigned>(tiles)), dim3(256),   \
                           0, s, n_rows, n_cols, n_right, n_diags, off, vals, B, C, tr);
        if (ept == 1) { DQ_SPMM_TILED(1) }
        else if (ept == 2) { DQ_SPMM_TILED(2) }
        else { DQ_SPMM_TILED(4) }
#undef DQ_SPMM_TILED
        hipError_t err = hipGetLastError();
        if (err != hipSuccess) throw HipError(std::string("spmm_dia_device: ") + hipGetErrorString(err));
        return;
    }
    const unsigned grid = blocks_for(total, 256, 1 << 20);
    if (small)
        hipLaunchKernelGGL((spmm_dia_kernel<T, uint32_t>), dim3(grid), dim3(256), 0, s, (uint32_t)n_rows,
                           (uint32_t)n_cols, (uint32_t)n_right, n_diags, off, vals, B, C);
    else
        hipLaunchKernelGGL((spmm_dia_kernel<T, int64_t>), dim3(grid), dim3(256), 0, s, n_rows, n_cols, n_right,
                           n_diags, off, vals, B, C);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) throw HipError(std::string("spmm_dia_device: ") + hipGetErrorString(err));
}

// Process-wide knob (dq_set_tuning "spmm_tiled"): 0 = flat kernel (default), 1, 2 or 4 = the tiled kernel with that
// many elements per lane.
static int g_spmm_tiled = 0;

bool spmm_set_tuning(const std::string &key, int value) {
    if (key == "spmm_cols_per_lane") {
        if (value != -1 && value != 0 && value != 1 && value != 2 && value != 4 && value != 8)
            throw ArgError("spmm_cols_per_lane must be -1, 0, 1, 2, 4 or 8");
        g_spmm_cols = value;
        return true;
    }
    if (key == "spmm_rows_per_lane") {
        if (value != -1 && value != 0 && value != 2 && value != 4 && value != 8)
            throw ArgError("spmm_rows_per_lane must be -1, 0, 2, 4 or 8");
        g_spmm_rows = value;
        return true;
    }
    if (key != "spmm_tiled") return false;
    if (value != 0 && value != 1 && value != 2 && value != 4) throw ArgError("spmm_tiled must be 0, 1, 2 or 4");
    g_spmm_tiled = value;
    return true;
}


}  // namespace dq

using namespace dq;

extern "C" int dq_spmm_dia_device(uintptr_t offsets, uintptr_t vals, int n_diags, int64_t n_rows, int64_t n_cols,
                                  uintptr_t B, int64_t n_right, uintptr_t C, int is_complex, uintptr_t stream) {
    return guard([&] {
        if (n_diags < 0 || n_rows < 0 || n_cols < 0 || n_right < 0) throw ArgError("spmm_dia_device: negative size");
        if (n_rows * n_right == 0) return;
        if (!C || (n_diags > 0 && (!offsets || !vals)) || (n_cols > 0 && !B))
            throw ArgError("spmm_dia_device: NULL pointer");
        const auto *off = reinterpret_cast<const int64_t *>(offsets);
        const int ept = g_spmm_tiled;  // 0: flat kernel; 1, 2, 4: tiled, elements per lane
        if (is_complex)
            launch<cplx>(n_rows, n_cols, n_right, n_diags, off, reinterpret_cast<const cplx *>(vals),
                         reinterpret_cast<const cplx *>(B), reinterpret_cast<cplx *>(C), ept, as_stream(stream));
        else
            launch<double>(n_rows, n_cols, n_right, n_diags, off, reinterpret_cast<const double *>(vals),
                           reinterpret_cast<const double *>(B), reinterpret_cast<double *>(C), ept,
                           as_stream(stream));
    });
}

// Storage and diagonal budgets (src/dia_viability.hpp:36-63) with a set of constants: DIA bytes <= max_ratio x CSR
// bytes, and n_diags <= max(min_budget, n_rows / rows_per_diagonal), capped at wide_budget when n_right >= wide_right.
struct ViabilityRule {
    double max_ratio;
    int64_t min_budget, rows_per_diagonal, wide_right, wide_budget;
};

// The reference's constants (an A100 sweep) and the ones fitted on MI355X by `bench.py viability-sweep` over the same
// 360-case grid: no case passing them ran slower in DIA than hipSPARSE's CSR x dense, and they keep 228 of the 305
// cases where DIA was faster (the reference's constants keep 63).
constexpr ViabilityRule REFERENCE_RULE{0.9, 16, 32, 32, 8};
constexpr ViabilityRule MI355X_RULE{6.0, 64, 128, 16, 64};

static dq_dia_viability_t viability(const ViabilityRule &R, int64_t n_rows, int64_t n_diags, int64_t nnz,
                                    int64_t n_right, int itemsize, int index_bytes) {
    dq_dia_viability_t v;
    std::memset(&v, 0, sizeof(v));
    const double stored = static_cast<double>(n_diags) * static_cast<double>(n_rows);
    v.dia_bytes = stored * itemsize + static_cast<double>(n_diags) * 8.0;
    v.csr_bytes = static_cast<double>(nnz) * (itemsize + index_bytes) + static_cast<double>(n_rows + 1) * index_bytes;
    v.density = stored > 0 ? static_cast<double>(nnz) / stored : 0.0;
    v.storage_ok = n_diags > 0 && v.dia_bytes <= R.max_ratio * v.csr_bytes;
    v.diagonal_budget = std::max<int64_t>(R.min_budget, n_rows / R.rows_per_diagonal);
    if (n_right >= R.wide_right) v.diagonal_budget = std::min<int64_t>(v.diagonal_budget, R.wide_budget);
    v.diagonals_ok = n_diags <= v.diagonal_budget;
    v.viable = v.storage_ok && v.diagonals_ok;
    std::string reason;
    if (!v.storage_ok) {
        char pct[32];
        std::snprintf(pct, sizeof(pct), "%g%%", R.max_ratio * 100.0);
        reason = std::string("storage: DIA would take more than ") + pct +
                 " of CSR's bytes (diagonals too sparsely filled)";
    } else if (!v.diagonals_ok) {
        reason = "diagonals: " + std::to_string(n_diags) + " over the budget of " + std::to_string(v.diagonal_budget);
    } else {
        reason = "viable";
    }
    std::strncpy(v.reason, reason.c_str(), sizeof(v.reason) - 1);
    return v;
}

extern "C" int dq_dia_viability(int64_t n_rows, int64_t n_cols, int64_t n_diags, int64_t nnz, int64_t n_right,
                                int itemsize, int index_bytes, dq_dia_viability_t *out) {
    return guard([&] {
        if (!out) throw ArgError("out is NULL");
        (void)n_cols;
        *out = viability(REFERENCE_RULE, n_rows, n_diags, nnz, n_right, itemsize, index_bytes);
    });
}

extern "C" int dq_dia_viability_mi355x(int64_t n_rows, int64_t n_cols, int64_t n_diags, int64_t nnz,
                                       int64_t n_right, int itemsize, int index_bytes, dq_dia_viability_t *out) {
    return guard([&] {
        if (!out) throw ArgError("out is NULL");
        (void)n_cols;
        *out = viability(MI355X_RULE, n_rows, n_diags, nnz, n_right, itemsize, index_bytes);
    });
}
