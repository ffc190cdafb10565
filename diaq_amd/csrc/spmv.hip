// DIA SpMV on gfx950: y = A x with A, x, y resident in HBM.
//
// Reference: spmv_single_diag_kernel / spmv_row_kernel / spmv_device (src/spMV_gpu.cu:17-168), spmv_host_vectors
// (:171-183), spMV_ntimes_cuda (src/spMV_ntimes_gpu.cu:61-263).
//
// One kernel serves every diagonal count (the reference's one-diagonal kernel gives the same values: y is zero
// except on the rows of the lone diagonal). A workgroup owns a tile of 256*RPT consecutive rows; lane l of the tile
// handles rows tile + l + 256*j. For each diagonal (a uniform loop, metadata through scalar loads) the wave reads 64
// consecutive values of the diagonal (one aligned 512-byte FP64 segment, row-aligned layout) and 64 consecutive
// values of the shifted x window, so every access is a full, coalesced line. Each row sums its diagonals in slot
// order in registers and y is written once: A is read once, y written once, and x is read once per diagonal (the
// compulsory bytes count it once; re-reads hit L2 or the Infinity Cache when the windows are near each other).
//
// Column bands: with n_bands > 1 the kernel runs once per band of columns [c0, c1), each pass adding only the
// elements whose column lies in the band (the first pass writes y, later passes add to it). A pass touches only
// 1/n_bands of x, so its x lines stay in the 256 MB Infinity Cache between the diagonals that share them, at the cost
// of one read and one write of y per extra pass.
//
// FP32: values and vectors are float; each row's diagonal sum is Kahan-compensated in FP32.
#include <algorithm>
#include <cstdlib>

#include "dia_rows.hpp"

namespace dq {

// Process-wide tuning knobs (set through dq_set_tuning).
static int g_spmv_rpt = 0;    // rows per lane; 0 = automatic
static int g_spmv_bands = 0;  // column bands; 0 = automatic

template <typename T, int RPT>
__global__ __launch_bounds__(256) void spmv_rows_kernel(int64_t rows, int n, const DiagMeta *__restrict__ meta,
                                                        const T *__restrict__ are, const T *__restrict__ aim,
                                                        const T *__restrict__ xr, const T *__restrict__ xi,
                                                        T *__restrict__ yr, T *__restrict__ yi, int64_t col_lo,
                                                        int64_t col_hi, int accumulate) {
    constexpr int TILE = 256 * RPT;
    const int64_t tile = static_cast<int64_t>(blockIdx.x) * TILE;
    const int64_t tile_end = tile + TILE < rows ? tile + TILE : rows;
    Acc<T> ar[RPT], ai[RPT];
    if (accumulate) {
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const int64_t r = tile + threadIdx.x + 256 * j;
            if (r < rows) { ar[j].add(yr[r]); ai[j].add(yi[r]); }
        }
    }
    dia_rows_tile<T, RPT>(tile, tile_end, n, meta, are, aim, xr, xi, col_lo, col_hi, ar, ai);
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        const int64_t r = tile + threadIdx.x + 256 * j;
        if (r < rows) { yr[r] = ar[j].get(); yi[r] = ai[j].get(); }
    }
}

template <typename T>
static void launch_rows(const dq_devmat &A, const T *xr, const T *xi, T *yr, T *yi, int rpt, int64_t c0, int64_t c1,
                        int accumulate, hipStream_t s) {
    const int64_t rows = A.rows;
    const DiagMeta *meta = A.d_meta.as<DiagMeta>();
    const T *are = static_cast<const T *>(A.re.p), *aim = static_cast<const T *>(A.im.p);
#define DQ_SPMV_CASE(R)                                                                                           \
    case R:                                                                                                       \
        hipLaunchKernelGGL((spmv_rows_kernel<T, R>), dim3(blocks_for(rows, 256 * R, 1LL << 31)), dim3(256), 0, s, \
                           rows, A.n(), meta, are, aim, xr, xi, yr, yi, c0, c1, accumulate);                      \
        break;
    switch (rpt) {
        DQ_SPMV_CASE(1)
        DQ_SPMV_CASE(2)
        DQ_SPMV_CASE(4)
        DQ_SPMV_CASE(8)
        default: throw ArgError("spmv: rows per lane must be 1, 2, 4 or 8");
    }
#undef DQ_SPMV_CASE
    DQ_LAUNCH_CHECK();
}

template <typename T>
__global__ void zero2_kernel(int64_t n, T *__restrict__ a, T *__restrict__ b) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        a[i] = T(0);
        b[i] = T(0);
    }
}

static int auto_bands(const dq_devmat &A) {
    if (g_spmv_bands > 0) return g_spmv_bands;
    return 1;
}

static int auto_rpt(const dq_devmat &A) {
    if (g_spmv_rpt > 0) return g_spmv_rpt;
    return A.precision == DQ_F64 ? 2 : 4;
}

void spmv_banded(const dq_devmat &A, const void *xr, const void *xi, void *yr, void *yi, int n_bands,
                 hipStream_t s) {
    if (A.rows == 0) return;
    if (A.n() == 0) {  // no diagonals: y = 0 (a fill kernel, safe inside graph capture)
        const unsigned g = blocks_for(A.rows, 256, 4096);
        if (A.precision == DQ_F64)
            hipLaunchKernelGGL(zero2_kernel<double>, dim3(g), dim3(256), 0, s, (int64_t)A.rows, (double *)yr,
                               (double *)yi);
        else
            hipLaunchKernelGGL(zero2_kernel<float>, dim3(g), dim3(256), 0, s, (int64_t)A.rows, (float *)yr,
                               (float *)yi);
        DQ_LAUNCH_CHECK();
        return;
    }
    if (n_bands <= 0) n_bands = auto_bands(A);
    if (n_bands > static_cast<int>(A.cols)) n_bands = std::max<int>(1, A.cols);
    const int rpt = auto_rpt(A);
    const int64_t cols = A.cols;
    for (int b = 0; b < n_bands; ++b) {
        const int64_t c0 = cols * b / n_bands, c1 = cols * (b + 1) / n_bands;
        if (A.precision == DQ_F64)
            launch_rows<double>(A, (const double *)xr, (const double *)xi, (double *)yr, (double *)yi, rpt, c0, c1,
                                b > 0, s);
        else
            launch_rows<float>(A, (const float *)xr, (const float *)xi, (float *)yr, (float *)yi, rpt, c0, c1, b > 0,
                               s);
    }
}

void spmv(const dq_devmat &A, const void *xr, const void *xi, void *yr, void *yi, hipStream_t s) {
    spmv_banded(A, xr, xi, yr, yi, 0, s);
}

}  // namespace dq

using namespace dq;

extern "C" int dq_set_tuning(const char *key, int value) {
    return guard([&] {
        if (!key) throw ArgError("key is NULL");
        const std::string k(key);
        if (k == "spmv_rows_per_lane") {
            if (value != 0 && value != 1 && value != 2 && value != 4 && value != 8)
                throw ArgError("spmv_rows_per_lane must be 0, 1, 2, 4 or 8");
            g_spmv_rpt = value;
        } else if (k == "spmv_bands") {
            if (value < 0 || value > 4096) throw ArgError("spmv_bands must be in [0, 4096]");
            g_spmv_bands = value;
        } else {
            throw ArgError("unknown tuning key: " + k);
        }
    });
}

extern "C" int dq_spmv_device(const dq_devmat *A, const void *xr, const void *xi, void *yr, void *yi,
                              uintptr_t stream) {
    return guard([&] {
        if (!A) throw ArgError("matrix is NULL");
        if (A->rows && (!yr || !yi)) throw ArgError("spmv_device: NULL y");
        if (A->n() && (!xr || !xi)) throw ArgError("spmv_device: NULL x");
        spmv(*A, xr, xi, yr, yi, as_stream(stream));
    });
}

extern "C" int dq_spmv_device_banded(const dq_devmat *A, const void *xr, const void *xi, void *yr, void *yi,
                                     int n_bands, uintptr_t stream) {
    return guard([&] {
        if (!A) throw ArgError("matrix is NULL");
        if (n_bands < 0) throw ArgError("n_bands must be >= 0");
        if (A->rows && (!yr || !yi)) throw ArgError("spmv_device: NULL y");
        if (A->n() && (!xr || !xi)) throw ArgError("spmv_device: NULL x");
        spmv_banded(*A, xr, xi, yr, yi, n_bands, as_stream(stream));
    });
}

// Host vectors of doubles through device buffers of the matrix precision.
namespace {
struct DevVec {
    DevBuf r, i;
    int64_t n;
    int prec;
    DevVec(int64_t n_, int prec_) : r(n_ * elem_bytes(prec_)), i(n_ * elem_bytes(prec_)), n(n_), prec(prec_) {}
    void put(const double *hr, const double *hi) {
        if (!n) return;
        if (prec == DQ_F64) {
            DQ_HIP(hipMemcpy(r.p, hr, n * 8, hipMemcpyHostToDevice));
            DQ_HIP(hipMemcpy(i.p, hi, n * 8, hipMemcpyHostToDevice));
        } else {
            std::vector<float> a(hr, hr + n), b(hi, hi + n);
            DQ_HIP(hipMemcpy(r.p, a.data(), n * 4, hipMemcpyHostToDevice));
            DQ_HIP(hipMemcpy(i.p, b.data(), n * 4, hipMemcpyHostToDevice));
        }
    }
    void get(double *hr, double *hi) const {
        if (!n) return;
        if (prec == DQ_F64) {
            DQ_HIP(hipMemcpy(hr, r.p, n * 8, hipMemcpyDeviceToHost));
            DQ_HIP(hipMemcpy(hi, i.p, n * 8, hipMemcpyDeviceToHost));
        } else {
            std::vector<float> a(n), b(n);
            DQ_HIP(hipMemcpy(a.data(), r.p, n * 4, hipMemcpyDeviceToHost));
            DQ_HIP(hipMemcpy(b.data(), i.p, n * 4, hipMemcpyDeviceToHost));
            std::copy(a.begin(), a.end(), hr);
            std::copy(b.begin(), b.end(), hi);
        }
    }
};
}  // namespace

extern "C" int dq_spmv_host(const dq_devmat *A, const double *xr, const double *xi, double *yr, double *yi) {
    return guard([&] {
        if (!A) throw ArgError("matrix is NULL");
        if ((A->cols && (!xr || !xi)) || (A->rows && (!yr || !yi))) throw ArgError("spmv_cuda: NULL vector");
        DevVec x(A->cols, A->precision), y(A->rows, A->precision);
        x.put(xr, xi);
        spmv(*A, x.r.p, x.i.p, y.r.p, y.i.p, 0);
        DQ_LAUNCH_CHECK();
        y.get(yr, yi);
    });
}

extern "C" int dq_spmv_ntimes_host(const dq_devmat *A, const double *xr, const double *xi, double *yr, double *yi,
                                   int n) {
    return guard([&] {
        if (!A) throw ArgError("matrix is NULL");
        if (n < 0) throw ArgError("spmv_ntimes_cuda: num_times must be >= 0");
        const int64_t M = A->rows, N = A->cols;
        if (n == 0) {
            if (M != N)
                throw ArgError("spmv_ntimes_cuda: num_times==0 requires square matrix (M==N) to copy x -> y");
            std::copy(xr, xr + M, yr);
            std::copy(xi, xi + M, yi);
            return;
        }
        if (A->n() == 0) {
            std::fill(yr, yr + M, 0.0);
            std::fill(yi, yi + M, 0.0);
            return;
        }
        if (n > 1 && M != N)
            throw ArgError("spmv_ntimes_cuda: repeated SpMV requires square matrix (M==N) for num_times > 1");
        DevVec a(N, A->precision), b(M, A->precision);
        a.put(xr, xi);
        DevVec *in = &a, *out = &b;
        for (int t = 0; t < n; ++t) {
            spmv(*A, in->r.p, in->i.p, out->r.p, out->i.p, 0);
            if (t != n - 1) std::swap(in, out);
        }
        out->get(yr, yi);
    });
}
