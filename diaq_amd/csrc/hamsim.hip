// Fused time evolution on gfx950: psi(t) = prod_r exp(-i H t/r) psi0 with a fixed-m Taylor action per step, all on
// the device.
//
// Reference: spmv_row_diag_sparse_kernel, vec_copy / vec_axpy / vec_mul_minus_i_scale, expm_action_taylor_device
// and hamiltonian_simulation_cuda (src/spHamSim_gpu_all_fused.cu:117-376).
//
// The reference runs, per Taylor term, an SpMV (tmp = H term), a scaling pass (term = (-i dt/k) tmp) and an axpy
// (out += term), plus two copies per step and a device synchronisation. Here one kernel does all three: it forms
// h = H term_{k-1} row by row in registers (the shared row-tile core), writes term_k = (-i dt/k) h, i.e.
// (Im h * s, -Re h * s) with s = dt/k, and adds term_k to the step's output; the first term reads psi directly and
// writes out = psi + term_1, so the copies disappear. Per term: H read once, term_{k-1} read, term_k written and the
// output read and written once (b*S + 4*b*N bytes). The same operation order as the reference per element:
// row sums in slot order, then the scaling, then the add. The term kernel stages the x window of each row tile in LDS
// when the matrix's SpMV schedule does (tuned, or some off-diagonals within half a tile): the +-2^k offsets of a qubit
// Hamiltonian then share one copy of the term vector per tile.
#include <algorithm>

#include "dia_rows.hpp"

namespace dq {

template <typename T, int RPT, bool LDS>
__global__ __launch_bounds__(256) void taylor_term_kernel(int64_t rows, int n, const DiagMeta *__restrict__ meta,
                                                          const T *__restrict__ are, const T *__restrict__ aim,
                                                          const T *__restrict__ tr, const T *__restrict__ ti,
                                                          T *__restrict__ nr, T *__restrict__ ni,
                                                          T *__restrict__ outr, T *__restrict__ outi, double scale,
                                                          int first, int64_t ntiles, int xcd) {
    constexpr int TILE = 256 * RPT;
    constexpr int HALO = LDS ? TILE / 2 : 0;
    constexpr int W = LDS ? TILE + 2 * HALO : 1;
    __shared__ T sr[W], si[W];
    const int64_t t = tile_of_block(blockIdx.x, ntiles, xcd);
    if (t >= ntiles) return;  // uniform: a surplus workgroup of the XCD-contiguous grid
    const int64_t tile = t * TILE;
    const int64_t tile_end = tile + TILE < rows ? tile + TILE : rows;
    if constexpr (LDS) stage_x_window<T, RPT, HALO>(tile, rows, tr, ti, sr, si);
    Acc<T> hr[RPT], hi[RPT];
    dia_rows_tile<T, RPT, true, true, HALO>(tile, tile_end, n, meta, are, aim, tr, ti, 0, rows, hr, hi, sr, si);
    const T s = static_cast<T>(scale);
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        const int64_t r = tile + threadIdx.x + 256 * j;
        if (r < rows) {
            const T a = hr[j].get(), b = hi[j].get();
            const T vr = b * s, vi = -a * s;  // (-i s)(a + ib)
            nr[r] = vr;
            ni[r] = vi;
            if (first) {  // out = psi + term_1 (psi is the term_0 input)
                outr[r] = tr[r] + vr;
                outi[r] = ti[r] + vi;
            } else {
                outr[r] += vr;
                outi[r] += vi;
            }
        }
    }
}

template <typename T>
__global__ void copy2_kernel(int64_t n, const T *__restrict__ ar, const T *__restrict__ ai, T *__restrict__ br,
                             T *__restrict__ bi) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        br[i] = ar[i];
        bi[i] = ai[i];
    }
}

// One step out = sum_{k=0}^{m} (-i dt H)^k psi / k! (src/spHamSim_gpu_all_fused.cu:235-277). t0/t1: scratch terms.
template <typename T>
static void taylor_step(const dq_devmat &H, const T *pr, const T *pi, T *outr, T *outi, T *t0r, T *t0i, T *t1r,
                        T *t1i, double dt, int m, hipStream_t s) {
    const int64_t N = H.rows;
    if (m <= 0 || H.n() == 0) {  // out = psi (+ zero terms)
        hipLaunchKernelGGL(copy2_kernel<T>, dim3(blocks_for(N, 256, 8192)), dim3(256), 0, s, N, pr, pi, outr, outi);
        DQ_LAUNCH_CHECK();
        return;
    }
    constexpr int RPT = 8;
    const int64_t ntiles = (N + 256 * RPT - 1) / (256 * RPT);
    const int xcd = xcd_tiles_for(H);
    const unsigned grid = static_cast<unsigned>(grid_for_tiles(ntiles, xcd));
    const bool lds = lds_window_for(H, RPT);
    const T *inr = pr, *ini = pi;
    T *bufr[2] = {t0r, t1r}, *bufi[2] = {t0i, t1i};
    for (int k = 1; k <= m; ++k) {
        T *nr = bufr[k & 1], *ni = bufi[k & 1];
        if (lds)
            hipLaunchKernelGGL((taylor_term_kernel<T, RPT, true>), dim3(grid), dim3(256), 0, s, N, H.n(),
                               H.d_meta.as<DiagMeta>(), static_cast<const T *>(H.re.p),
                               static_cast<const T *>(H.im.p), inr, ini, nr, ni, outr, outi,
                               dt / static_cast<double>(k), k == 1 ? 1 : 0, ntiles, xcd);
        else
            hipLaunchKernelGGL((taylor_term_kernel<T, RPT, false>), dim3(grid), dim3(256), 0, s, N, H.n(),
                               H.d_meta.as<DiagMeta>(), static_cast<const T *>(H.re.p),
                               static_cast<const T *>(H.im.p), inr, ini, nr, ni, outr, outi,
                               dt / static_cast<double>(k), k == 1 ? 1 : 0, ntiles, xcd);
        DQ_LAUNCH_CHECK();
        inr = nr;
        ini = ni;
    }
}

// Evolve the device state (pr, pi) in place over r steps of dt = t / r.
template <typename T>
static void evolve(const dq_devmat &H, T *pr, T *pi, double t, uint64_t r, int m, hipStream_t s) {
    const int64_t N = H.rows;
    if (r == 0 || N == 0) return;
    const double dt = t / static_cast<double>(r);
    DevBuf br(N * sizeof(T)), bi(N * sizeof(T)), t0r(N * sizeof(T)), t0i(N * sizeof(T)), t1r(N * sizeof(T)),
        t1i(N * sizeof(T));
    T *ar = pr, *ai = pi, *xr = br.as<T>(), *xi = bi.as<T>();
    for (uint64_t step = 0; step < r; ++step) {
        taylor_step<T>(H, ar, ai, xr, xi, t0r.as<T>(), t0i.as<T>(), t1r.as<T>(), t1i.as<T>(), dt, m, s);
        std::swap(ar, xr);
        std::swap(ai, xi);
    }
    if (ar != pr) {  // odd number of steps: the result sits in the scratch pair
        hipLaunchKernelGGL(copy2_kernel<T>, dim3(blocks_for(N, 256, 8192)), dim3(256), 0, s, N, (const T *)ar,
                           (const T *)ai, pr, pi);
        DQ_LAUNCH_CHECK();
    }
    DQ_HIP(hipStreamSynchronize(s));  // scratch buffers are freed on return
}

void hamsim_device(const dq_devmat &H, void *pr, void *pi, double t, uint64_t r, int m, hipStream_t s) {
    if (H.rows != H.cols) throw ArgError("hamiltonian_simulation_cuda: H must be square");
    if (H.precision == DQ_F64)
        evolve<double>(H, static_cast<double *>(pr), static_cast<double *>(pi), t, r, m, s);
    else
        evolve<float>(H, static_cast<float *>(pr), static_cast<float *>(pi), t, r, m, s);
}

}  // namespace dq

using namespace dq;

extern "C" int dq_hamsim_device(const dq_devmat *H, void *pr, void *pi, double t, uint64_t r, int m,
                                uintptr_t stream) {
    return guard([&] {
        if (!H) throw ArgError("hamiltonian_simulation_cuda: H is null");
        if (H->rows && (!pr || !pi)) throw ArgError("hamiltonian_simulation_cuda: psi0 is null");
        hamsim_device(*H, pr, pi, t, r, m, as_stream(stream));
    });
}

extern "C" int dq_hamsim_host(const dq_devmat *H, const double *psi_r, const double *psi_i, double t, uint64_t r,
                              int max_iterations, double *out_r, double *out_i) {
    return guard([&] {
        if (!H) throw ArgError("hamiltonian_simulation_cuda: H is null");
        if (H->rows != H->cols) throw ArgError("hamiltonian_simulation_cuda: H must be square");
        const int64_t N = H->rows;
        if (N == 0) return;
        if (!psi_r || !psi_i || !out_r || !out_i) throw ArgError("hamiltonian_simulation_cuda: psi0 is null");
        const size_t eb = elem_bytes(H->precision);
        DevBuf dr(N * eb), di(N * eb);
        if (H->precision == DQ_F64) {
            DQ_HIP(hipMemcpy(dr.p, psi_r, N * 8, hipMemcpyHostToDevice));
            DQ_HIP(hipMemcpy(di.p, psi_i, N * 8, hipMemcpyHostToDevice));
        } else {
            std::vector<float> a(psi_r, psi_r + N), b(psi_i, psi_i + N);
            DQ_HIP(hipMemcpy(dr.p, a.data(), N * 4, hipMemcpyHostToDevice));
            DQ_HIP(hipMemcpy(di.p, b.data(), N * 4, hipMemcpyHostToDevice));
        }
        hamsim_device(*H, dr.p, di.p, t, r, max_iterations, 0);
        if (H->precision == DQ_F64) {
            DQ_HIP(hipMemcpy(out_r, dr.p, N * 8, hipMemcpyDeviceToHost));
            DQ_HIP(hipMemcpy(out_i, di.p, N * 8, hipMemcpyDeviceToHost));
        } else {
            std::vector<float> a(N), b(N);
            DQ_HIP(hipMemcpy(a.data(), dr.p, N * 4, hipMemcpyDeviceToHost));
            DQ_HIP(hipMemcpy(b.data(), di.p, N * 4, hipMemcpyDeviceToHost));
            std::copy(a.begin(), a.end(), out_r);
            std::copy(b.begin(), b.end(), out_i);
        }
    });
}
