// Gates on a state vector without forming the n-qubit matrix: y (+)= (I_{2^before} (x) G (x) I_{2^after}) x for a
// square DIA gate G of size g (src/spKron.cpp:286-335, 433-496, 498-571 on the host).
//
// Index the state as x[b][c][a] (b < 2^before, c < g, a < 2^after): the operator mixes only the g entries of one
// column (b, a), y[b][r][a] = sum over G's diagonals d of G[r][r + d] x[b][r + d][a].
//
// Gates up to 16 x 16 (up to four qubits): one lane per column (b, a). It loads the column's g entries once into
// registers, then adds each diagonal's terms in slot order -- the reference's order -- and stores the g results.
// Neighbouring lanes own neighbouring a, so every load and store is coalesced over a when 2^after >= 64, and x and y
// each cross HBM once. Because a lane reads its whole column before it writes, x and y may be the same vector.
// Larger gates: one lane per output element, looping over the slots (x re-read through the caches).
#include <algorithm>
#include <cmath>

#include "dq_common.hpp"
#include "kron_core.hpp"

namespace dq {

constexpr int KRON_REG_MAX = 16;     // largest gate of the register path
constexpr int KRON_MAX_SLOTS = 64;   // slots staged in LDS on the register path

template <typename T, int GM>
__global__ __launch_bounds__(256) void kron_apply_reg_kernel(int64_t ncols, int after, int g, int nd,
                                                             const DiagMeta *__restrict__ meta,
                                                             const T *__restrict__ gre, const T *__restrict__ gim,
                                                             const T *xr_, const T *xi_, T *yr_, T *yi_,
                                                             int accumulate) {
    __shared__ T vr[KRON_MAX_SLOTS][GM], vi[KRON_MAX_SLOTS][GM];
    __shared__ int soff[KRON_MAX_SLOTS], slo[KRON_MAX_SLOTS], shi[KRON_MAX_SLOTS];
    // stage the gate: slot k's value at row r (zero outside the diagonal)
    for (int t = threadIdx.x; t < nd * GM; t += blockDim.x) {
        const int k = t / GM, r = t % GM;
        const DiagMeta m = meta[k];
        const bool on = r >= m.rlo && r < m.rhi;
        vr[k][r] = on ? gre[m.rowbase + r] : T(0);
        vi[k][r] = on ? gim[m.rowbase + r] : T(0);
        if (r == 0) {
            soff[k] = m.off;
            slo[k] = static_cast<int>(m.rlo);
            shi[k] = static_cast<int>(m.rhi);
        }
    }
    __syncthreads();
    const int64_t A = int64_t(1) << after;
    for (int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; q < ncols;
         q += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t b = q >> after, a = q & (A - 1);
        const int64_t base = b * g * A + a;
        T xr[GM], xi[GM];
#pragma unroll
        for (int c = 0; c < GM; ++c) {
            xr[c] = c < g ? xr_[base + c * A] : T(0);
            xi[c] = c < g ? xi_[base + c * A] : T(0);
        }
        Acc<T> ar[GM], ai[GM];
        if (accumulate) {
#pragma unroll
            for (int r = 0; r < GM; ++r) {
                if (r < g) {
                    ar[r].s = yr_[base + r * A];
                    ai[r].s = yi_[base + r * A];
                }
            }
        }
        for (int k = 0; k < nd; ++k)
            kron_diag_dispatch<T, GM, -(GM - 1)>(soff[k], g, slo[k], shi[k], vr[k], vi[k], xr, xi, ar, ai);
#pragma unroll
        for (int r = 0; r < GM; ++r) {
            if (r < g) {
                yr_[base + r * A] = ar[r].get();
                yi_[base + r * A] = ai[r].get();
            }
        }
    }
}

// Any gate: one lane per output element y[b][r][a], slots in order.
template <typename T>
__global__ __launch_bounds__(256) void kron_apply_elem_kernel(int64_t n, int after, int g, int nd,
                                                              const DiagMeta *__restrict__ meta,
                                                              const T *__restrict__ gre, const T *__restrict__ gim,
                                                              const T *__restrict__ xr_, const T *__restrict__ xi_,
                                                              T *__restrict__ yr_, T *__restrict__ yi_,
                                                              int accumulate) {
    const int64_t A = int64_t(1) << after;
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t a = e & (A - 1), br = e >> after;
        const int64_t r = br % g, colbase = (br - r) * A + a;  // x[b][c][a] = colbase + c * A
        Acc<T> sr, si;
        if (accumulate) {
            sr.s = yr_[e];
            si.s = yi_[e];
        }
        for (int k = 0; k < nd; ++k) {
            const DiagMeta m = meta[k];
            if (r >= m.rlo && r < m.rhi) {
                const int64_t c = r + m.off;
                cmac(sr, si, gre[m.rowbase + r], gim[m.rowbase + r], xr_[colbase + c * A], xi_[colbase + c * A]);
            }
        }
        yr_[e] = sr.get();
        yi_[e] = si.get();
    }
}

template <typename T>
static void kron_launch(const dq_devmat &G, int before, int after, const void *xr, const void *xi, void *yr,
                        void *yi, int accumulate, hipStream_t s) {
    const int g = static_cast<int>(G.rows), nd = G.n();
    const DiagMeta *meta = G.d_meta.as<DiagMeta>();
    const T *gre = G.re.as<T>(), *gim = G.im.as<T>();
    const T *x_r = static_cast<const T *>(xr), *x_i = static_cast<const T *>(xi);
    T *y_r = static_cast<T *>(yr), *y_i = static_cast<T *>(yi);
    const int64_t ncols = int64_t(1) << (before + after);
    if (g <= KRON_REG_MAX && nd <= KRON_MAX_SLOTS) {
        const unsigned grid = blocks_for(ncols, 256, 1 << 16);
#define DQ_KRON_REG(GM)                                                                                       \
    hipLaunchKernelGGL((kron_apply_reg_kernel<T, GM>), dim3(grid), dim3(256), 0, s, ncols, after, g, nd, meta, gre, \
                       gim, x_r, x_i, y_r, y_i, accumulate)
        if (g <= 2) {
            DQ_KRON_REG(2);
        } else if (g <= 4) {
            DQ_KRON_REG(4);
        } else if (g <= 8) {
            DQ_KRON_REG(8);
        } else {
            DQ_KRON_REG(16);
        }
#undef DQ_KRON_REG
    } else {
        const int64_t n = ncols * g;
        hipLaunchKernelGGL(kron_apply_elem_kernel<T>, dim3(blocks_for(n, 256, 1 << 16)), dim3(256), 0, s, n, after, g,
                           nd, meta, gre, gim, x_r, x_i, y_r, y_i, accumulate);
    }
    DQ_LAUNCH_CHECK();
}

// Shape checks shared by the entry points; returns the vector length 2^before * g * 2^after.
static int64_t kron_check(const dq_devmat *G, uint32_t before, uint32_t after) {
    if (!G) throw ArgError("gate is NULL");
    if (G->rows != G->cols) throw ArgError("kron_apply: the gate must be square");
    if (G->rows == 0) throw ArgError("kron_apply: empty gate");
    if (before > 40 || after > 40 || before + after > 40)
        throw ArgError("kron_apply: before + after must be <= 40 qubits");
    const double len = std::ldexp(static_cast<double>(G->rows), static_cast<int>(before + after));
    if (len > 9.0e15) throw ArgError("kron_apply: vector too long");
    return static_cast<int64_t>(G->rows) << (before + after);
}

static bool kron_in_place_ok(const dq_devmat &G) { return G.rows <= KRON_REG_MAX && G.n() <= KRON_MAX_SLOTS; }

void kron_apply(const dq_devmat &G, int before, int after, const void *xr, const void *xi, void *yr, void *yi,
                bool accumulate, hipStream_t s) {
    if (G.precision == DQ_F64) kron_launch<double>(G, before, after, xr, xi, yr, yi, accumulate, s);
    else kron_launch<float>(G, before, after, xr, xi, yr, yi, accumulate, s);
}

}  // namespace dq

using namespace dq;

extern "C" int dq_kron_apply_device(const dq_devmat *G, uint32_t before, uint32_t after, const void *xr,
                                    const void *xi, void *yr, void *yi, int accumulate, uintptr_t stream) {
    return guard([&] {
        kron_check(G, before, after);
        if (!xr || !xi || !yr || !yi) throw ArgError("kron_apply: NULL vector");
        if ((xr == yr || xi == yi) && !kron_in_place_ok(*G))
            throw ArgError("kron_apply: in place (x == y) needs a gate of at most 16 x 16 with at most 64 slots");
        hipStream_t s = as_stream(stream);
        if (G->precision == DQ_F64) kron_launch<double>(*G, before, after, xr, xi, yr, yi, accumulate != 0, s);
        else kron_launch<float>(*G, before, after, xr, xi, yr, yi, accumulate != 0, s);
    });
}

extern "C" int dq_kron_apply_host(const dq_devmat *G, uint32_t before, uint32_t after, const double *xr,
                                  const double *xi, double *yr, double *yi) {
    return guard([&] {
        const int64_t n = kron_check(G, before, after);
        if (G->precision != DQ_F64) throw ArgError("kron_apply_host: the gate must be FP64 (host vectors are FP64)");
        if (!xr || !xi || !yr || !yi) throw ArgError("kron_apply: NULL vector");
        hipStream_t s = nullptr;
        const size_t nb = static_cast<size_t>(n) * sizeof(double);
        DevBuf dxr(nb), dxi(nb), dyr(nb), dyi(nb);
        DQ_HIP(hipMemcpyAsync(dxr.p, xr, nb, hipMemcpyHostToDevice, s));
        DQ_HIP(hipMemcpyAsync(dxi.p, xi, nb, hipMemcpyHostToDevice, s));
        DQ_HIP(hipMemcpyAsync(dyr.p, yr, nb, hipMemcpyHostToDevice, s));
        DQ_HIP(hipMemcpyAsync(dyi.p, yi, nb, hipMemcpyHostToDevice, s));
        kron_launch<double>(*G, before, after, dxr.p, dxi.p, dyr.p, dyi.p, 1, s);
        DQ_HIP(hipMemcpyAsync(yr, dyr.p, nb, hipMemcpyDeviceToHost, s));
        DQ_HIP(hipMemcpyAsync(yi, dyi.p, nb, hipMemcpyDeviceToHost, s));
        DQ_HIP(hipStreamSynchronize(s));
    });
}
