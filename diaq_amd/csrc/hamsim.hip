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
// row sums in slot order, then the scaling, then the add. The term kernel may stage the x window of each row tile in
// LDS: the +-2^k offsets of a qubit Hamiltonian then share one copy of the term vector per tile. Its rows per lane
// and LDS window are timed on the first evolution of a matrix and kept in the handle.
//
// Each call first maps the 256-row blocks of every diagonal whose values are all zero (one read of H), and every term
// skips them: such a block adds nothing to a row sum (the products it would add are zeros, exact in FP64 for finite
// inputs), and a qubit Hamiltonian's flip diagonals are zero on alternate runs of 2^k rows (three runs in four for a
// flip-flop term). H cannot change during the call, so the map stays valid for all of its terms.
#include <algorithm>

#include "dia_rows.hpp"

namespace dq {

static int g_hamsim_skip = 1;  // "hamsim_skip_zero_blocks"
static int g_hamsim_rpt = 0;   // "hamsim_rows_per_lane": 0 = the matrix's timed choice, else 2, 4 or 8
static int g_hamsim_lds = -1;  // "hamsim_lds": -1 = the matrix's timed choice, else 0 or 1

bool hamsim_set_tuning(const std::string &key, int value) {
    if (key == "hamsim_skip_zero_blocks") {
        if (value != 0 && value != 1) throw ArgError("hamsim_skip_zero_blocks must be 0 or 1");
        g_hamsim_skip = value;
        return true;
    }
    if (key == "hamsim_rows_per_lane") {
        if (value != 0 && value != 2 && value != 4 && value != 8)
            throw ArgError("hamsim_rows_per_lane must be 0, 2, 4 or 8");
        g_hamsim_rpt = value;
        return true;
    }
    if (key == "hamsim_lds") {
        if (value < -1 || value > 1) throw ArgError("hamsim_lds must be -1, 0 or 1");
        g_hamsim_lds = value;
        return true;
    }
    return false;
}

// TO: the precision of the step output (and of psi, read by the first term). TO = T for an FP64 Hamiltonian. An FP32
// Hamiltonian (the precision form) runs its terms in FP32 but adds them into an FP64 output; the step's last term
// also writes the output rounded to FP32 (o32), the next step's FP32 input.
template <typename T, typename TO, int RPT, bool LDS, bool ZM>
__global__ __launch_bounds__(256) void taylor_term_kernel(int64_t rows, int n, const DiagMeta *__restrict__ meta,
                                                          const T *__restrict__ are, const T *__restrict__ aim,
                                                          const T *__restrict__ tr, const T *__restrict__ ti,
                                                          T *__restrict__ nr, T *__restrict__ ni,
                                                          TO *__restrict__ outr, TO *__restrict__ outi,
                                                          const TO *__restrict__ p0r, const TO *__restrict__ p0i,
                                                          T *__restrict__ o32r, T *__restrict__ o32i, double scale,
                                                          int first, int64_t ntiles, int xcd,
                                                          const uint32_t *__restrict__ zm, int64_t zwords) {
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
    dia_rows_tile<T, RPT, true, true, HALO, ZM>(tile, tile_end, n, meta, are, aim, tr, ti, 0, rows, hr, hi, sr, si,
                                                zm, zwords);
    const T s = static_cast<T>(scale);
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        const int64_t r = tile + threadIdx.x + 256 * j;
        if (r < rows) {
            const T a = hr[j].get(), b = hi[j].get();
            const T vr = b * s, vi = -a * s;  // (-i s)(a + ib)
            nr[r] = vr;
            ni[r] = vi;
            // out = psi + term_1 at the first term (psi is the term_0 input), out += term_k after
            const TO orr = (first ? p0r[r] : outr[r]) + static_cast<TO>(vr);
            const TO oii = (first ? p0i[r] : outi[r]) + static_cast<TO>(vi);
            outr[r] = orr;
            outi[r] = oii;
            if (o32r) {
                o32r[r] = static_cast<T>(orr);
                o32i[r] = static_cast<T>(oii);
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

template <typename TD, typename TS>
__global__ void convert2_kernel(int64_t n, const TS *__restrict__ ar, const TS *__restrict__ ai, TD *__restrict__ br,
                                TD *__restrict__ bi) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        br[i] = static_cast<TD>(ar[i]);
        bi[i] = static_cast<TD>(ai[i]);
    }
}

template <typename TD, typename TS>
static void convert2(int64_t n, const TS *ar, const TS *ai, TD *br, TD *bi, hipStream_t s) {
    hipLaunchKernelGGL((convert2_kernel<TD, TS>), dim3(blocks_for(n, 256, 8192)), dim3(256), 0, s, n, ar, ai, br, bi);
    DQ_LAUNCH_CHECK();
}

// One term: term_k = (-i scale) H in, out (+)= term_k, on tiles of 256 * RPT rows.
template <typename T, typename TO, int RPT>
static void launch_term(const dq_devmat &H, bool lds, const T *inr, const T *ini, T *nr, T *ni, TO *outr, TO *outi,
                        const TO *p0r, const TO *p0i, T *o32r, T *o32i, double scale, int first, const uint32_t *zm,
                        int64_t zwords, hipStream_t s) {
    const int64_t N = H.rows;
    const int64_t ntiles = (N + 256 * RPT - 1) / (256 * RPT);
    const int xcd = xcd_tiles_for(H);
    const dim3 grid(static_cast<unsigned>(grid_for_tiles(ntiles, xcd)));
    const DiagMeta *meta = H.d_meta.as<DiagMeta>();
    const T *are = static_cast<const T *>(H.re.p), *aim = static_cast<const T *>(H.im.p);
#define DQ_TERM(L, Z)                                                                                             \
    hipLaunchKernelGGL((taylor_term_kernel<T, TO, RPT, L, Z>), grid, dim3(256), 0, s, N, H.n(), meta, are, aim, inr, \
                       ini, nr, ni, outr, outi, p0r, p0i, o32r, o32i, scale, first, ntiles, xcd, zm, zwords)
    if (lds && zm) DQ_TERM(true, true);
    else if (lds) DQ_TERM(true, false);
    else if (zm) DQ_TERM(false, true);
    else DQ_TERM(false, false);
#undef DQ_TERM
    DQ_LAUNCH_CHECK();
}

template <typename T, typename TO>
using TermFn = void (*)(const dq_devmat &, bool, const T *, const T *, T *, T *, TO *, TO *, const TO *, const TO *,
                        T *, T *, double, int, const uint32_t *, int64_t, hipStream_t);

template <typename T, typename TO>
static TermFn<T, TO> term_fn(int rpt) {
    return rpt == 2 ? launch_term<T, TO, 2> : rpt == 4 ? launch_term<T, TO, 4> : launch_term<T, TO, 8>;
}

// The term kernel's schedule for H: rows per lane {2, 4, 8} x LDS window {off, on where some off-diagonal lies
// within half a tile}, each timed on two first terms (psi in, scratch out) the first time H is evolved, and kept in
// the handle. Every schedule adds the same products in the same order, so the choice changes the speed only.
template <typename T, typename TO>
static void term_schedule(const dq_devmat &H, const T *pr, const T *pi, const TO *p0r, const TO *p0i, TO *outr,
                          TO *outi, T *tr, T *ti, const uint32_t *zm, int64_t zwords, hipStream_t s, int &rpt,
                          bool &lds) {
    if (H.term_rpt == 0 && (g_hamsim_rpt == 0 || g_hamsim_lds < 0)) {
        hipEvent_t e0, e1;
        DQ_HIP(hipEventCreate(&e0));
        DQ_HIP(hipEventCreate(&e1));
        float best = 1e30f;
        int brpt = 8, blds = 0;
        for (int r : {2, 4, 8}) {
            for (int l = 0; l < 2; ++l) {
                if (l && !lds_window_pays(H, r)) continue;
                const TermFn<T, TO> f = term_fn<T, TO>(r);
                f(H, l == 1, pr, pi, tr, ti, outr, outi, p0r, p0i, nullptr, nullptr, 0.1, 1, zm, zwords, s);  // warm
                DQ_HIP(hipEventRecord(e0, s));
                for (int q = 0; q < 2; ++q)
                    f(H, l == 1, pr, pi, tr, ti, outr, outi, p0r, p0i, nullptr, nullptr, 0.1, 1, zm, zwords, s);
                DQ_HIP(hipEventRecord(e1, s));
                DQ_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                DQ_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) { best = ms; brpt = r; blds = l; }
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        H.term_lds = blds;  // before term_rpt, which marks the choice as made
        H.term_rpt = brpt;
    }
    rpt = g_hamsim_rpt ? g_hamsim_rpt : H.term_rpt.load();
    lds = g_hamsim_lds >= 0 ? g_hamsim_lds == 1 : H.term_lds.load() == 1;
}

// One step out = sum_{k=0}^{m} (-i dt H)^k psi / k! (src/spHamSim_gpu_all_fused.cu:235-277). (pr, pi): psi in the
// term precision, the first term's input; (p0r, p0i): psi in the output precision (the same arrays when T = TO).
// t0/t1: scratch terms. (o32r, o32i), for T != TO: receives the output rounded to T.
template <typename T, typename TO>
static void taylor_step(const dq_devmat &H, const T *pr, const T *pi, const TO *p0r, const TO *p0i, TO *outr,
                        TO *outi, T *o32r, T *o32i, T *t0r, T *t0i, T *t1r, T *t1i, double dt, int m,
                        TermFn<T, TO> term, bool lds, const uint32_t *zm, int64_t zwords, hipStream_t s) {
    const int64_t N = H.rows;
    if (m <= 0 || H.n() == 0) {  // out = psi (+ zero terms)
        hipLaunchKernelGGL(copy2_kernel<TO>, dim3(blocks_for(N, 256, 8192)), dim3(256), 0, s, N, p0r, p0i, outr,
                           outi);
        DQ_LAUNCH_CHECK();
        if (o32r) {
            hipLaunchKernelGGL(copy2_kernel<T>, dim3(blocks_for(N, 256, 8192)), dim3(256), 0, s, N, pr, pi, o32r,
                               o32i);
            DQ_LAUNCH_CHECK();
        }
        return;
    }
    const T *inr = pr, *ini = pi;
    T *bufr[2] = {t0r, t1r}, *bufi[2] = {t0i, t1i};
    for (int k = 1; k <= m; ++k) {
        T *nr = bufr[k & 1], *ni = bufi[k & 1];
        term(H, lds, inr, ini, nr, ni, outr, outi, p0r, p0i, k == m ? o32r : nullptr, k == m ? o32i : nullptr,
             dt / static_cast<double>(k), k == 1 ? 1 : 0, zm, zwords, s);
        inr = nr;
        ini = ni;
    }
}

// ||v||^2 of an FP64 complex vector as RENORM_PARTS block partials (each block sums a fixed, grid-strided share of
// the elements in a fixed order, so the partials -- and the fold of them -- have the same bits on every run).
constexpr int RENORM_PARTS = 512;

__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline double block_sum256(double v) {
    __shared__ double part[4];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    const double r = (part[0] + part[1]) + (part[2] + part[3]);
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(256) void sumsq_parts_kernel(int64_t n, const double *__restrict__ re,
                                                          const double *__restrict__ im, double *__restrict__ part) {
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        acc += re[i] * re[i] + im[i] * im[i];
    acc = block_sum256(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// psi *= ||psi0|| / ||psi||, both squared norms folded from their partials by every block in the same order.
__global__ __launch_bounds__(256) void renorm_kernel(int64_t n, double *__restrict__ re, double *__restrict__ im,
                                                     const double *__restrict__ part0,
                                                     const double *__restrict__ part1) {
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < RENORM_PARTS; i += 256) {
        a += part0[i];
        b += part1[i];
    }
    a = block_sum256(a);
    b = block_sum256(b);
    const double f = b > 0.0 ? sqrt(a / b) : 1.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        re[i] *= f;
        im[i] *= f;
    }
}

static void sumsq_parts(int64_t n, const double *re, const double *im, double *part, hipStream_t s) {
    hipLaunchKernelGGL(sumsq_parts_kernel, dim3(RENORM_PARTS), dim3(256), 0, s, n, re, im, part);
    DQ_LAUNCH_CHECK();
}

// The zero-block map of H (256-row blocks), read by every term of one evolution.
template <typename T>
static DevBuf zero_map(const dq_devmat &H, int m, int64_t zwords, hipStream_t s) {
    DevBuf zmap;
    if (g_hamsim_skip && H.n() > 0 && m > 0) {
        zmap = DevBuf(sizeof(uint32_t) * zwords * H.n());
        hipLaunchKernelGGL((zero_blocks_kernel<T, 256>), dim3(static_cast<unsigned>(zwords), zero_map_grid_y(H.n()), 1),
                           dim3(256), 0, s, (int64_t)H.rows, H.n(), H.d_meta.as<DiagMeta>(), int64_t(0),
                           static_cast<const T *>(H.re.p), static_cast<const T *>(H.im.p), zwords,
                           zmap.as<uint32_t>());
        DQ_LAUNCH_CHECK();
    }
    return zmap;
}

// Evolve the device state (pr, pi) in place over r steps of dt = t / r, all in H's precision.
template <typename T>
static void evolve(const dq_devmat &H, T *pr, T *pi, double t, uint64_t r, int m, hipStream_t s) {
    const int64_t N = H.rows;
    if (r == 0 || N == 0) return;
    const double dt = t / static_cast<double>(r);
    DevBuf br(N * sizeof(T)), bi(N * sizeof(T)), t0r(N * sizeof(T)), t0i(N * sizeof(T)), t1r(N * sizeof(T)),
        t1i(N * sizeof(T));
    const int64_t zwords = zero_block_words(N, 256);
    DevBuf zmap = zero_map<T>(H, m, zwords, s);
    const uint32_t *zm = zmap.p ? zmap.as<uint32_t>() : nullptr;
    T *ar = pr, *ai = pi, *xr = br.as<T>(), *xi = bi.as<T>();
    int rpt = 8;
    bool lds = false;
    if (m > 0 && H.n() > 0)
        term_schedule<T, T>(H, pr, pi, pr, pi, xr, xi, t0r.as<T>(), t0i.as<T>(), zm, zwords, s, rpt, lds);
    const TermFn<T, T> term = term_fn<T, T>(rpt);
    for (uint64_t step = 0; step < r; ++step) {
        taylor_step<T, T>(H, ar, ai, ar, ai, xr, xi, nullptr, nullptr, t0r.as<T>(), t0i.as<T>(), t1r.as<T>(),
                          t1i.as<T>(), dt, m, term, lds, zm, zwords, s);
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

// The precision form for an FP32 H: the state (pr, pi) stays FP64. Each step's terms run in FP32 from an FP32 copy
// of the state; they are added into an FP64 output, which the step's last term also rounds to FP32 for the next
// step. At the end the state is renormalised in FP64 to the norm it started with (the exact evolution is unitary).
static void evolve_mixed(const dq_devmat &H, double *pr, double *pi, double t, uint64_t r, int m, hipStream_t s) {
    const int64_t N = H.rows;
    if (r == 0 || N == 0) return;
    const double dt = t / static_cast<double>(r);
    const size_t b32 = N * sizeof(float), b64 = N * sizeof(double);
    DevBuf xr(b64), xi(b64), p32r(b32), p32i(b32), q32r(b32), q32i(b32), t0r(b32), t0i(b32), t1r(b32), t1i(b32);
    DevBuf parts(2 * RENORM_PARTS * sizeof(double));
    double *part0 = parts.as<double>(), *part1 = part0 + RENORM_PARTS;
    sumsq_parts(N, pr, pi, part0, s);
    convert2<float, double>(N, pr, pi, p32r.as<float>(), p32i.as<float>(), s);
    const int64_t zwords = zero_block_words(N, 256);
    DevBuf zmap = zero_map<float>(H, m, zwords, s);
    const uint32_t *zm = zmap.p ? zmap.as<uint32_t>() : nullptr;
    double *ar = pr, *ai = pi, *or_ = xr.as<double>(), *oi = xi.as<double>();
    float *in_r = p32r.as<float>(), *in_i = p32i.as<float>(), *nx_r = q32r.as<float>(), *nx_i = q32i.as<float>();
    int rpt = 8;
    bool lds = false;
    if (m > 0 && H.n() > 0)
        term_schedule<float, double>(H, in_r, in_i, pr, pi, or_, oi, t0r.as<float>(), t0i.as<float>(), zm, zwords, s,
                                     rpt, lds);
    const TermFn<float, double> term = term_fn<float, double>(rpt);
    for (uint64_t step = 0; step < r; ++step) {
        taylor_step<float, double>(H, in_r, in_i, ar, ai, or_, oi, nx_r, nx_i, t0r.as<float>(), t0i.as<float>(),
                                   t1r.as<float>(), t1i.as<float>(), dt, m, term, lds, zm, zwords, s);
        std::swap(ar, or_);
        std::swap(ai, oi);
        std::swap(in_r, nx_r);
        std::swap(in_i, nx_i);
    }
    if (ar != pr) {
        hipLaunchKernelGGL(copy2_kernel<double>, dim3(blocks_for(N, 256, 8192)), dim3(256), 0, s, N,
                           (const double *)ar, (const double *)ai, pr, pi);
        DQ_LAUNCH_CHECK();
    }
    sumsq_parts(N, pr, pi, part1, s);
    hipLaunchKernelGGL(renorm_kernel, dim3(blocks_for(N, 256, 8192)), dim3(256), 0, s, N, pr, pi,
                       (const double *)part0, (const double *)part1);
    DQ_LAUNCH_CHECK();
    DQ_HIP(hipStreamSynchronize(s));  // scratch buffers are freed on return
}

// An FP32 H on an FP32 state: the state is widened to FP64 for the evolution and rounded back once at the end.
static void evolve_fp32_state(const dq_devmat &H, float *pr, float *pi, double t, uint64_t r, int m, hipStream_t s) {
    const int64_t N = H.rows;
    if (r == 0 || N == 0) return;
    DevBuf wr(N * sizeof(double)), wi(N * sizeof(double));
    convert2<double, float>(N, pr, pi, wr.as<double>(), wi.as<double>(), s);
    evolve_mixed(H, wr.as<double>(), wi.as<double>(), t, r, m, s);
    convert2<float, double>(N, wr.as<double>(), wi.as<double>(), pr, pi, s);
    DQ_HIP(hipStreamSynchronize(s));
}

void hamsim_device(const dq_devmat &H, void *pr, void *pi, double t, uint64_t r, int m, hipStream_t s) {
    if (H.rows != H.cols) throw ArgError("hamiltonian_simulation_cuda: H must be square");
    if (H.precision == DQ_F64)
        evolve<double>(H, static_cast<double *>(pr), static_cast<double *>(pi), t, r, m, s);
    else
        evolve_fp32_state(H, static_cast<float *>(pr), static_cast<float *>(pi), t, r, m, s);
}

void hamsim_device_f64(const dq_devmat &H, double *pr, double *pi, double t, uint64_t r, int m, hipStream_t s) {
    if (H.rows != H.cols) throw ArgError("hamiltonian_simulation_cuda: H must be square");
    if (H.precision == DQ_F64)
        evolve<double>(H, pr, pi, t, r, m, s);
    else
        evolve_mixed(H, pr, pi, t, r, m, s);
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

extern "C" int dq_hamsim_device_f64state(const dq_devmat *H, double *pr, double *pi, double t, uint64_t r, int m,
                                         uintptr_t stream) {
    return guard([&] {
        if (!H) throw ArgError("hamiltonian_simulation_cuda: H is null");
        if (H->rows && (!pr || !pi)) throw ArgError("hamiltonian_simulation_cuda: psi0 is null");
        hamsim_device_f64(*H, pr, pi, t, r, m, as_stream(stream));
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
        DevBuf dr(N * 8), di(N * 8);
        DQ_HIP(hipMemcpy(dr.p, psi_r, N * 8, hipMemcpyHostToDevice));
        DQ_HIP(hipMemcpy(di.p, psi_i, N * 8, hipMemcpyHostToDevice));
        hamsim_device_f64(*H, dr.as<double>(), di.as<double>(), t, r, max_iterations, 0);
        DQ_HIP(hipMemcpy(out_r, dr.p, N * 8, hipMemcpyDeviceToHost));
        DQ_HIP(hipMemcpy(out_i, di.p, N * 8, hipMemcpyDeviceToHost));
    });
}
