// Batched time evolution on gfx950: many Hamiltonians, several states each, evolved to their own times in one
// launch of a HIP graph. No reference counterpart (the reference evolves one state of one Hamiltonian per call,
// src/spHamSim_gpu_all_fused.cu:287-376); the arithmetic per sub-step is the Taylor action of that file.
//
// Plan: jobs (Hamiltonians) of the same size and diagonal set form a class; a class stores its Hamiltonians' values
// back to back in one layout (shared metadata), in FP32 or FP64, and its states row-major with the states of a row
// adjacent ([job][row][state]) so the NS lanes of a row share each matrix value and read NS adjacent vector entries.
//
// Job j is evolved to t_j by 2^{s_j} sub-steps of tau_j = t_j / 2^{s_j}, each a degree-m_j Taylor polynomial of
// Y = -i tau_j H_j evaluated by Horner's rule:
//     b_m = psi;  b_k = psi + (Y / (k+1)) b_{k+1}  (k = m-1 .. 1);  psi <- psi + Y b_1,
// one SpMV per level. The state psi is kept in FP64; the Horner vectors b_k and an FP32 copy of psi are FP32, and the
// row sums are Kahan-compensated in FP32 (or plain FP64 when the values are FP64). Only the correction Y b_1 -- small
// next to psi -- carries FP32 rounding, so the error per sub-step is FP32 eps times ||Y b_1||, not times ||psi||.
//
// Launch g of a class advances every job of the class by one Horner level (jobs that are done return at once); the
// class needs max_j m_j 2^{s_j} launches, captured once into a graph with one branch per class on its own stream.
// A final FP64 pass renormalises every state.
//
// A-posteriori check: the probe runs the forward Taylor terms t_k = (Y/k) t_{k-1} of the first sub-step from the
// initial states up to k = m_j + 1 and records ||t_{m_j}|| and ||t_{m_j+1}||; the estimate of the total truncation
// error is 2^{s_j} ||t_{m+1}|| / (1 - ||t_{m+1}|| / ||t_m||) (a geometric tail), as in expmv-style stopping tests.
#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>
#include <vector>

#include "dia_rows.hpp"

namespace dq {

struct JobParam {
    double tau;      // sub-step length t / 2^s
    int32_t m;       // Taylor degree
    int32_t s;       // squaring count: 2^s sub-steps
    int64_t steps;   // m * 2^s Horner levels
};

template <typename TV>
struct AccT {
    using type = Acc<TV>;
};

// One Horner level for every job of a class. blockIdx.y = job in class; tiles of 256/NS rows x RPT.
// LDSW: the workgroup first stages the input rows [tile - T/2, tile + 3T/2) (T rows per tile, all NS states, zero
// outside the matrix) in LDS, and the diagonals with |offset| <= T/2 gather from there: the +-2^k offsets of a qubit
// Hamiltonian then share one copy of the 8-state vectors per tile.
// zmask (NULL = off): bit b of word [job][diagonal][b / 32] is clear when every value of the diagonal in rows
// [b * 256/NS, (b + 1) * 256/NS) is zero. Such a block adds nothing, so its value loads and x gathers are skipped;
// a qubit Hamiltonian's +-2^k flip diagonals are zero on alternate runs of 2^k rows, half of what they store.
template <typename TV, int NS, int RPT, bool LDSW>
__global__ __launch_bounds__(256) void horner_kernel(int64_t N, int n, const DiagMeta *__restrict__ meta,
                                                     int64_t total, const TV *__restrict__ hre,
                                                     const TV *__restrict__ him, const JobParam *__restrict__ params,
                                                     int64_t g, double *__restrict__ p64r, double *__restrict__ p64i,
                                                     float *__restrict__ p32r, float *__restrict__ p32i,
                                                     const float *__restrict__ bin_r, const float *__restrict__ bin_i,
                                                     float *__restrict__ bout_r, float *__restrict__ bout_i,
                                                     int64_t ntiles, int xcd, const uint32_t *__restrict__ zmask,
                                                     int64_t zwords) {
    const int j = blockIdx.y;
    const JobParam P = params[j];
    if (g >= P.steps) return;
    const int64_t t = tile_of_block(blockIdx.x, ntiles, xcd);
    if (t >= ntiles) return;  // uniform: a surplus workgroup of the XCD-contiguous grid
    const int m = P.m;
    const int k = m - 1 - static_cast<int>(g % m);  // Horner level
    constexpr int RPP = 256 / NS;                   // rows per pass of the workgroup
    const int st = threadIdx.x % NS;
    const int rl = threadIdx.x / NS;
    const int64_t tile = t * RPP * RPT;
    const int64_t vb = static_cast<int64_t>(j) * total;
    const int64_t sb = static_cast<int64_t>(j) * N * NS;
    // gathers always come from the previous level's output: b_{k+1}, or at the top level the FP32 state written by
    // the previous sub-step's last level (or by the run's prologue); psi's own entries are read from p32 / p64
    const float *xr = bin_r + sb;
    const float *xi = bin_i + sb;
    constexpr int T = RPP * RPT;
    constexpr int H = LDSW ? T / 2 : 0;
    constexpr int WE = LDSW ? (T + 2 * H) * NS : 1;  // staged elements per array
    __shared__ float sxr[WE], sxi[WE];
    if constexpr (LDSW) {
        const int64_t ws = (tile - H) * NS;  // first staged element (row tile - H, state 0)
        for (int e = threadIdx.x; e < WE; e += 256) {
            const int64_t gi = ws + e;
            float a = 0.f, b = 0.f;
            if (gi >= 0 && gi < N * NS) {
                a = xr[gi];
                b = xi[gi];
            }
            sxr[e] = a;
            sxi[e] = b;
        }
        __syncthreads();
    }
    Acc<TV> hr[RPT], hi[RPT];
    for (int d = 0; d < n; ++d) {
        const DiagMeta md = meta[d];
        const int64_t lo = md.rlo, hi_ = md.rhi;
        if (hi_ <= tile || lo >= tile + T) continue;  // uniform
        // pass q of the tile covers block t*RPT + q of RPP rows; RPT divides 32, so the tile's bits share one word
        uint32_t live = (1u << RPT) - 1;
        if (zmask) {
            const uint32_t w = zmask[(static_cast<int64_t>(j) * n + d) * zwords + ((t * RPT) >> 5)];
            live = (w >> ((t * RPT) & 31)) & ((1u << RPT) - 1);
            if (!live) continue;  // uniform: every value of the diagonal in this tile is zero
        }
        const TV *ar = hre + vb + md.rowbase;
        const TV *ai = him + vb + md.rowbase;
        TV va_r[RPT] = {}, va_i[RPT] = {}, vx_r[RPT] = {}, vx_i[RPT] = {};
        if (LDSW && md.off >= -H && md.off <= H) {  // uniform: gather from the staged window
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                if (!((live >> q) & 1u)) continue;  // uniform
                const int64_t r = tile + rl + RPP * q;
                const int64_t rc = r < lo ? lo : (r >= hi_ ? hi_ - 1 : r);  // stays inside the tile
                va_r[q] = ar[rc];
                va_i[q] = ai[rc];
                const int e = (static_cast<int>(rc - tile) + H + md.off) * NS + st;
                vx_r[q] = static_cast<TV>(sxr[e]);
                vx_i[q] = static_cast<TV>(sxi[e]);
            }
        } else {
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                if (!((live >> q) & 1u)) continue;  // uniform
                const int64_t r = tile + rl + RPP * q;
                const int64_t rc = r < lo ? lo : (r >= hi_ ? hi_ - 1 : r);
                va_r[q] = ar[rc];
                va_i[q] = ai[rc];
                const int64_t xo = (rc + md.off) * NS + st;
                vx_r[q] = static_cast<TV>(xr[xo]);
                vx_i[q] = static_cast<TV>(xi[xo]);
            }
        }
#pragma unroll
        for (int q = 0; q < RPT; ++q) {  // rows off the diagonal add nothing (a compensated sum is not inert to zeros)
            const int64_t r = tile + rl + RPP * q;
            if (((live >> q) & 1u) && r >= lo && r < hi_) cmac(hr[q], hi[q], va_r[q], va_i[q], vx_r[q], vx_i[q]);
        }
    }
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int64_t r = tile + rl + RPP * q;
        if (r >= N) continue;
        const int64_t o = sb + r * NS + st;
        const double a = static_cast<double>(hr[q].get()), b = static_cast<double>(hi[q].get());
        if (k > 0) {  // b_k = psi + (Y/(k+1)) b_{k+1}, (Y/c) h = (tau/c) (Im h, -Re h)
            const double c = P.tau / static_cast<double>(k + 1);
            bout_r[o] = static_cast<float>(static_cast<double>(p32r[o]) + c * b);
            bout_i[o] = static_cast<float>(static_cast<double>(p32i[o]) - c * a);
        } else {  // psi <- psi + Y b_1 in FP64; its FP32 copy goes to p32 (own entries) and to the next level's input
            const double nr = p64r[o] + P.tau * b;
            const double ni = p64i[o] - P.tau * a;
            p64r[o] = nr;
            p64i[o] = ni;
            p32r[o] = static_cast<float>(nr);
            p32i[o] = static_cast<float>(ni);
            bout_r[o] = static_cast<float>(nr);
            bout_i[o] = static_cast<float>(ni);
        }
    }
}

// The same Horner level with four states per lane (NS = 4, 8 or 16: NS / 4 lanes per row): each diagonal's gather is
// one 16-byte load of 4 adjacent states' real parts and one of their imaginary parts, and a lane's matrix value serves
// its 4 states, so a wave issues a quarter of the vector memory instructions of horner_kernel for the same bytes. A
// wave covers 256 / NS rows per pass -- exactly one block of the zero-block map -- and the workgroup RPT passes of
// 1024 / NS rows. Each (row, state) sums its diagonals in the same order with the same operations as horner_kernel,
// so the two kernels give the same bits.
template <typename TV, int NS, int RPT, int DU = 1>
__global__ __launch_bounds__(256) void horner4_kernel(int64_t N, int n, const DiagMeta *__restrict__ meta,
                                                      int64_t total, const TV *__restrict__ hre,
                                                      const TV *__restrict__ him, const JobParam *__restrict__ params,
                                                      int64_t g, double *__restrict__ p64r, double *__restrict__ p64i,
                                                      float *__restrict__ p32r, float *__restrict__ p32i,
                                                      const float *__restrict__ bin_r, const float *__restrict__ bin_i,
                                                      float *__restrict__ bout_r, float *__restrict__ bout_i,
                                                      int64_t ntiles, int xcd, const uint32_t *__restrict__ zmask,
                                                      int64_t zwords) {
    static_assert(NS == 4 || NS == 8 || NS == 16, "four states per lane");
    constexpr int L = NS / 4;         // lanes per row
    constexpr int WR = 64 / L;        // rows per wave and pass: one zero-map block (256 / NS rows)
    constexpr int RPP = 4 * WR;       // rows per workgroup pass
    constexpr int T = RPP * RPT;      // rows per tile: 4 RPT blocks, inside one 32-bit word of the map
    const int j = blockIdx.y;
    const JobParam P = params[j];
    if (g >= P.steps) return;
    const int64_t t = tile_of_block(blockIdx.x, ntiles, xcd);
    if (t >= ntiles) return;  // uniform: a surplus workgroup of the XCD-contiguous grid
    const int m = P.m;
    const int k = m - 1 - static_cast<int>(g % m);  // Horner level
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int h = lane % L;                          // states 4h .. 4h + 3 of the row
    const int rw = lane / L;
    const int64_t tile = t * T;
    const int64_t vb = static_cast<int64_t>(j) * total;
    const int64_t sb = static_cast<int64_t>(j) * N * NS;
    const float *xr = bin_r + sb;
    const float *xi = bin_i + sb;
    const int64_t blk0 = t * 4 * RPT;  // the tile's first map block
    Acc<TV> hr[RPT][4], hi[RPT][4];
    // DU diagonals per batch: their loads are all issued before the first of them is used (more loads in flight per
    // wave); the products are still added diagonal after diagonal, in slot order
    for (int d0 = 0; d0 < n; d0 += DU) {
        TV va_r[DU][RPT], va_i[DU][RPT];
        float4 vx_r[DU][RPT], vx_i[DU][RPT];
        uint32_t lv[DU];
        int64_t lo_[DU], hi_[DU];
#pragma unroll
        for (int u = 0; u < DU; ++u) {
            lv[u] = 0;
            lo_[u] = 0;
            hi_[u] = 0;
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                va_r[u][q] = TV(0);
                va_i[u][q] = TV(0);
                vx_r[u][q] = make_float4(0.f, 0.f, 0.f, 0.f);
                vx_i[u][q] = vx_r[u][q];
            }
            const int d = d0 + u;
            if (d >= n) continue;  // uniform
            const DiagMeta md = meta[d];
            const int64_t rlo = md.rlo, rhi = md.rhi;
            if (rhi <= tile || rlo >= tile + T) continue;  // uniform
            uint32_t live = (4 * RPT >= 32) ? ~0u : ((1u << (4 * RPT)) - 1);
            if (zmask) {
                const uint32_t w = zmask[(static_cast<int64_t>(j) * n + d) * zwords + (blk0 >> 5)];
                live = (4 * RPT >= 32) ? w : ((w >> (blk0 & 31)) & ((1u << (4 * RPT)) - 1));
                if (!live) continue;  // uniform: every value of the diagonal in this tile is zero
            }
            lv[u] = live;
            lo_[u] = rlo;
            hi_[u] = rhi;
            const TV *ar = hre + vb + md.rowbase;
            const TV *ai = him + vb + md.rowbase;
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                if (!((live >> (4 * q + wave)) & 1u)) continue;  // uniform in the wave
                const int64_t r = tile + RPP * q + WR * wave + rw;
                const int64_t rc = r < rlo ? rlo : (r >= rhi ? rhi - 1 : r);  // stays inside the tile
                va_r[u][q] = ar[rc];
                va_i[u][q] = ai[rc];
                const int64_t xo = (rc + md.off) * NS + 4 * h;
                vx_r[u][q] = *reinterpret_cast<const float4 *>(xr + xo);
                vx_i[u][q] = *reinterpret_cast<const float4 *>(xi + xo);
            }
        }
#pragma unroll
        for (int u = 0; u < DU; ++u) {
            // rows off the diagonal add nothing (a compensated sum is not inert to zeros)
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const int64_t r = tile + RPP * q + WR * wave + rw;
                if (((lv[u] >> (4 * q + wave)) & 1u) && r >= lo_[u] && r < hi_[u]) {
                    cmac(hr[q][0], hi[q][0], va_r[u][q], va_i[u][q], static_cast<TV>(vx_r[u][q].x),
                         static_cast<TV>(vx_i[u][q].x));
                    cmac(hr[q][1], hi[q][1], va_r[u][q], va_i[u][q], static_cast<TV>(vx_r[u][q].y),
                         static_cast<TV>(vx_i[u][q].y));
                    cmac(hr[q][2], hi[q][2], va_r[u][q], va_i[u][q], static_cast<TV>(vx_r[u][q].z),
                         static_cast<TV>(vx_i[u][q].z));
                    cmac(hr[q][3], hi[q][3], va_r[u][q], va_i[u][q], static_cast<TV>(vx_r[u][q].w),
                         static_cast<TV>(vx_i[u][q].w));
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int64_t r = tile + RPP * q + WR * wave + rw;
        if (r >= N) continue;
        const int64_t o = sb + r * NS + 4 * h;
        double a[4], b[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            a[v] = static_cast<double>(hr[q][v].get());
            b[v] = static_cast<double>(hi[q][v].get());
        }
        if (k > 0) {  // b_k = psi + (Y/(k+1)) b_{k+1}, (Y/c) h = (tau/c) (Im h, -Re h)
            const double c = P.tau / static_cast<double>(k + 1);
            const float4 pr = *reinterpret_cast<const float4 *>(p32r + o);
            const float4 pi = *reinterpret_cast<const float4 *>(p32i + o);
            float4 outr, outi;
            outr.x = static_cast<float>(static_cast<double>(pr.x) + c * b[0]);
            outr.y = static_cast<float>(static_cast<double>(pr.y) + c * b[1]);
            outr.z = static_cast<float>(static_cast<double>(pr.z) + c * b[2]);
            outr.w = static_cast<float>(static_cast<double>(pr.w) + c * b[3]);
            outi.x = static_cast<float>(static_cast<double>(pi.x) - c * a[0]);
            outi.y = static_cast<float>(static_cast<double>(pi.y) - c * a[1]);
            outi.z = static_cast<float>(static_cast<double>(pi.z) - c * a[2]);
            outi.w = static_cast<float>(static_cast<double>(pi.w) - c * a[3]);
            *reinterpret_cast<float4 *>(bout_r + o) = outr;
            *reinterpret_cast<float4 *>(bout_i + o) = outi;
        } else {  // psi <- psi + Y b_1 in FP64; its FP32 copy goes to p32 (own entries) and to the next level's input
            float nr32[4], ni32[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const double nr = p64r[o + v] + P.tau * b[v];
                const double ni = p64i[o + v] - P.tau * a[v];
                p64r[o + v] = nr;
                p64i[o + v] = ni;
                nr32[v] = static_cast<float>(nr);
                ni32[v] = static_cast<float>(ni);
            }
            const float4 fr = make_float4(nr32[0], nr32[1], nr32[2], nr32[3]);
            const float4 fi = make_float4(ni32[0], ni32[1], ni32[2], ni32[3]);
            *reinterpret_cast<float4 *>(p32r + o) = fr;
            *reinterpret_cast<float4 *>(p32i + o) = fi;
            *reinterpret_cast<float4 *>(bout_r + o) = fr;
            *reinterpret_cast<float4 *>(bout_i + o) = fi;
        }
    }
}

// Forward Taylor terms of the first sub-step for the a-posteriori estimate: t_k = (Y/k) t_{k-1}, t_0 = psi (FP32);
// the squared norms of t_m and t_{m+1} of every state are accumulated into norms[(j*NS + state)*2 + (k - m)].
template <typename TV, int NS, int RPT>
__global__ __launch_bounds__(256) void probe_kernel(int64_t N, int n, const DiagMeta *__restrict__ meta, int64_t total,
                                                    const TV *__restrict__ hre, const TV *__restrict__ him,
                                                    const JobParam *__restrict__ params, int kstep,
                                                    const float *__restrict__ p32r, const float *__restrict__ p32i,
                                                    const float *__restrict__ tin_r, const float *__restrict__ tin_i,
                                                    float *__restrict__ tout_r, float *__restrict__ tout_i,
                                                    double *__restrict__ norms, const uint32_t *__restrict__ zmask,
                                                    int64_t zwords) {
    const int j = blockIdx.y;
    const JobParam P = params[j];
    if (kstep > P.m + 1) return;
    constexpr int RPP = 256 / NS;
    const int st = threadIdx.x % NS;
    const int rl = threadIdx.x / NS;
    const int64_t tile = static_cast<int64_t>(blockIdx.x) * RPP * RPT;
    const int64_t vb = static_cast<int64_t>(j) * total;
    const int64_t sb = static_cast<int64_t>(j) * N * NS;
    const float *xr = (kstep == 1) ? p32r + sb : tin_r + sb;
    const float *xi = (kstep == 1) ? p32i + sb : tin_i + sb;
    Acc<TV> hr[RPT], hi[RPT];
    for (int d = 0; d < n; ++d) {
        const DiagMeta md = meta[d];
        const int64_t lo = md.rlo, hi_ = md.rhi;
        if (hi_ <= tile || lo >= tile + RPP * RPT) continue;
        uint32_t live = (1u << RPT) - 1;  // the zero-block map, as in horner_kernel
        if (zmask) {
            const int64_t b0 = static_cast<int64_t>(blockIdx.x) * RPT;
            live = (zmask[(static_cast<int64_t>(j) * n + d) * zwords + (b0 >> 5)] >> (b0 & 31)) & ((1u << RPT) - 1);
            if (!live) continue;  // uniform
        }
        const TV *ar = hre + vb + md.rowbase;
        const TV *ai = him + vb + md.rowbase;
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            if (!((live >> q) & 1u)) continue;  // uniform
            const int64_t r = tile + rl + RPP * q;
            const int64_t rc = r < lo ? lo : (r >= hi_ ? hi_ - 1 : r);
            TV a_r = ar[rc], a_i = ai[rc];
            const int64_t xo = (rc + md.off) * NS + st;
            TV x_r = static_cast<TV>(xr[xo]), x_i = static_cast<TV>(xi[xo]);
            if (r >= lo && r < hi_) cmac(hr[q], hi[q], a_r, a_i, x_r, x_i);
        }
    }
    const double c = P.tau / static_cast<double>(kstep);
    double sq = 0.0;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int64_t r = tile + rl + RPP * q;
        if (r >= N) continue;
        const int64_t o = sb + r * NS + st;
        const double vr = c * static_cast<double>(hi[q].get()), vi = -c * static_cast<double>(hr[q].get());
        tout_r[o] = static_cast<float>(vr);
        tout_i[o] = static_cast<float>(vi);
        sq += vr * vr + vi * vi;
    }
    if (kstep >= P.m) {
        // lanes of one state are NS apart in the wave: fold the other lanes onto lanes 0..NS-1
#pragma unroll
        for (int w = 32; w >= NS; w >>= 1) sq += __shfl_xor(sq, w, 64);
        if ((threadIdx.x & 63) < NS) atomicAdd(&norms[(static_cast<int64_t>(j) * NS + st) * 2 + (kstep - P.m)], sq);
    }
}

// The probe with four states per lane (NS = 4, 8 or 16), the gathers as 16-byte loads, as horner4_kernel; R rows per
// lane, tiles of 4 R zero-map blocks.
template <typename TV, int NS, int R>
__global__ __launch_bounds__(256) void probe4_kernel(int64_t N, int n, const DiagMeta *__restrict__ meta,
                                                     int64_t total, const TV *__restrict__ hre,
                                                     const TV *__restrict__ him, const JobParam *__restrict__ params,
                                                     int kstep, const float *__restrict__ p32r,
                                                     const float *__restrict__ p32i, const float *__restrict__ tin_r,
                                                     const float *__restrict__ tin_i, float *__restrict__ tout_r,
                                                     float *__restrict__ tout_i, double *__restrict__ norms,
                                                     const uint32_t *__restrict__ zmask, int64_t zwords) {
    static_assert(NS == 4 || NS == 8 || NS == 16, "four states per lane");
    constexpr int L = NS / 4, WR = 64 / L, RPP = 4 * WR, T = RPP * R;
    const int j = blockIdx.y;
    const JobParam P = params[j];
    if (kstep > P.m + 1) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int h = lane % L, rw = lane / L;
    const int64_t t = blockIdx.x;
    const int64_t tile = t * T;
    const int64_t vb = static_cast<int64_t>(j) * total;
    const int64_t sb = static_cast<int64_t>(j) * N * NS;
    const float *xr = (kstep == 1) ? p32r + sb : tin_r + sb;
    const float *xi = (kstep == 1) ? p32i + sb : tin_i + sb;
    const int64_t blk0 = t * 4 * R;
    Acc<TV> hr[R][4], hi[R][4];
    for (int d = 0; d < n; ++d) {
        const DiagMeta md = meta[d];
        const int64_t lo = md.rlo, hi_ = md.rhi;
        if (hi_ <= tile || lo >= tile + T) continue;  // uniform
        uint32_t live = (4 * R >= 32) ? ~0u : ((1u << (4 * R)) - 1);
        if (zmask) {
            const uint32_t w = zmask[(static_cast<int64_t>(j) * n + d) * zwords + (blk0 >> 5)];
            live = (4 * R >= 32) ? w : ((w >> (blk0 & 31)) & ((1u << (4 * R)) - 1));
            if (!live) continue;  // uniform
        }
        const TV *ar = hre + vb + md.rowbase;
        const TV *ai = him + vb + md.rowbase;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            if (!((live >> (4 * q + wave)) & 1u)) continue;  // uniform in the wave
            const int64_t r = tile + RPP * q + WR * wave + rw;
            const int64_t rc = r < lo ? lo : (r >= hi_ ? hi_ - 1 : r);
            const TV a_r = ar[rc], a_i = ai[rc];
            const int64_t xo = (rc + md.off) * NS + 4 * h;
            const float4 x_r = *reinterpret_cast<const float4 *>(xr + xo);
            const float4 x_i = *reinterpret_cast<const float4 *>(xi + xo);
            if (r >= lo && r < hi_) {
                cmac(hr[q][0], hi[q][0], a_r, a_i, static_cast<TV>(x_r.x), static_cast<TV>(x_i.x));
                cmac(hr[q][1], hi[q][1], a_r, a_i, static_cast<TV>(x_r.y), static_cast<TV>(x_i.y));
                cmac(hr[q][2], hi[q][2], a_r, a_i, static_cast<TV>(x_r.z), static_cast<TV>(x_i.z));
                cmac(hr[q][3], hi[q][3], a_r, a_i, static_cast<TV>(x_r.w), static_cast<TV>(x_i.w));
            }
        }
    }
    const double c = P.tau / static_cast<double>(kstep);
    double sq[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int64_t r = tile + RPP * q + WR * wave + rw;
        if (r >= N) continue;
        const int64_t o = sb + r * NS + 4 * h;
        float vr32[4], vi32[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const double vr = c * static_cast<double>(hi[q][v].get()), vi = -c * static_cast<double>(hr[q][v].get());
            vr32[v] = static_cast<float>(vr);
            vi32[v] = static_cast<float>(vi);
            sq[v] += vr * vr + vi * vi;
        }
        *reinterpret_cast<float4 *>(tout_r + o) = make_float4(vr32[0], vr32[1], vr32[2], vr32[3]);
        *reinterpret_cast<float4 *>(tout_i + o) = make_float4(vi32[0], vi32[1], vi32[2], vi32[3]);
    }
    if (kstep >= P.m) {
        // lanes holding the same four states are L apart in the wave: fold them onto lanes 0 .. L-1
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int w = 32; w >= L; w >>= 1) sq[v] += __shfl_xor(sq[v], w, 64);
        if (lane < L)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                atomicAdd(&norms[(static_cast<int64_t>(j) * NS + 4 * h + v) * 2 + (kstep - P.m)], sq[v]);
    }
}

// Squared norms of every state (FP64): per-workgroup partials partial[(j * gridDim.x + block) * NS + state], folded in
// a fixed order by fold_states_kernel (no atomics: every replay gives the same bits), then psi /= ||psi||.
template <int NS>
__global__ __launch_bounds__(256) void sumsq_states_kernel(int64_t N, const double *__restrict__ pr,
                                                           const double *__restrict__ pi, double *__restrict__ partial) {
    __shared__ double red[4][NS];
    const int j = blockIdx.y;
    constexpr int RPP = 256 / NS;
    const int st = threadIdx.x % NS;
    const int64_t sb = static_cast<int64_t>(j) * N * NS;
    double sq = 0.0;
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * RPP + threadIdx.x / NS; r < N;
         r += static_cast<int64_t>(gridDim.x) * RPP) {
        const double a = pr[sb + r * NS + st], b = pi[sb + r * NS + st];
        sq += a * a + b * b;
    }
#pragma unroll
    for (int w = 32; w >= NS; w >>= 1) sq += __shfl_xor(sq, w, 64);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) < NS) red[wave][st] = sq;
    __syncthreads();
    if (threadIdx.x < NS) {
        const double v = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
        partial[(static_cast<int64_t>(j) * gridDim.x + blockIdx.x) * NS + threadIdx.x] = v;
    }
}

template <int NS>
__global__ void fold_states_kernel(int nblocks, const double *__restrict__ partial, double *__restrict__ out) {
    const int j = blockIdx.x;
    for (int st = threadIdx.x; st < NS; st += blockDim.x) {
        double v = 0.0;
        for (int b = 0; b < nblocks; ++b) v += partial[(static_cast<int64_t>(j) * nblocks + b) * NS + st];
        out[static_cast<int64_t>(j) * NS + st] = v;
    }
}

template <int NS>
__global__ __launch_bounds__(256) void scale_states_kernel(int64_t N, double *__restrict__ pr, double *__restrict__ pi,
                                                           float *__restrict__ qr, float *__restrict__ qi,
                                                           const double *__restrict__ sq) {
    const int j = blockIdx.y;
    constexpr int RPP = 256 / NS;
    const int st = threadIdx.x % NS;
    const int64_t sb = static_cast<int64_t>(j) * N * NS;
    const double s2 = sq[static_cast<int64_t>(j) * NS + st];
    const double f = s2 > 0.0 ? 1.0 / sqrt(s2) : 1.0;
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * RPP + threadIdx.x / NS; r < N;
         r += static_cast<int64_t>(gridDim.x) * RPP) {
        const int64_t o = sb + r * NS + st;
        const double a = pr[o] * f, b = pi[o] * f;
        pr[o] = a;
        pi[o] = b;
        qr[o] = static_cast<float>(a);
        qi[o] = static_cast<float>(b);
    }
}

// dst = src over n floats (the FP32 state into the first level's input buffer).
__global__ void copyf_kernel(int64_t n, const float *__restrict__ ar, const float *__restrict__ ai,
                             float *__restrict__ br, float *__restrict__ bi) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        br[i] = ar[i];
        bi[i] = ai[i];
    }
}

__global__ void zero_kernel(int64_t n, double *p) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

struct EvoClass {
    uint32_t N = 0;
    std::vector<int32_t> offs;
    std::vector<DiagMeta> meta;
    DevBuf d_meta;
    int64_t total = 0;        // value elements per job
    std::vector<int> jobs;    // global job ids
    DevBuf hre, him;          // [job][total]
    DevBuf p64r, p64i;        // [job][N][NS] FP64
    DevBuf p32r, p32i;        // FP32 copy
    DevBuf b0r, b0i, b1r, b1i;
    DevBuf d_params;
    DevBuf d_sq;              // [job][NS] squared norms
    DevBuf d_partial;         // [job][block][NS] partial squared norms
    DevBuf d_probe;           // [job][NS][2]
    DevBuf zmask;             // [job][diagonal][zwords] live-block bits of horner_kernel
    int64_t zwords = 0;
    std::vector<JobParam> params;
    int64_t G = 0;            // launches: max steps
    int rpt = 4;              // tuned rows per lane
    int xcd = 0;              // tuned row-tile order: 0 round-robin over the XCDs, 1 XCD-contiguous
    int lds = 0;              // tuned: 1 = stage each tile's input window in LDS
    int vec = 0;              // tuned: 1 = horner4_kernel (four states per lane; NS = 4, 8, 16, no LDS window)
    int du = 1;               // tuned: horner4_kernel's diagonals per load batch (1, 2 or 4; 1 or 2 rows per lane)
    hipStream_t stream = nullptr;
    int jobs_n() const { return static_cast<int>(jobs.size()); }
};

}  // namespace dq

struct dq_evo {
    int n_states = 1;
    int value_precision = DQ_F32;
    int n_jobs = 0;
    std::vector<std::pair<int, int>> where;  // job -> (class, index in class)
    std::vector<dq::EvoClass> classes;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<int64_t> graph_sig;  // what the built graph's launches depend on (graph_signature)
    std::vector<hipStream_t> lanes;  // streams of the lane plan (lane_plan), made on first use
    ~dq_evo() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto &c : classes)
            if (c.stream) (void)hipStreamDestroy(c.stream);
        for (auto l : lanes) (void)hipStreamDestroy(l);
    }
};

namespace dq {

// Process-wide overrides of every class's tuned Horner schedule (dq_set_tuning "evo_rows_per_lane", "evo_xcd",
// "evo_lds").
static int g_evo_rpt = 0;   // 0 = the class's, else 1, 2, 4 or 8
static int g_evo_xcd = -1;  // -1 = the class's, else 0 or 1
static int g_evo_lds = -1;  // -1 = the class's, else 0 or 1
static int g_evo_skip = 1;  // skip the all-zero blocks of a diagonal ("evo_skip_zero_blocks")
static int g_evo_vec = -1;  // -1 = the class's, else 0 (a lane per state) or 1 (four states per lane) ("evo_vec")
static int g_evo_lanes = 4; // streams the classes are spread over in dq_evo_run; 0 = one per class ("evo_lanes")
static int g_evo_du = 0;    // 0 = the class's, else the diagonals whose loads horner4_kernel issues together, for
                            // 1 or 2 rows per lane ("evo_diag_batch": 0, 1, 2 or 4)

bool evo_set_tuning(const std::string &key, int value) {
    if (key == "evo_rows_per_lane") {
        if (value != 0 && value != 1 && value != 2 && value != 4 && value != 8)
            throw ArgError("evo_rows_per_lane must be 0, 1, 2, 4 or 8");
        g_evo_rpt = value;
        return true;
    }
    if (key == "evo_xcd") {
        if (value < -1 || value > 1) throw ArgError("evo_xcd must be -1, 0 or 1");
        g_evo_xcd = value;
        return true;
    }
    if (key == "evo_lds") {
        if (value < -1 || value > 1) throw ArgError("evo_lds must be -1, 0 or 1");
        g_evo_lds = value;
        return true;
    }
    if (key == "evo_vec") {
        if (value < -1 || value > 1) throw ArgError("evo_vec must be -1, 0 or 1");
        g_evo_vec = value;
        return true;
    }
    if (key == "evo_diag_batch") {
        if (value != 0 && value != 1 && value != 2 && value != 4)
            throw ArgError("evo_diag_batch must be 0, 1, 2 or 4");
        g_evo_du = value;
        return true;
    }
    if (key == "evo_lanes") {
        if (value < 0 || value > 8) throw ArgError("evo_lanes must be 0 (a stream per class) or 1 to 8");
        g_evo_lanes = value;
        return true;
    }
    if (key == "evo_skip_zero_blocks") {
        if (value != 0 && value != 1) throw ArgError("evo_skip_zero_blocks must be 0 or 1");
        g_evo_skip = value;
        return true;
    }
    return false;
}

template <typename TV, int NS>
static void launch_level(const EvoClass &c, int64_t g, hipStream_t s) {
    const int64_t N = c.N;
    constexpr int RPP = 256 / NS;
    const int rpt = g_evo_rpt > 0 ? g_evo_rpt : c.rpt;
    const int xcd = g_evo_xcd >= 0 ? g_evo_xcd : c.xcd;
    const bool lds = (g_evo_lds >= 0 ? g_evo_lds : c.lds) == 1;
    float *bout_r = (g & 1) ? c.b1r.as<float>() : c.b0r.as<float>();
    float *bout_i = (g & 1) ? c.b1i.as<float>() : c.b0i.as<float>();
    const float *bin_r = (g & 1) ? c.b0r.as<float>() : c.b1r.as<float>();
    const float *bin_i = (g & 1) ? c.b0i.as<float>() : c.b1i.as<float>();
    const uint32_t *zmask = g_evo_skip ? c.zmask.as<uint32_t>() : nullptr;
    if constexpr (NS == 4 || NS == 8 || NS == 16) {
        if ((g_evo_vec >= 0 ? g_evo_vec : c.vec) == 1) {  // four states per lane (no LDS window)
            const int64_t rows4 = static_cast<int64_t>(1024 / NS) * rpt;
            const int64_t nt4 = (N + rows4 - 1) / rows4;
            const dim3 grid4(static_cast<unsigned>(grid_for_tiles(nt4, xcd)), static_cast<unsigned>(c.jobs_n()));
#define DQ_HORNER4(R, DU)                                                                                          \
    hipLaunchKernelGGL((horner4_kernel<TV, NS, R, DU>), grid4, dim3(256), 0, s, N, (int)c.offs.size(),              \
                       c.d_meta.as<DiagMeta>(), c.total, c.hre.as<TV>(), c.him.as<TV>(), c.d_params.as<JobParam>(), \
                       g, c.p64r.as<double>(), c.p64i.as<double>(), c.p32r.as<float>(), c.p32i.as<float>(), bin_r,   \
                       bin_i, bout_r, bout_i, nt4, xcd, zmask, c.zwords)
            const int du = rpt <= 2 ? (g_evo_du > 0 ? g_evo_du : c.du) : 1;
            switch (rpt * 8 + du) {
                case 9: DQ_HORNER4(1, 1); break;
                case 10: DQ_HORNER4(1, 2); break;
                case 12: DQ_HORNER4(1, 4); break;
                case 17: DQ_HORNER4(2, 1); break;
                case 18: DQ_HORNER4(2, 2); break;
                case 20: DQ_HORNER4(2, 4); break;
                case 33: DQ_HORNER4(4, 1); break;
                case 65: DQ_HORNER4(8, 1); break;
                default: throw ArgError("evolve: rows per lane must be 1, 2, 4 or 8");
            }
#undef DQ_HORNER4
            DQ_LAUNCH_CHECK();
            return;
        }
    }
    const int64_t rows_per_block = static_cast<int64_t>(RPP) * rpt;
    const int64_t ntiles = (N + rows_per_block - 1) / rows_per_block;
    const dim3 grid(static_cast<unsigned>(grid_for_tiles(ntiles, xcd)), static_cast<unsigned>(c.jobs_n()));
#define DQ_HORNER_L(R, L)                                                                                        \
    hipLaunchKernelGGL((horner_kernel<TV, NS, R, L>), grid, dim3(256), 0, s, N, (int)c.offs.size(),               \
                       c.d_meta.as<DiagMeta>(), c.total, c.hre.as<TV>(), c.him.as<TV>(), c.d_params.as<JobParam>(), \
                       g, c.p64r.as<double>(), c.p64i.as<double>(), c.p32r.as<float>(), c.p32i.as<float>(), bin_r,   \
                       bin_i, bout_r, bout_i, ntiles, xcd, zmask, c.zwords)
#define DQ_HORNER(R)                         \
    if (lds) { DQ_HORNER_L(R, true); }       \
    else { DQ_HORNER_L(R, false); }
    switch (rpt) {
        case 1: DQ_HORNER(1); break;
        case 2: DQ_HORNER(2); break;
        case 4: DQ_HORNER(4); break;
        case 8: DQ_HORNER(8); break;
        default: throw ArgError("evolve: rows per lane must be 1, 2, 4 or 8");
    }
#undef DQ_HORNER
#undef DQ_HORNER_L
    DQ_LAUNCH_CHECK();
}

constexpr int RENORM_BLOCKS = 256;

template <int NS>
static void launch_renorm(const EvoClass &c, hipStream_t s) {
    const int64_t N = c.N;
    const int nj = c.jobs_n();
    const unsigned gx = static_cast<unsigned>(std::min<int64_t>(RENORM_BLOCKS, (N + (256 / NS) - 1) / (256 / NS)));
    hipLaunchKernelGGL(sumsq_states_kernel<NS>, dim3(gx, nj), dim3(256), 0, s, N, c.p64r.as<double>(),
                       c.p64i.as<double>(), c.d_partial.as<double>());
    hipLaunchKernelGGL(fold_states_kernel<NS>, dim3(nj), dim3(64), 0, s, (int)gx, c.d_partial.as<double>(),
                       c.d_sq.as<double>());
    hipLaunchKernelGGL(scale_states_kernel<NS>, dim3(gx, nj), dim3(256), 0, s, N, c.p64r.as<double>(),
                       c.p64i.as<double>(), c.p32r.as<float>(), c.p32i.as<float>(), c.d_sq.as<double>());
    DQ_LAUNCH_CHECK();
}

template <typename TV, int NS>
static void launch_probe(const EvoClass &c, int kstep, hipStream_t s) {
    const int64_t N = c.N;
    constexpr int RPP = 256 / NS;
    constexpr int R = 4;
    float *tout_r = (kstep & 1) ? c.b1r.as<float>() : c.b0r.as<float>();
    float *tout_i = (kstep & 1) ? c.b1i.as<float>() : c.b0i.as<float>();
    const float *tin_r = (kstep & 1) ? c.b0r.as<float>() : c.b1r.as<float>();
    const float *tin_i = (kstep & 1) ? c.b0i.as<float>() : c.b1i.as<float>();
    if constexpr (NS == 4 || NS == 8 || NS == 16) {  // four states per lane: a quarter of the load instructions
        constexpr int T4 = (1024 / NS) * R;
        const dim3 grid4(static_cast<unsigned>((N + T4 - 1) / T4), static_cast<unsigned>(c.jobs_n()));
        hipLaunchKernelGGL((probe4_kernel<TV, NS, R>), grid4, dim3(256), 0, s, N, (int)c.offs.size(),
                           c.d_meta.as<DiagMeta>(), c.total, c.hre.as<TV>(), c.him.as<TV>(),
                           c.d_params.as<JobParam>(), kstep, c.p32r.as<float>(), c.p32i.as<float>(), tin_r, tin_i,
                           tout_r, tout_i, c.d_probe.as<double>(), g_evo_skip ? c.zmask.as<uint32_t>() : nullptr,
                           c.zwords);
        DQ_LAUNCH_CHECK();
        return;
    }
    const dim3 grid(static_cast<unsigned>((N + RPP * R - 1) / (RPP * R)), static_cast<unsigned>(c.jobs_n()));
    hipLaunchKernelGGL((probe_kernel<TV, NS, R>), grid, dim3(256), 0, s, N, (int)c.offs.size(),
                       c.d_meta.as<DiagMeta>(), c.total, c.hre.as<TV>(), c.him.as<TV>(), c.d_params.as<JobParam>(),
                       kstep, c.p32r.as<float>(), c.p32i.as<float>(), tin_r, tin_i, tout_r, tout_i,
                       c.d_probe.as<double>(), g_evo_skip ? c.zmask.as<uint32_t>() : nullptr, c.zwords);
    DQ_LAUNCH_CHECK();
}

// Dispatch on (value precision, states per job).
template <typename F>
static void dispatch(int prec, int ns, F &&f) {
    auto go = [&](auto tv) {
        using TV = decltype(tv);
        switch (ns) {
            case 1: f(TV{}, std::integral_constant<int, 1>{}); break;
            case 2: f(TV{}, std::integral_constant<int, 2>{}); break;
            case 4: f(TV{}, std::integral_constant<int, 4>{}); break;
            case 8: f(TV{}, std::integral_constant<int, 8>{}); break;
            case 16: f(TV{}, std::integral_constant<int, 16>{}); break;
            default: throw ArgError("evolve: states per job must be 1, 2, 4, 8 or 16");
        }
    };
    if (prec == DQ_F32) go(float{}); else go(double{});
}

static void run_class_levels(dq_evo &E, EvoClass &c, hipStream_t s, bool renorm) {
    dispatch(E.value_precision, E.n_states, [&](auto tv, auto ns) {
        using TV = decltype(tv);
        constexpr int NS = decltype(ns)::value;
        // prologue: the FP32 state into the input buffer of level 0 (b1: level g reads buffer (g+1) mod 2)
        const int64_t sv = static_cast<int64_t>(c.N) * NS * c.jobs_n();
        hipLaunchKernelGGL(copyf_kernel, dim3(blocks_for(sv, 256, 4096)), dim3(256), 0, s, sv, c.p32r.as<float>(),
                           c.p32i.as<float>(), c.b1r.as<float>(), c.b1i.as<float>());
        DQ_LAUNCH_CHECK();
        for (int64_t g = 0; g < c.G; ++g) launch_level<TV, NS>(c, g, s);
        if (renorm) launch_renorm<NS>(c, s);
    });
}

// Which classes run on which stream, in order. Branches of a graph share the process's few hardware queues, and a
// queue runs what it is given in order, so with a stream per class the classes mostly ran one after another, the
// longest often late (256-TFIM batch: one kernel in flight 79 % of the graph's time). Instead the classes go to
// g_evo_lanes streams by longest-first list scheduling on an estimate of their time -- per level a launch cost, per
// job-level the bytes of the values and the 8-byte complex states at ~3 TB/s -- and each lane runs its classes
// longest first. g_evo_lanes = 0 keeps a stream per class.
static std::vector<std::vector<int>> lane_plan(const dq_evo &E) {
    const int nc = static_cast<int>(E.classes.size());
    std::vector<std::vector<int>> plan;
    if (g_evo_lanes == 0) {
        for (int ci = 0; ci < nc; ++ci) plan.push_back({ci});
        return plan;
    }
    std::vector<double> cost(nc, 0.0);
    for (int ci = 0; ci < nc; ++ci) {
        const EvoClass &c = E.classes[ci];
        int64_t job_levels = 0;
        for (const auto &p : c.params) job_levels += p.steps;
        // per row: the values in live blocks (about half of the 2 x elem_bytes per diagonal stored) and per state
        // the gathered input, psi's copy and the output
        const double bytes_per_job_level =
            static_cast<double>(c.N) * (static_cast<double>(elem_bytes(E.value_precision)) *
                                            static_cast<double>(c.offs.size()) +
                                        24.0 * E.n_states);
        cost[ci] = 0.012 * static_cast<double>(c.G) + static_cast<double>(job_levels) * bytes_per_job_level / 3e9;
    }
    std::vector<int> order(nc);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    const int k = std::min(g_evo_lanes, std::max(1, nc));
    plan.assign(k, {});
    std::vector<double> load(k, 0.0);
    for (int ci : order) {
        const int l = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
        plan[l].push_back(ci);
        load[l] += cost[ci];
    }
    return plan;
}

// What a captured graph's launches depend on: the renormalisation flag, the process-wide Horner overrides, per class
// its number of levels and its tuned schedule, and the lane plan. The kernels read each job's (tau, m, s) from
// d_params when they run, so a new schedule that keeps every class's number of levels and the plan replays the same
// graph.
static std::vector<int64_t> graph_signature(const dq_evo &E, bool renorm) {
    std::vector<int64_t> sig{renorm ? 1 : 0, g_evo_rpt, g_evo_xcd, g_evo_lds, g_evo_skip, g_evo_vec, g_evo_lanes,
                             g_evo_du};
    for (const auto &c : E.classes) sig.insert(sig.end(), {c.G, c.rpt, c.xcd, c.lds, c.vec, c.du});
    for (const auto &lane : lane_plan(E)) {
        sig.push_back(-1);
        sig.insert(sig.end(), lane.begin(), lane.end());
    }
    return sig;
}

}  // namespace dq

using namespace dq;

extern "C" int dq_evo_create(int n_jobs, dq_devmat *const *H, int n_states, int value_precision, dq_evo **out) {
    return guard([&] {
        if (!out || n_jobs < 0 || (n_jobs > 0 && !H)) throw ArgError("dq_evo_create: bad arguments");
        if (value_precision != DQ_F32 && value_precision != DQ_F64) throw ArgError("value precision: DQ_F32 or DQ_F64");
        if (n_states != 1 && n_states != 2 && n_states != 4 && n_states != 8 && n_states != 16)
            throw ArgError("states per job must be 1, 2, 4, 8 or 16");
        auto E = std::make_unique<dq_evo>();
        E->n_states = n_states;
        E->value_precision = value_precision;
        E->n_jobs = n_jobs;
        std::map<std::pair<uint32_t, std::vector<int32_t>>, int> key;
        E->where.resize(n_jobs);
        for (int j = 0; j < n_jobs; ++j) {
            const dq_devmat *A = H[j];
            if (!A) throw ArgError("dq_evo_create: NULL matrix");
            if (A->rows != A->cols) throw ArgError("dq_evo_create: Hamiltonians must be square");
            auto kk = std::make_pair(A->rows, A->offs);
            auto it = key.find(kk);
            int ci;
            if (it == key.end()) {
                ci = static_cast<int>(E->classes.size());
                key[kk] = ci;
                E->classes.emplace_back();
                EvoClass &c = E->classes.back();
                c.N = A->rows;
                c.offs = A->offs;
                // the class layout: the matrix layout in the plan's value precision
                auto L = make_devmat(A->rows, A->cols, A->offs.data(), A->n(), value_precision, 0);
                c.meta = L->meta;
                c.total = L->total;
            } else {
                ci = it->second;
            }
            E->where[j] = {ci, static_cast<int>(E->classes[ci].jobs.size())};
            E->classes[ci].jobs.push_back(j);
            if (E->classes[ci].jobs.size() > 65535)  // a class's jobs index the grid's y (or z) dimension
                throw ArgError("dq_evo_create: more than 65535 Hamiltonians of one size and diagonal set");
        }
        const size_t vb = elem_bytes(value_precision);
        for (auto &c : E->classes) {
            const int nj = c.jobs_n();
            c.d_meta = DevBuf(sizeof(DiagMeta) * std::max<size_t>(1, c.meta.size()));
            if (!c.meta.empty())
                DQ_HIP(hipMemcpy(c.d_meta.p, c.meta.data(), sizeof(DiagMeta) * c.meta.size(), hipMemcpyHostToDevice));
            c.hre = DevBuf(vb * c.total * nj);
            c.him = DevBuf(vb * c.total * nj);
            DQ_HIP(hipMemset(c.hre.p, 0, c.hre.bytes));
            DQ_HIP(hipMemset(c.him.p, 0, c.him.bytes));
            const size_t sv = static_cast<size_t>(c.N) * n_states * nj;
            c.p64r = DevBuf(sv * 8); c.p64i = DevBuf(sv * 8);
            c.p32r = DevBuf(sv * 4); c.p32i = DevBuf(sv * 4);
            c.b0r = DevBuf(sv * 4); c.b0i = DevBuf(sv * 4); c.b1r = DevBuf(sv * 4); c.b1i = DevBuf(sv * 4);
            for (DevBuf *b : {&c.p64r, &c.p64i}) DQ_HIP(hipMemset(b->p, 0, b->bytes));
            for (DevBuf *b : {&c.p32r, &c.p32i, &c.b0r, &c.b0i, &c.b1r, &c.b1i}) DQ_HIP(hipMemset(b->p, 0, b->bytes));
            c.d_params = DevBuf(sizeof(JobParam) * nj);
            c.params.assign(nj, JobParam{0.0, 1, 0, 0});
            DQ_HIP(hipMemcpy(c.d_params.p, c.params.data(), sizeof(JobParam) * nj, hipMemcpyHostToDevice));
            c.d_sq = DevBuf(sizeof(double) * nj * n_states);
            c.d_partial = DevBuf(sizeof(double) * nj * n_states * RENORM_BLOCKS);
            c.d_probe = DevBuf(sizeof(double) * nj * n_states * 2);
            c.rpt = c.N >= (1u << 14) ? 4 : 1;
            DQ_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
        }
        // copy each Hamiltonian's values into its class slot (converted to the plan's precision: same layout)
        for (int j = 0; j < n_jobs; ++j) {
            EvoClass &c = E->classes[E->where[j].first];
            const int64_t slot = E->where[j].second;
            dq_devmat *T = nullptr;
            if (dq_devmat_convert(H[j], value_precision, &T) != DQ_OK) throw HipError(dq_last_error());
            std::unique_ptr<dq_devmat, void (*)(dq_devmat *)> hold(T, dq_devmat_destroy);
            if (T->total != c.total) throw ArgError("dq_evo_create: layout mismatch inside a class");
            const size_t bytes = vb * c.total;
            DQ_HIP(hipMemcpy(static_cast<char *>(c.hre.p) + slot * bytes, T->re.p, bytes, hipMemcpyDeviceToDevice));
            DQ_HIP(hipMemcpy(static_cast<char *>(c.him.p) + slot * bytes, T->im.p, bytes, hipMemcpyDeviceToDevice));
        }
        // the zero-block maps, from the values just copied
        for (auto &c : E->classes) {
            const int nd = static_cast<int>(c.offs.size());
            const int64_t rpp = 256 / n_states;
            c.zwords = zero_block_words(c.N, static_cast<int>(rpp));
            c.zmask = DevBuf(sizeof(uint32_t) * std::max<int64_t>(1, c.zwords * nd * c.jobs_n()));
            if (nd == 0 || c.jobs_n() == 0) continue;
            const dim3 grid(static_cast<unsigned>(c.zwords), zero_map_grid_y(nd), static_cast<unsigned>(c.jobs_n()));
            dispatch(value_precision, n_states, [&](auto tv, auto ns) {
                using TV = decltype(tv);
                constexpr int RPP = 256 / decltype(ns)::value;
                hipLaunchKernelGGL((zero_blocks_kernel<TV, RPP>), grid, dim3(256), 0, nullptr,
                                   static_cast<int64_t>(c.N), nd, c.d_meta.as<DiagMeta>(), c.total, c.hre.as<TV>(),
                                   c.him.as<TV>(), c.zwords, c.zmask.as<uint32_t>());
            });
            DQ_LAUNCH_CHECK();
        }
        DQ_HIP(hipDeviceSynchronize());
        *out = E.release();
    });
}

extern "C" void dq_evo_destroy(dq_evo *E) { delete E; }

extern "C" int dq_evo_info(const dq_evo *E, int *n_classes, int *n_jobs, int *n_states) {
    return guard([&] {
        if (!E) throw ArgError("plan is NULL");
        if (n_classes) *n_classes = static_cast<int>(E->classes.size());
        if (n_jobs) *n_jobs = E->n_jobs;
        if (n_states) *n_states = E->n_states;
    });
}

// States of job j from host arrays, state-major: psi_s[r] = re[s*N + r] + i im[s*N + r].
extern "C" int dq_evo_set_states(dq_evo *E, int job, const double *re, const double *im) {
    return guard([&] {
        if (!E || job < 0 || job >= E->n_jobs || !re || !im) throw ArgError("dq_evo_set_states: bad arguments");
        EvoClass &c = E->classes[E->where[job].first];
        const int64_t slot = E->where[job].second, N = c.N, NS = E->n_states;
        std::vector<double> r(N * NS), i(N * NS);
        std::vector<float> fr(N * NS), fi(N * NS);
        for (int64_t s = 0; s < NS; ++s)
            for (int64_t k = 0; k < N; ++k) {
                r[k * NS + s] = re[s * N + k];
                i[k * NS + s] = im[s * N + k];
                fr[k * NS + s] = static_cast<float>(re[s * N + k]);
                fi[k * NS + s] = static_cast<float>(im[s * N + k]);
            }
        const size_t off = static_cast<size_t>(slot) * N * NS;
        DQ_HIP(hipMemcpy(c.p64r.as<double>() + off, r.data(), N * NS * 8, hipMemcpyHostToDevice));
        DQ_HIP(hipMemcpy(c.p64i.as<double>() + off, i.data(), N * NS * 8, hipMemcpyHostToDevice));
        DQ_HIP(hipMemcpy(c.p32r.as<float>() + off, fr.data(), N * NS * 4, hipMemcpyHostToDevice));
        DQ_HIP(hipMemcpy(c.p32i.as<float>() + off, fi.data(), N * NS * 4, hipMemcpyHostToDevice));
    });
}

extern "C" int dq_evo_get_states(const dq_evo *E, int job, double *re, double *im) {
    return guard([&] {
        if (!E || job < 0 || job >= E->n_jobs || !re || !im) throw ArgError("dq_evo_get_states: bad arguments");
        const EvoClass &c = E->classes[E->where[job].first];
        const int64_t slot = E->where[job].second, N = c.N, NS = E->n_states;
        std::vector<double> r(N * NS), i(N * NS);
        const size_t off = static_cast<size_t>(slot) * N * NS;
        DQ_HIP(hipMemcpy(r.data(), c.p64r.as<double>() + off, N * NS * 8, hipMemcpyDeviceToHost));
        DQ_HIP(hipMemcpy(i.data(), c.p64i.as<double>() + off, N * NS * 8, hipMemcpyDeviceToHost));
        for (int64_t s = 0; s < NS; ++s)
            for (int64_t k = 0; k < N; ++k) {
                re[s * N + k] = r[k * NS + s];
                im[s * N + k] = i[k * NS + s];
            }
    });
}

// Per-job schedule: evolve to t[j] with 2^s[j] sub-steps of degree m[j]. The parameters go to device memory, where
// the kernels read them; dq_evo_run rebuilds its graph only when some class's number of levels changed.
extern "C" int dq_evo_set_schedule(dq_evo *E, const double *t, const int32_t *m, const int32_t *s) {
    return guard([&] {
        if (!E || (E->n_jobs && (!t || !m || !s))) throw ArgError("dq_evo_set_schedule: bad arguments");
        for (int j = 0; j < E->n_jobs; ++j) {
            if (m[j] < 1 || m[j] > 64) throw ArgError("dq_evo_set_schedule: m must be in [1, 64]");
            if (s[j] < 0 || s[j] > 24) throw ArgError("dq_evo_set_schedule: s must be in [0, 24]");
        }
        for (auto &c : E->classes) {
            c.G = 0;
            for (int q = 0; q < c.jobs_n(); ++q) {
                const int j = c.jobs[q];
                const int64_t nsub = int64_t(1) << s[j];
                c.params[q] = JobParam{std::ldexp(t[j], -s[j]), m[j], s[j], nsub * m[j]};
                c.G = std::max(c.G, c.params[q].steps);
            }
            DQ_HIP(hipMemcpy(c.d_params.p, c.params.data(), sizeof(JobParam) * c.jobs_n(), hipMemcpyHostToDevice));
        }
    });
}

// Rows per lane, row-tile order and LDS window of each class, chosen by timing a few Horner levels of each
// candidate (the schedule must be set).
extern "C" int dq_evo_tune(dq_evo *E, int *chosen_rpt, int *chosen_xcd, int *chosen_lds) {
    return guard([&] {
        if (!E) throw ArgError("plan is NULL");
        hipEvent_t e0, e1;
        DQ_HIP(hipEventCreate(&e0));
        DQ_HIP(hipEventCreate(&e1));
        for (size_t ci = 0; ci < E->classes.size(); ++ci) {
            EvoClass &c = E->classes[ci];
            float best = 1e30f;
            int brpt = c.rpt, bxcd = c.xcd, blds = c.lds;
            // scratch copy of the states: tuning must not advance them
            const size_t sv = static_cast<size_t>(c.N) * E->n_states * c.jobs_n();
            DevBuf sr(sv * 8), si(sv * 8);
            DQ_HIP(hipMemcpy(sr.p, c.p64r.p, sv * 8, hipMemcpyDeviceToDevice));
            DQ_HIP(hipMemcpy(si.p, c.p64i.p, sv * 8, hipMemcpyDeviceToDevice));
            DevBuf fr(sv * 4), fi(sv * 4);
            DQ_HIP(hipMemcpy(fr.p, c.p32r.p, sv * 4, hipMemcpyDeviceToDevice));
            DQ_HIP(hipMemcpy(fi.p, c.p32i.p, sv * 4, hipMemcpyDeviceToDevice));
            const bool vec_ok = E->n_states == 4 || E->n_states == 8 || E->n_states == 16;
            int bvec = c.vec, bdu = c.du;
            for (int cand = 0; cand < 28; ++cand) {
                // 0-15: a lane per state (rows per lane x tile order x LDS window); 16-23: four states per lane;
                // 24-27: four states per lane, 1 or 2 rows per lane, the loads of 2 diagonals batched (batches of
                // 4 time about as well on a level with every job active but ran the 256-TFIM graph ~3 ms slower)
                const int vec = cand >= 16 ? 1 : 0;
                if (vec && !vec_ok) break;
                const int rpt = cand >= 24 ? 1 << (cand & 1) : 1 << (cand & 3);
                const int xcd = (cand >> (cand >= 24 ? 1 : 2)) & 1, lds = vec ? 0 : cand >> 3;
                const int du = cand >= 24 ? 2 : 1;
                c.rpt = rpt;
                c.xcd = xcd;
                c.lds = lds;
                c.vec = vec;
                c.du = du;
                dispatch(E->value_precision, E->n_states, [&](auto tv, auto ns) {
                    using TV = decltype(tv);
                    constexpr int NS = decltype(ns)::value;
                    launch_level<TV, NS>(c, 0, c.stream);  // warm; the states are restored below
                    DQ_HIP(hipEventRecord(e0, c.stream));
                    for (int r = 0; r < 4; ++r) launch_level<TV, NS>(c, 0, c.stream);
                    DQ_HIP(hipEventRecord(e1, c.stream));
                });
                DQ_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                DQ_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) { best = ms; brpt = rpt; bxcd = xcd; blds = lds; bvec = vec; bdu = du; }
            }
            c.rpt = brpt;
            c.xcd = bxcd;
            c.lds = blds;
            c.vec = bvec;
            c.du = bdu;
            if (chosen_rpt) chosen_rpt[ci] = brpt;
            if (chosen_xcd) chosen_xcd[ci] = bxcd;
            if (chosen_lds) chosen_lds[ci] = bvec ? (bdu == 1 ? 2 : (bdu == 2 ? 3 : 4)) : blds;
            DQ_HIP(hipMemcpy(c.p64r.p, sr.p, sv * 8, hipMemcpyDeviceToDevice));
            DQ_HIP(hipMemcpy(c.p64i.p, si.p, sv * 8, hipMemcpyDeviceToDevice));
            DQ_HIP(hipMemcpy(c.p32r.p, fr.p, sv * 4, hipMemcpyDeviceToDevice));
            DQ_HIP(hipMemcpy(c.p32i.p, fi.p, sv * 4, hipMemcpyDeviceToDevice));
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (E->exec) { (void)hipGraphExecDestroy(E->exec); E->exec = nullptr; }
        if (E->graph) { (void)hipGraphDestroy(E->graph); E->graph = nullptr; }
    });
}

// Evolve every job by its schedule. use_graph: capture all classes (one stream each, forked from and joined to
// `stream`) into one graph on the first call and replay it afterwards; otherwise launch directly. renormalize: end
// with the FP64 renormalisation of every state. ms_out (may be NULL): elapsed GPU time.
extern "C" int dq_evo_run(dq_evo *E, int use_graph, int renormalize, uintptr_t stream, float *ms_out) {
    return guard([&] {
        if (!E) throw ArgError("plan is NULL");
        hipStream_t s0 = as_stream(stream);
        hipEvent_t e0, e1, fork;
        DQ_HIP(hipEventCreate(&e0));
        DQ_HIP(hipEventCreate(&e1));
        DQ_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
        const std::vector<std::vector<int>> plan = lane_plan(*E);
        if (g_evo_lanes > 0)
            while (E->lanes.size() < plan.size()) {
                hipStream_t l;
                DQ_HIP(hipStreamCreateWithFlags(&l, hipStreamNonBlocking));
                E->lanes.push_back(l);
            }
        std::vector<hipEvent_t> joins(plan.size());
        for (auto &j : joins) DQ_HIP(hipEventCreateWithFlags(&j, hipEventDisableTiming));
        auto body = [&](hipStream_t base) {
            DQ_HIP(hipEventRecord(fork, base));
            for (size_t li = 0; li < plan.size(); ++li) {
                if (plan[li].empty()) continue;
                hipStream_t ls = g_evo_lanes > 0 ? E->lanes[li] : E->classes[plan[li][0]].stream;
                DQ_HIP(hipStreamWaitEvent(ls, fork, 0));
                for (int ci : plan[li]) run_class_levels(*E, E->classes[ci], ls, renormalize != 0);
                DQ_HIP(hipEventRecord(joins[li], ls));
                DQ_HIP(hipStreamWaitEvent(base, joins[li], 0));
            }
        };
        if (use_graph) {
            std::vector<int64_t> sig = graph_signature(*E, renormalize != 0);
            if (E->exec && E->graph_sig != sig) {
                (void)hipGraphExecDestroy(E->exec); E->exec = nullptr;
                (void)hipGraphDestroy(E->graph); E->graph = nullptr;
            }
            if (!E->exec) {
                hipStream_t cap;
                DQ_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
                DQ_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
                try {
                    body(cap);
                } catch (...) {
                    hipGraph_t g = nullptr;
                    (void)hipStreamEndCapture(cap, &g);
                    if (g) (void)hipGraphDestroy(g);
                    (void)hipStreamDestroy(cap);
                    throw;
                }
                DQ_HIP(hipStreamEndCapture(cap, &E->graph));
                DQ_HIP(hipStreamDestroy(cap));
                DQ_HIP(hipGraphInstantiate(&E->exec, E->graph, nullptr, nullptr, 0));
                E->graph_sig = std::move(sig);
            }
            DQ_HIP(hipEventRecord(e0, s0));
            DQ_HIP(hipGraphLaunch(E->exec, s0));
            DQ_HIP(hipEventRecord(e1, s0));
        } else {
            DQ_HIP(hipEventRecord(e0, s0));
            body(s0);
            DQ_HIP(hipEventRecord(e1, s0));
        }
        DQ_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        DQ_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms_out) *ms_out = ms;
        for (auto &j : joins) (void)hipEventDestroy(j);
        (void)hipEventDestroy(fork);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    });
}

// A-posteriori truncation estimates from the current (initial) states: est[j] = 2^{s_j} ||t_{m+1}|| /
// (1 - ||t_{m+1}|| / ||t_m||) with the largest value over the job's states (infinity when the terms do not shrink).
// Must run before dq_evo_run (it reads the states as the initial ones; it does not change them).
extern "C" int dq_evo_probe(dq_evo *E, double *est) {
    return guard([&] {
        if (!E || (E->n_jobs && !est)) throw ArgError("dq_evo_probe: bad arguments");
        // every class's terms first, each class on its own stream (they run concurrently), then the reads
        for (auto &c : E->classes) {
            const int nj = c.jobs_n();
            hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(256), 0, c.stream, (int64_t)nj * E->n_states * 2,
                               c.d_probe.as<double>());
            int kmax = 0;
            for (auto &p : c.params) kmax = std::max(kmax, p.m + 1);
            dispatch(E->value_precision, E->n_states, [&](auto tv, auto ns) {
                using TV = decltype(tv);
                constexpr int NS = decltype(ns)::value;
                for (int k = 1; k <= kmax; ++k) launch_probe<TV, NS>(c, k, c.stream);
            });
        }
        for (auto &c : E->classes) {
            const int nj = c.jobs_n();
            std::vector<double> h(static_cast<size_t>(nj) * E->n_states * 2);
            DQ_HIP(hipMemcpyAsync(h.data(), c.d_probe.p, h.size() * 8, hipMemcpyDeviceToHost, c.stream));
            DQ_HIP(hipStreamSynchronize(c.stream));
            for (int q = 0; q < nj; ++q) {
                double worst = 0.0;
                for (int st = 0; st < E->n_states; ++st) {
                    const double tm = std::sqrt(h[(static_cast<size_t>(q) * E->n_states + st) * 2 + 0]);
                    const double t1 = std::sqrt(h[(static_cast<size_t>(q) * E->n_states + st) * 2 + 1]);
                    const double ratio = tm > 0.0 ? t1 / tm : 0.0;
                    const double e = ratio < 1.0 ? t1 / (1.0 - ratio) : INFINITY;
                    worst = std::max(worst, e * std::ldexp(1.0, c.params[q].s));
                }
                est[c.jobs[q]] = worst;
            }
        }
    });
}
