// The row-tile core shared by SpMV, the fused Taylor-action terms and the batched evolution.
//
// A workgroup of 256 lanes owns the rows [tile, tile + 256*RPT); lane l holds rows tile + l + 256*j, j < RPT, so each
// of the RPT loads of a wave covers 64 consecutive rows: one aligned line per diagonal and one contiguous window of x.
// The diagonal loop is uniform across the workgroup; metadata reads are scalar loads. A diagonal that misses the tile
// (or the column band [col_lo, col_hi)) is skipped by the whole workgroup.
//
// LDS x window (HALO > 0): the workgroup first copies x[tile - HALO, tile + TILE + HALO) into LDS (zero outside the
// matrix columns); every diagonal with |offset| <= HALO then reads its shifted x slice from LDS instead of memory.
// Neighbouring diagonals (a band, or the +-2^k offsets of a qubit Hamiltonian) share that window, so x crosses the
// memory hierarchy once per tile instead of once per near diagonal.
#pragma once

#include "dq_common.hpp"

namespace dq {

template <typename T>
struct Acc {
    T s = T(0);
    __device__ inline void add(T v) { s += v; }
    __device__ inline T get() const { return s; }
};

template <>
struct Acc<float> {  // Kahan-compensated FP32 sum
    float s = 0.0f, c = 0.0f;
    __device__ inline void add(float v) {
        const float y = v - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ inline float get() const { return s; }
};

// Complex product (a_r + i a_i)(x_r + i x_i) with one fixed rounding sequence -- re = fma(a_r, x_r, -(a_i x_i)),
// im = fma(a_r, x_i, a_i x_r) -- so every kernel instantiation and schedule (tile size, load order) rounds each term
// identically; left to itself the compiler may contract a*b - c*d either way round.
template <typename T>
__device__ inline void cmul(T a_r, T a_i, T x_r, T x_i, T &re, T &im) {
    re = __builtin_fma(a_r, x_r, -(a_i * x_i));
    im = __builtin_fma(a_r, x_i, a_i * x_r);
}
template <>
__device__ inline void cmul<float>(float a_r, float a_i, float x_r, float x_i, float &re, float &im) {
    re = __builtin_fmaf(a_r, x_r, -(a_i * x_i));
    im = __builtin_fmaf(a_r, x_i, a_i * x_r);
}

// Accumulate the complex product into a row's (real, imaginary) sums.
template <typename T>
__device__ inline void cmac(Acc<T> &sr, Acc<T> &si, T a_r, T a_i, T x_r, T x_i) {
    T re, im;
    cmul(a_r, a_i, x_r, x_i, re, im);
    sr.add(re);
    si.add(im);
}

// Loads of the matrix values: plain, or non-temporal (the values are streamed once, so they need not displace x from
// L2 and the Infinity Cache).
template <typename T, bool NT>
__device__ inline T ld_val(const T *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

// Row tile of workgroup b. xcd = 0: tile b. xcd = 1: the workgroups of one XCD (b mod 8: the hardware hands
// consecutive workgroups to the 8 XCDs in turn) take one contiguous eighth of the tiles, [x*per, (x+1)*per), so an
// XCD's L2 serves the x rows its neighbouring diagonals share; the grid is then 8*per workgroups and the surplus ones
// (tile >= ntiles) have no work.
__device__ inline int64_t tile_of_block(int64_t b, int64_t ntiles, int xcd) {
    if (!xcd) return b;
    const int64_t per = (ntiles + 7) >> 3;
    return (b & 7) * per + (b >> 3);
}

inline int64_t grid_for_tiles(int64_t ntiles, int xcd) { return xcd ? 8 * ((ntiles + 7) >> 3) : ntiles; }

// Copy x[tile - HALO, tile + TILE + HALO) into LDS (sr, si), zero outside [0, cols), then synchronise the workgroup.
template <typename T, int RPT, int HALO>
__device__ inline void stage_x_window(int64_t tile, int64_t cols, const T *__restrict__ xr, const T *__restrict__ xi,
                                      T *__restrict__ sr, T *__restrict__ si) {
    constexpr int W = 256 * RPT + 2 * HALO;
    const int64_t ws = tile - HALO;
    for (int e = threadIdx.x; e < W; e += 256) {
        const int64_t c = ws + e;
        T a = T(0), b = T(0);
        if (c >= 0 && c < cols) {
            a = xr[c];
            b = xi[c];
        }
        sr[e] = a;
        si[e] = b;
    }
    __syncthreads();
}

// Zero-block map of a DIA layout: bit b of word [job][diagonal][b / 32] is set when some value of the diagonal in
// rows [b * BR, (b + 1) * BR) is nonzero (a NaN counts as nonzero). One workgroup per (word, diagonal, job), the
// diagonals strided over gridDim.y (at most 65535); the values of job j start at j * total. A block whose bit is
// clear adds nothing to any row sum, so the row-tile loops can skip its value loads and x gathers: a qubit
// Hamiltonian's +-2^k flip diagonals are zero on alternate runs of 2^k rows (three runs in four for a flip-flop
// term), half or more of what they store.
template <typename T, int BR>
__global__ __launch_bounds__(256) void zero_blocks_kernel(int64_t rows, int n, const DiagMeta *__restrict__ meta,
                                                          int64_t total, const T *__restrict__ are,
                                                          const T *__restrict__ aim, int64_t zwords,
                                                          uint32_t *__restrict__ zmask) {
    __shared__ int live[32];
    const int64_t w = blockIdx.x;
    const int j = blockIdx.z;
    const int64_t r0 = w * 32 * BR;
    for (int d = blockIdx.y; d < n; d += gridDim.y) {  // uniform
        if (threadIdx.x < 32) live[threadIdx.x] = 0;
        __syncthreads();
        const DiagMeta md = meta[d];
        const T *pr = are + static_cast<int64_t>(j) * total + md.rowbase;
        const T *pi = aim + static_cast<int64_t>(j) * total + md.rowbase;
        for (int e = threadIdx.x; e < 32 * BR; e += 256) {
            const int64_t r = r0 + e;
            if (r >= md.rlo && r < md.rhi && r < rows && !(pr[r] == T(0) && pi[r] == T(0))) live[e / BR] = 1;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t m = 0;
            for (int b = 0; b < 32; ++b)
                if (live[b]) m |= 1u << b;
            zmask[(static_cast<int64_t>(j) * n + d) * zwords + w] = m;
        }
        __syncthreads();  // live[] is cleared for the next diagonal
    }
}

// Grid rows for a zero-block map over n diagonals.
inline unsigned zero_map_grid_y(int n) { return static_cast<unsigned>(n < 65535 ? n : 65535); }

inline int64_t zero_block_words(int64_t rows, int block_rows) {
    return ((rows + block_rows - 1) / block_rows + 31) / 32;
}

// PIPE = true: each diagonal's 4*RPT loads are issued together (full tiles unconditionally, boundary tiles with
// clamped addresses), the schedule that keeps the most bytes in flight. Every schedule adds exactly the products of
// the rows on the diagonal, in slot order, so a row's sum has the same bits whatever the schedule (a compensated
// FP32 sum is not left unchanged by adding a zero). PIPE = false: per-row predicated loads, which
// the compiler issues row after row -- fewer loads in flight per wave but half the registers, so twice the waves;
// with scattered offsets (no x reuse between diagonals) that variant can fill L2 faster.
//
// With HALO > 0 the caller has staged the x window in (sr, si) with stage_x_window; diagonals with |offset| <= HALO
// read x from there (row r of the tile at LDS index r - tile + HALO + offset), the others from memory.
//
// ZM = true: zm is the layout's zero-block map at 256-row blocks (zero_blocks_kernel<T, 256>, zwords words per
// diagonal; tiles start at a multiple of 256 * RPT rows) and the rows of a block whose values are all zero are
// skipped -- the same blocks under every RPT, so every schedule still adds the same products.
template <typename T, int RPT, bool NT = false, bool PIPE = true, int HALO = 0, bool ZM = false>
__device__ inline void dia_rows_tile(int64_t tile, int64_t tile_end, int n, const DiagMeta *__restrict__ meta,
                                     const T *__restrict__ are, const T *__restrict__ aim, const T *__restrict__ xr,
                                     const T *__restrict__ xi, int64_t col_lo, int64_t col_hi, Acc<T> (&ar)[RPT],
                                     Acc<T> (&ai)[RPT], const T *__restrict__ sr = nullptr,
                                     const T *__restrict__ si = nullptr, const uint32_t *__restrict__ zm = nullptr,
                                     int64_t zwords = 0) {
    constexpr int TILE = 256 * RPT;
    static_assert(!ZM || (RPT <= 16 && 32 % RPT == 0), "a tile's block bits must sit in one map word");
    for (int k = 0; k < n; ++k) {
        const DiagMeta m = meta[k];
        const int64_t lo = max(m.rlo, col_lo - m.off);
        const int64_t hi = min(m.rhi, col_hi - m.off);
        if (hi <= tile || lo >= tile_end) continue;  // uniform: the diagonal misses this tile
        uint32_t live = (1u << RPT) - 1;             // bit j: rows tile + 256 j .. + 255 hold a nonzero value
        if constexpr (ZM) {
            const int64_t b0 = tile / 256;
            live = (zm[static_cast<int64_t>(k) * zwords + (b0 >> 5)] >> (b0 & 31)) & ((1u << RPT) - 1);
            if (!live) continue;  // uniform: the diagonal is all zero in this tile
        }
        const T *pr = are + m.rowbase;
        const T *pi = aim + m.rowbase;
        if (HALO > 0 && m.off >= -HALO && m.off <= HALO) {  // uniform: x from the LDS window
            const int base = HALO + m.off;
            T va_r[RPT], va_i[RPT], vx_r[RPT], vx_i[RPT];
            if (lo <= tile && hi >= tile + TILE) {
#pragma unroll
                for (int j = 0; j < RPT; ++j) {
                    if (ZM && !((live >> j) & 1u)) continue;  // uniform
                    const int64_t r = tile + threadIdx.x + 256 * j;
                    va_r[j] = ld_val<T, NT>(pr + r);
                    va_i[j] = ld_val<T, NT>(pi + r);
                    vx_r[j] = sr[threadIdx.x + 256 * j + base];
                    vx_i[j] = si[threadIdx.x + 256 * j + base];
                }
#pragma unroll
                for (int j = 0; j < RPT; ++j)
                    if (!ZM || ((live >> j) & 1u)) cmac(ar[j], ai[j], va_r[j], va_i[j], vx_r[j], vx_i[j]);
            } else {  // boundary tile: clamped rows stay inside the tile, so their LDS index stays in the window
#pragma unroll
                for (int j = 0; j < RPT; ++j) {
                    if (ZM && !((live >> j) & 1u)) continue;  // uniform
                    const int64_t r = tile + threadIdx.x + 256 * j;
                    const int64_t rc = r < lo ? lo : (r >= hi ? hi - 1 : r);
                    va_r[j] = ld_val<T, NT>(pr + rc);
                    va_i[j] = ld_val<T, NT>(pi + rc);
                    const int e = static_cast<int>(rc - tile) + base;
                    vx_r[j] = sr[e];
                    vx_i[j] = si[e];
                }
#pragma unroll
                for (int j = 0; j < RPT; ++j) {  // rows off the diagonal add nothing (not even a zero)
                    const int64_t r = tile + threadIdx.x + 256 * j;
                    if ((!ZM || ((live >> j) & 1u)) && r >= lo && r < hi)
                        cmac(ar[j], ai[j], va_r[j], va_i[j], vx_r[j], vx_i[j]);
                }
            }
            continue;
        }
        const T *qr = xr + m.off;
        const T *qi = xi + m.off;
        if constexpr (!PIPE) {
#pragma unroll
            for (int j = 0; j < RPT; ++j) {
                const int64_t r = tile + threadIdx.x + 256 * j;
                if ((!ZM || ((live >> j) & 1u)) && r >= lo && r < hi) {
                    const T a_r = ld_val<T, NT>(pr + r), a_i = ld_val<T, NT>(pi + r), x_r = qr[r], x_i = qi[r];
                    cmac(ar[j], ai[j], a_r, a_i, x_r, x_i);
                }
            }
            continue;
        }
        // All 4*RPT loads of this diagonal are issued before the first use, so a wave keeps them in flight together.
        T va_r[RPT], va_i[RPT], vx_r[RPT], vx_i[RPT];
        if (lo <= tile && hi >= tile + TILE) {  // uniform: every row of the tile is on the diagonal
#pragma unroll
            for (int j = 0; j < RPT; ++j) {
                if (ZM && !((live >> j) & 1u)) continue;  // uniform
                const int64_t r = tile + threadIdx.x + 256 * j;
                va_r[j] = ld_val<T, NT>(pr + r);
                va_i[j] = ld_val<T, NT>(pi + r);
                vx_r[j] = qr[r];
                vx_i[j] = qi[r];
            }
#pragma unroll
            for (int j = 0; j < RPT; ++j)
                if (!ZM || ((live >> j) & 1u)) cmac(ar[j], ai[j], va_r[j], va_i[j], vx_r[j], vx_i[j]);
        } else {  // boundary tile: clamp the loads to a valid row; rows off the diagonal add nothing (not even a zero)
#pragma unroll
            for (int j = 0; j < RPT; ++j) {
                if (ZM && !((live >> j) & 1u)) continue;  // uniform
                const int64_t r = tile + threadIdx.x + 256 * j;
                const int64_t rc = r < lo ? lo : (r >= hi ? hi - 1 : r);
                va_r[j] = ld_val<T, NT>(pr + rc);
                va_i[j] = ld_val<T, NT>(pi + rc);
                vx_r[j] = qr[rc];
                vx_i[j] = qi[rc];
            }
#pragma unroll
            for (int j = 0; j < RPT; ++j) {
                const int64_t r = tile + threadIdx.x + 256 * j;
                if ((!ZM || ((live >> j) & 1u)) && r >= lo && r < hi)
                    cmac(ar[j], ai[j], va_r[j], va_i[j], vx_r[j], vx_i[j]);
            }
        }
    }
}

}  // namespace dq
