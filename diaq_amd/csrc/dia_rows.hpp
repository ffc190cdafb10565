// The row-tile core shared by SpMV, the fused Taylor-action terms and the batched evolution.
//
// A workgroup of 256 lanes owns the rows [tile, tile + 256*RPT); lane l holds rows tile + l + 256*j, j < RPT, so each
// of the RPT loads of a wave covers 64 consecutive rows: one aligned line per diagonal and one contiguous window of x.
// The diagonal loop is uniform across the workgroup; metadata reads are scalar loads. A diagonal that misses the tile
// (or the column band [col_lo, col_hi)) is skipped by the whole workgroup.
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

template <typename T, int RPT>
__device__ inline void dia_rows_tile(int64_t tile, int64_t tile_end, int n, const DiagMeta *__restrict__ meta,
                                     const T *__restrict__ are, const T *__restrict__ aim, const T *__restrict__ xr,
                                     const T *__restrict__ xi, int64_t col_lo, int64_t col_hi, Acc<T> (&ar)[RPT],
                                     Acc<T> (&ai)[RPT]) {
    for (int k = 0; k < n; ++k) {
        const DiagMeta m = meta[k];
        const int64_t lo = max(m.rlo, col_lo - m.off);
        const int64_t hi = min(m.rhi, col_hi - m.off);
        if (hi <= tile || lo >= tile_end) continue;  // uniform: the diagonal misses this tile
        const T *pr = are + m.rowbase;
        const T *pi = aim + m.rowbase;
        const T *qr = xr + m.off;
        const T *qi = xi + m.off;
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const int64_t r = tile + threadIdx.x + 256 * j;
            if (r >= lo && r < hi) {
                const T a_r = pr[r], a_i = pi[r], x_r = qr[r], x_i = qi[r];
                ar[j].add(a_r * x_r - a_i * x_i);
                ai[j].add(a_r * x_i + a_i * x_r);
            }
        }
    }
}

}  // namespace dq
