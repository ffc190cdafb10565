// The arithmetic of one gate on one column of amplitudes, shared by the single-gate kernels (kron.hip) and the fused
// gate passes (gates.hip): x[0..g) in registers, each of the gate's diagonals adds its terms in slot order into the
// row sums ar / ai. Only static register indexes: a runtime offset is dispatched to a template instance per offset.
#pragma once

#include "dia_rows.hpp"

namespace dq {

// The terms of one diagonal with offset D for every row r of the column: static register indexes only.
template <typename T, int GM, int D>
__device__ inline void kron_diag_terms(int g, int rlo, int rhi, const T *vr, const T *vi, const T (&xr)[GM],
                                       const T (&xi)[GM], Acc<T> (&ar)[GM], Acc<T> (&ai)[GM]) {
#pragma unroll
    for (int r = 0; r < GM; ++r) {
        if (r + D >= 0 && r + D < GM) {
            if (r >= rlo && r < rhi && r < g && r + D < g) cmac(ar[r], ai[r], vr[r], vi[r], xr[r + D], xi[r + D]);
        }
    }
}

// Dispatch a runtime offset d in [-(GM - 1), GM - 1] to kron_diag_terms<D>.
template <typename T, int GM, int D>
__device__ inline void kron_diag_dispatch(int d, int g, int rlo, int rhi, const T *vr, const T *vi,
                                          const T (&xr)[GM], const T (&xi)[GM], Acc<T> (&ar)[GM], Acc<T> (&ai)[GM]) {
    if constexpr (D < GM) {
        if (d == D) kron_diag_terms<T, GM, D>(g, rlo, rhi, vr, vi, xr, xi, ar, ai);
        else kron_diag_dispatch<T, GM, D + 1>(d, g, rlo, rhi, vr, vi, xr, xi, ar, ai);
    }
}

}  // namespace dq
