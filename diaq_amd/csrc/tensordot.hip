// tensordot on the device: the contraction np.tensordot(A, B, axes) of two DIA-stored tensors, as one device product.
//
// Reference: tensordot (src/tensordotKernels.cpp:21-74), with transpose (src/transposeKernels.cpp:19-110) and
// matrix_reshape (src/reshapeKernels.cpp:21-73) on the host and product() (the GPU product with the zero-diagonal
// prune, src/spGEMM.cpp:516-534) in between.
//
// A DIA matrix with tensor shape (n_0, ..., n_{l-1}) holds the tensor's entry of row-major flat index f at matrix
// position (f / cols, f % cols). The contraction permutes A's axes to (free axes, contracted axes) and reads the
// result as a free_A x sum_A matrix, B's to (contracted, free) as sum_B x free_B, and multiplies. Permuting and
// reshaping only move stored entries, so here one kernel maps each stored nonzero of A (row r, column c) straight to
// its place in the permuted matrix: f = r cols + c, the multi-index of f, its permuted flat index f', and
// (r', c') = (f' / cols', f' % cols') on diagonal c' - r'. A first pass marks the diagonals that receive a nonzero,
// the host lays them out in ascending order (as the reference's rebuild does), and a second pass writes the values.
// Entries that are exactly zero are dropped, as the reference's transpose and reshape drop them.
#include <algorithm>
#include <string>

#include "dq_common.hpp"

namespace dq {

constexpr int TD_MAX_RANK = 32;

// Per input axis m of the source tensor: its extent, its row-major stride in the source, and the stride its index
// gets in the permuted tensor's row-major order.
struct AxisMap {
    int rank;
    int64_t dim[TD_MAX_RANK];
    int64_t src_stride[TD_MAX_RANK];
    int64_t dst_stride[TD_MAX_RANK];
};

__device__ inline int64_t permuted_flat(int64_t f, const AxisMap &am) {
    int64_t g = 0;
    for (int m = 0; m < am.rank; ++m) g += ((f / am.src_stride[m]) % am.dim[m]) * am.dst_stride[m];
    return g;
}

// Pass 1: used[d' + rows' - 1] = 1 for every diagonal d' of the permuted matrix that receives a nonzero. One
// workgroup row (grid y) per diagonal of A, grid-strided over its rows.
template <typename T>
__global__ __launch_bounds__(256) void permute_mark_kernel(int n, const DiagMeta *__restrict__ meta,
                                                           const T *__restrict__ re, const T *__restrict__ im,
                                                           int64_t cols, int64_t cols2, int64_t rows2, AxisMap am,
                                                           uint8_t *__restrict__ used) {
    for (int k = blockIdx.y; k < n; k += gridDim.y) {
        const DiagMeta m = meta[k];
        for (int64_t r = m.rlo + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < m.rhi;
             r += (int64_t)gridDim.x * blockDim.x) {
            if (re[m.rowbase + r] == T(0) && im[m.rowbase + r] == T(0)) continue;
            const int64_t g = permuted_flat(r * cols + r + m.off, am);
            const int64_t r2 = g / cols2, c2 = g % cols2;
            used[c2 - r2 + rows2 - 1] = 1;
        }
    }
}

// Pass 2: write each nonzero into the permuted matrix C (slot_of[d' + rows' - 1] is the slot of diagonal d').
template <typename T>
__global__ __launch_bounds__(256) void permute_scatter_kernel(int n, const DiagMeta *__restrict__ meta,
                                                              const T *__restrict__ re, const T *__restrict__ im,
                                                              int64_t cols, int64_t cols2, int64_t rows2, AxisMap am,
                                                              const int32_t *__restrict__ slot_of,
                                                              const DiagMeta *__restrict__ cmeta, T *__restrict__ cre,
                                                              T *__restrict__ cim) {
    for (int k = blockIdx.y; k < n; k += gridDim.y) {
        const DiagMeta m = meta[k];
        for (int64_t r = m.rlo + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < m.rhi;
             r += (int64_t)gridDim.x * blockDim.x) {
            const T vr = re[m.rowbase + r], vi = im[m.rowbase + r];
            if (vr == T(0) && vi == T(0)) continue;
            const int64_t g = permuted_flat(r * cols + r + m.off, am);
            const int64_t r2 = g / cols2, c2 = g % cols2;
            const int32_t slot = slot_of[c2 - r2 + rows2 - 1];
            const int64_t at = cmeta[slot].rowbase + r2;
            cre[at] = vr;
            cim[at] = vi;
        }
    }
}

// A's entries (tensor shape `shape`, its row-major order over A's matrix positions) with the axes permuted to `order`
// (result axis j is source axis order[j]), as a rows2 x cols2 DIA matrix of the nonzeros.
static std::unique_ptr<dq_devmat> permute_reshape(const dq_devmat &A, const std::vector<int64_t> &shape,
                                                  const std::vector<int> &order, int64_t rows2, int64_t cols2,
                                                  hipStream_t s) {
    const int l = static_cast<int>(shape.size());
    if (l > TD_MAX_RANK) throw ArgError("tensordot: tensors of rank above 32 are not supported");
    if (rows2 > 0xffffffffLL || cols2 > 0xffffffffLL)
        throw ArgError("tensordot: a permuted operand has more than 2^32 - 1 rows or columns");
    AxisMap am{};
    am.rank = l;
    int64_t st = 1;
    for (int m = l - 1; m >= 0; --m) {
        am.dim[m] = shape[m];
        am.src_stride[m] = st;
        st *= shape[m];
    }
    int64_t dst = 1;
    for (int j = l - 1; j >= 0; --j) {
        am.dst_stride[order[j]] = dst;
        dst *= shape[order[j]];
    }
    const int64_t n_slots = rows2 + cols2 - 1;
    const int n = A.n();
    const bool f64 = A.precision == DQ_F64;
    std::vector<int32_t> diags;
    DevBuf used(n_slots > 0 ? static_cast<size_t>(n_slots) : 1);
    if (n_slots > 0) zero_bytes(used.p, static_cast<size_t>(n_slots), s);
    const dim3 grid(256, static_cast<unsigned>(std::max(1, std::min(n, 65535))));
    if (n > 0 && n_slots > 0) {
        if (f64)
            hipLaunchKernelGGL(permute_mark_kernel<double>, grid, dim3(256), 0, s, n, A.d_meta.as<DiagMeta>(),
                               A.re.as<double>(), A.im.as<double>(), (int64_t)A.cols, cols2, rows2, am,
                               used.as<uint8_t>());
        else
            hipLaunchKernelGGL(permute_mark_kernel<float>, grid, dim3(256), 0, s, n, A.d_meta.as<DiagMeta>(),
                               A.re.as<float>(), A.im.as<float>(), (int64_t)A.cols, cols2, rows2, am,
                               used.as<uint8_t>());
        DQ_LAUNCH_CHECK();
        std::vector<uint8_t> h(static_cast<size_t>(n_slots));
        DQ_HIP(hipMemcpyAsync(h.data(), used.p, h.size(), hipMemcpyDeviceToHost, s));
        DQ_HIP(hipStreamSynchronize(s));
        for (int64_t i = 0; i < n_slots; ++i)
            if (h[static_cast<size_t>(i)]) diags.push_back(static_cast<int32_t>(i - (rows2 - 1)));
    }
    auto C = make_devmat(static_cast<uint32_t>(rows2), static_cast<uint32_t>(cols2), diags.data(),
                         static_cast<int>(diags.size()), A.precision, s);
    if (diags.empty()) return C;
    std::vector<int32_t> slot_of(static_cast<size_t>(n_slots), -1);
    for (size_t k = 0; k < diags.size(); ++k) slot_of[static_cast<size_t>(diags[k] + rows2 - 1)] = static_cast<int32_t>(k);
    DevBuf d_slot(sizeof(int32_t) * slot_of.size());
    DQ_HIP(hipMemcpyAsync(d_slot.p, slot_of.data(), d_slot.bytes, hipMemcpyHostToDevice, s));
    if (f64)
        hipLaunchKernelGGL(permute_scatter_kernel<double>, grid, dim3(256), 0, s, n, A.d_meta.as<DiagMeta>(),
                           A.re.as<double>(), A.im.as<double>(), (int64_t)A.cols, cols2, rows2, am,
                           d_slot.as<int32_t>(), C->d_meta.as<DiagMeta>(), C->re.as<double>(), C->im.as<double>());
    else
        hipLaunchKernelGGL(permute_scatter_kernel<float>, grid, dim3(256), 0, s, n, A.d_meta.as<DiagMeta>(),
                           A.re.as<float>(), A.im.as<float>(), (int64_t)A.cols, cols2, rows2, am,
                           d_slot.as<int32_t>(), C->d_meta.as<DiagMeta>(), C->re.as<float>(), C->im.as<float>());
    DQ_LAUNCH_CHECK();
    DQ_HIP(hipStreamSynchronize(s));  // the slot map is freed on return
    return C;
}

static std::vector<int64_t> tensor_shape_of(const dq_devmat &M, int rank, const uint32_t *shape, const char *who) {
    std::vector<int64_t> sh;
    if (rank == 0) {  // no tensor shape given: the matrix itself
        sh = {static_cast<int64_t>(M.rows), static_cast<int64_t>(M.cols)};
    } else {
        if (!shape) throw ArgError(std::string("tensordot: NULL shape for ") + who);
        int64_t prod = 1;
        for (int i = 0; i < rank; ++i) {
            sh.push_back(shape[i]);
            prod *= shape[i];
        }
        if (prod != static_cast<int64_t>(M.rows) * static_cast<int64_t>(M.cols))
            throw ArgError(std::string("tensordot: the tensor shape of ") + who + " does not match its matrix size");
    }
    return sh;
}

}  // namespace dq

using namespace dq;

extern "C" int dq_tensordot(const dq_devmat *A, int rank_a, const uint32_t *shape_a, const dq_devmat *B, int rank_b,
                            const uint32_t *shape_b, int n_axes, const uint32_t *axes_a, const uint32_t *axes_b,
                            dq_devmat **out, uint32_t *out_shape, int *out_rank) {
    return guard([&] {
        if (!A || !B || !out) throw ArgError("NULL argument");
        if (n_axes < 0 || (n_axes > 0 && (!axes_a || !axes_b)))
            throw ArgError("tensordot: axes must be two lists of the same length");
        if (A->precision != B->precision) throw ArgError("tensordot: operands have different precisions");
        const std::vector<int64_t> sa = tensor_shape_of(*A, rank_a, shape_a, "A");
        const std::vector<int64_t> sb = tensor_shape_of(*B, rank_b, shape_b, "B");
        const int la = static_cast<int>(sa.size()), lb = static_cast<int>(sb.size());
        std::vector<char> ca(la, 0), cb(lb, 0);
        for (int k = 0; k < n_axes; ++k) {
            if (axes_a[k] >= static_cast<uint32_t>(la)) throw ArgError("tensordot: axis out of range for A");
            if (axes_b[k] >= static_cast<uint32_t>(lb)) throw ArgError("tensordot: axis out of range for B");
            ca[axes_a[k]] = 1;
            cb[axes_b[k]] = 1;
        }
        // A: free axes in order, then the contracted axes in the order given; B: contracted axes, then free axes
        std::vector<int> order_a, order_b;
        std::vector<int64_t> result_shape;
        int64_t free_a = 1, free_b = 1, sum_a = 1, sum_b = 1;
        for (int i = 0; i < la; ++i)
            if (!ca[i]) { order_a.push_back(i); result_shape.push_back(sa[i]); free_a *= sa[i]; }
        for (int k = 0; k < n_axes; ++k) { order_a.push_back(static_cast<int>(axes_a[k])); sum_a *= sa[axes_a[k]]; }
        for (int k = 0; k < n_axes; ++k) { order_b.push_back(static_cast<int>(axes_b[k])); sum_b *= sb[axes_b[k]]; }
        for (int i = 0; i < lb; ++i)
            if (!cb[i]) { order_b.push_back(i); result_shape.push_back(sb[i]); free_b *= sb[i]; }
        for (int k = 0; k < n_axes; ++k)
            if (sa[axes_a[k]] != sb[axes_b[k]]) throw ArgError("tensordot: contracted axes have different sizes");
        if (sum_a != sum_b) throw ArgError("tensordot: contracted sizes differ");
        if (static_cast<int>(order_a.size()) != la || static_cast<int>(order_b.size()) != lb)
            throw ArgError("tensordot: an axis is contracted twice");
        hipStream_t s = 0;
        auto Ap = permute_reshape(*A, sa, order_a, free_a, sum_a, s);
        auto Bp = permute_reshape(*B, sb, order_b, sum_b, free_b, s);
        auto C = product(*Ap, *Bp, 1.0, s, false);
        Ap.reset();
        Bp.reset();
        auto P = prune(*C, s);
        DQ_HIP(hipStreamSynchronize(s));
        if (result_shape.empty()) result_shape = {1, 1};
        if (out_rank) *out_rank = static_cast<int>(result_shape.size());
        if (out_shape)
            for (size_t i = 0; i < result_shape.size(); ++i) out_shape[i] = static_cast<uint32_t>(result_shape[i]);
        *out = P.release();
    });
}
