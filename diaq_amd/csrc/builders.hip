// Building DIA Hamiltonians on the device: a Pauli sum H = sum_j c_j P_j over n qubits goes straight into the device
// layout, with no dense matrix and no host-side diagonal arrays (the reference builds a diaQMatrix on the host from a
// dense array or element by element, src/diaQMatrix.hpp:666-674, include/pythonWrappers.cpp:64).
//
// A Pauli string is given by its flip mask x (qubits acted on by X or Y) and its phase mask z (qubits acted on by Z or
// Y); qubit k is bit k of the basis index. P|c> = i^{popcount(x & z)} (-1)^{popcount(c & z)} |c xor x> (per qubit Y =
// i X Z), so H[r, c] collects c_j times that phase for every term with c = r xor x_j. The entry lies on diagonal
// d = c - r = sum over the flipped qubits of +-2^k, so a term with w flips can touch 2^w diagonals; the result keeps
// exactly that structural set (ascending; a diagonal whose terms cancel stays, all zero -- dq_prune drops it).
//
// One lane owns one row and adds the terms in order, so every element's sum has a fixed order and no two lanes write
// the same element.
#include <algorithm>
#include <vector>

#include "dq_common.hpp"

namespace dq {

__global__ void pauli_rows_kernel(int64_t N, int n_terms, const uint64_t *__restrict__ xm,
                                  const uint64_t *__restrict__ zm, const double *__restrict__ cre,
                                  const double *__restrict__ cim, int n_diags, const DiagMeta *__restrict__ meta,
                                  double *__restrict__ re, double *__restrict__ im) {
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < N;
         r += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        for (int j = 0; j < n_terms; ++j) {
            const uint64_t x = xm[j], z = zm[j];
            const int64_t c = static_cast<int64_t>(static_cast<uint64_t>(r) ^ x);
            const int64_t d = c - r;
            int lo = 0, hi = n_diags - 1;  // the slot of offset d (offsets ascend; d is always present)
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (meta[mid].off < d) lo = mid + 1;
                else hi = mid;
            }
            const int ny = __popcll(x & z) & 3;
            double pr = ny == 0 ? 1.0 : (ny == 2 ? -1.0 : 0.0);
            double pi = ny == 1 ? 1.0 : (ny == 3 ? -1.0 : 0.0);
            if (__popcll(static_cast<uint64_t>(c) & z) & 1) { pr = -pr; pi = -pi; }
            const int64_t idx = meta[lo].rowbase + r;
            re[idx] += cre[j] * pr - cim[j] * pi;
            im[idx] += cre[j] * pi + cim[j] * pr;
        }
    }
}

// The structural diagonal set of the sum: every +-2^k combination over each term's flipped qubits.
static std::vector<int32_t> pauli_offsets(int n_qubits, int n_terms, const uint64_t *xm) {
    std::vector<int64_t> all;
    for (int j = 0; j < n_terms; ++j) {
        std::vector<int> bits;
        for (int k = 0; k < n_qubits; ++k)
            if ((xm[j] >> k) & 1) bits.push_back(k);
        if (bits.size() > 20) throw ArgError("pauli_sum_device: a term flips more than 20 qubits (2^20+ diagonals)");
        const uint64_t combos = uint64_t(1) << bits.size();
        for (uint64_t s = 0; s < combos; ++s) {
            int64_t d = 0;
            for (size_t b = 0; b < bits.size(); ++b) d += ((s >> b) & 1) ? -(int64_t(1) << bits[b]) : (int64_t(1) << bits[b]);
            all.push_back(d);
        }
        if (all.size() > (size_t(1) << 26)) {
            std::sort(all.begin(), all.end());
            all.erase(std::unique(all.begin(), all.end()), all.end());
        }
    }
    std::sort(all.begin(), all.end());
    all.erase(std::unique(all.begin(), all.end()), all.end());
    if (all.size() > (size_t(1) << 22)) throw ArgError("pauli_sum_device: more than 2^22 diagonals");
    return std::vector<int32_t>(all.begin(), all.end());
}

}  // namespace dq

using namespace dq;

extern "C" int dq_devmat_from_pauli_sum(int n_qubits, int n_terms, const uint64_t *x_masks, const uint64_t *z_masks,
                                        const double *coef_re, const double *coef_im, int precision,
                                        uintptr_t stream, dq_devmat **out) {
    return guard([&] {
        if (!out) throw ArgError("out is NULL");
        if (n_qubits < 1 || n_qubits > 31) throw ArgError("pauli_sum_device: n_qubits must be in [1, 31]");
        if (n_terms < 0 || (n_terms && (!x_masks || !z_masks || !coef_re || !coef_im)))
            throw ArgError("pauli_sum_device: bad term arrays");
        if (precision != DQ_F64 && precision != DQ_F32) throw ArgError("precision must be DQ_F64 or DQ_F32");
        const uint64_t lim = uint64_t(1) << n_qubits;
        for (int j = 0; j < n_terms; ++j)
            if (x_masks[j] >= lim || z_masks[j] >= lim) throw ArgError("pauli_sum_device: mask beyond n_qubits");
        hipStream_t s = as_stream(stream);
        const int64_t N = static_cast<int64_t>(lim);
        const std::vector<int32_t> offs = pauli_offsets(n_qubits, n_terms, x_masks);
        auto M = make_devmat(static_cast<uint32_t>(N), static_cast<uint32_t>(N), offs.data(),
                             static_cast<int>(offs.size()), DQ_F64, s);  // zeroed
        if (n_terms && !offs.empty()) {
            DevBuf dx(8 * n_terms), dz(8 * n_terms), dr(8 * n_terms), di(8 * n_terms);
            DQ_HIP(hipMemcpyAsync(dx.p, x_masks, dx.bytes, hipMemcpyHostToDevice, s));
            DQ_HIP(hipMemcpyAsync(dz.p, z_masks, dz.bytes, hipMemcpyHostToDevice, s));
            DQ_HIP(hipMemcpyAsync(dr.p, coef_re, dr.bytes, hipMemcpyHostToDevice, s));
            DQ_HIP(hipMemcpyAsync(di.p, coef_im, di.bytes, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(pauli_rows_kernel, dim3(blocks_for(N, 256, 1 << 16)), dim3(256), 0, s, N, n_terms,
                               dx.as<uint64_t>(), dz.as<uint64_t>(), dr.as<double>(), di.as<double>(), M->n(),
                               M->d_meta.as<DiagMeta>(), M->re.as<double>(), M->im.as<double>());
            DQ_LAUNCH_CHECK();
            DQ_HIP(hipStreamSynchronize(s));  // the term arrays are freed on return
        }
        if (precision == DQ_F32) {
            dq_devmat *c = nullptr;
            if (dq_devmat_convert(M.get(), DQ_F32, &c) != DQ_OK) throw HipError(dq_last_error());
            *out = c;
        } else {
            *out = M.release();
        }
    });
}
