// Dense path for near-dense DIA products: the squarings of a matrix exponential whose result fills (nearly) every
// diagonal, where diagonal-pair enumeration degenerates into a dense matrix product done one pair at a time.
//
// A and B are expanded into dense row-major complex double matrices (the tiled conversion kernels of builders.hip),
// multiplied by the complex FP64 MFMA GEMM of gemm.hip (which also divides by the divisor), and C's diagonals --
// exactly the structural dry-run's offsets, every in-range dA + dB -- are gathered back. An FP32 matrix is widened
// to FP64 for the product and rounded once at the end. The values agree with the pair-by-pair sums to rounding (a
// different summation order); the offsets are the same. Used only inside pow and the matrix exponential (where no
// prune follows), never by the user-facing product, whose pair order matches the reference's.

#include <algorithm>
#include <climits>

#include "dq_common.hpp"

namespace dq {

static int g_prod_dense = 1;  // dq_set_tuning "product_dense": 0 = never, 1 = when it pays

bool dense_set_tuning(const std::string &key, int value) {
    if (key != "product_dense") return false;
    if (value < 0 || value > 1) throw ArgError("product_dense must be 0 or 1");
    g_prod_dense = value;
    return true;
}

// The dense form pays when the pair work (at most nA * nB pairs of up to min(M, N) elements) reaches a quarter of
// the dense product's M * K * N, and the dense operands stay small (every dimension <= 8192).
bool dense_product_pays(const dq_devmat &A, const dq_devmat &B) {
    if (!g_prod_dense) return false;
    const int64_t M = A.rows, K = A.cols, N = B.cols;
    if (M == 0 || K == 0 || N == 0 || M > 8192 || K > 8192 || N > 8192) return false;
    const double pair_work = static_cast<double>(A.n()) * B.n() * static_cast<double>(std::min(M, N));
    return pair_work >= 0.25 * static_cast<double>(M) * K * N;
}

namespace {
// Result offsets of A B: every in-range dA + dB, ascending (the structural dry-run without its pair lists). The
// result is a bitset over [lo, hi]; each of A's offsets ORs in B's offset bitset shifted by dA, a whole word at a
// time (nA x nB / 64 word operations: 8191 x 8191 offsets take ~1 M).
std::vector<int32_t> sum_offsets(const dq_devmat &A, const dq_devmat &B) {
    const int64_t lo = -static_cast<int64_t>(A.rows) + 1, hi = static_cast<int64_t>(B.cols) - 1;
    const int64_t range = hi - lo + 1;
    std::vector<int32_t> out;
    if (range <= 0 || A.n() == 0 || B.n() == 0) return out;
    // B's offsets as bits over [lo_b, lo_b + 64 * nbw)
    int64_t lo_b = INT64_MAX, hi_b = INT64_MIN;
    for (int32_t d : B.offs) { lo_b = std::min<int64_t>(lo_b, d); hi_b = std::max<int64_t>(hi_b, d); }
    const int64_t nbw = (hi_b - lo_b + 1 + 63) / 64;
    std::vector<uint64_t> bb(static_cast<size_t>(nbw), 0);
    for (int32_t d : B.offs) bb[(d - lo_b) / 64] |= uint64_t(1) << ((d - lo_b) % 64);
    // result bits, offset by `base` bits so that every shifted word lands inside the array: bit i of bb (offset
    // lo_b + i) goes to result bit da + lo_b + i - lo, which may fall outside [0, range) for out-of-range sums
    int64_t pmin = INT64_MAX, pmax = INT64_MIN;
    for (int32_t da : A.offs) {
        pmin = std::min<int64_t>(pmin, static_cast<int64_t>(da) + lo_b - lo);
        pmax = std::max<int64_t>(pmax, static_cast<int64_t>(da) + lo_b - lo);
    }
    const int64_t base = pmin < 0 ? ((-pmin + 63) / 64) * 64 : 0;
    const int64_t words = (range + 63) / 64;
    const int64_t total = std::max<int64_t>(base / 64 + words, (pmax + base) / 64 + nbw + 1) + 1;
    std::vector<uint64_t> res(static_cast<size_t>(total), 0);
    for (int32_t da : A.offs) {
        const int64_t shift = static_cast<int64_t>(da) + lo_b - lo + base;  // >= 0
        const int64_t w0 = shift / 64;
        const int sh = static_cast<int>(shift % 64);
        uint64_t *r = res.data() + w0;  // r[0 .. nbw] stay below total
        if (sh == 0) {
            for (int64_t w = 0; w < nbw; ++w) r[w] |= bb[w];
        } else {
            for (int64_t w = 0; w < nbw; ++w) {
                r[w] |= bb[w] << sh;
                r[w + 1] |= bb[w] >> (64 - sh);
            }
        }
    }
    // result bit j (offset lo + j) sits at bit base + j; keep j in [0, range)
    for (int64_t w = 0; w < words; ++w) {
        uint64_t x = res[static_cast<size_t>(w + base / 64)];
        if (w == words - 1 && (range % 64)) x &= (uint64_t(1) << (range % 64)) - 1;
        for (; x; x &= x - 1) out.push_back(static_cast<int32_t>(lo + w * 64 + __builtin_ctzll(x)));
    }
    return out;
}
}  // namespace

std::unique_ptr<dq_devmat> product_dense(const dq_devmat &A, const dq_devmat &B, double divisor, hipStream_t s) {
    if (A.cols != B.rows) throw ArgError("diaQMatrix product_cuda: dimension mismatch.");
    if (A.precision != B.precision) throw ArgError("product: operands have different precisions");
    const std::vector<int32_t> res = sum_offsets(A, B);
    // every stored element of C is written from the dense product; only the gaps between diagonals are zeroed
    auto C = make_devmat(A.rows, B.cols, res.data(), static_cast<int>(res.size()), A.precision, s, false);
    if (res.empty()) return C;
    zero_gaps(*C, s);
    const int64_t M = A.rows, K = A.cols, N = B.cols;
    const size_t cb = 2 * sizeof(double);  // an interleaved complex double
    DevBuf Ad(static_cast<size_t>(M * K) * cb), Cd(static_cast<size_t>(M * N) * cb);
    DevBuf Bd;
    dia_to_dense(A, Ad.as<double>(), K, s);
    const double *bd = Ad.as<double>();
    if (&A != &B) {  // a squaring reads one dense copy twice
        Bd = DevBuf(static_cast<size_t>(K * N) * cb);
        dia_to_dense(B, Bd.as<double>(), N, s);
        bd = Bd.as<double>();
    }
    zgemm_dense(M, N, K, Ad.as<double>(), bd, Cd.as<double>(), divisor, s);
    dense_to_dia(*C, Cd.as<double>(), N, s);
    DQ_HIP(hipStreamSynchronize(s));  // before the dense buffers go back to the cache
    return C;
}

}  // namespace dq
