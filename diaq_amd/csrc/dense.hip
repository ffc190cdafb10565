// Dense path for near-dense DIA products: the squarings of a matrix exponential whose result fills (nearly) every
// diagonal, where diagonal-pair enumeration degenerates into a dense matrix product done one pair at a time.
//
// A and B are scattered into dense row-major complex matrices, multiplied by the library GEMM on the matrix cores
// (rocBLAS zgemm / cgemm: FP64 / FP32 MFMA), and C's diagonals -- exactly the structural dry-run's offsets, every
// in-range dA + dB -- are gathered back, divided by the divisor. The values agree with the pair-by-pair sums to
// rounding (a different summation order); the offsets are the same. Used only inside pow and the matrix
// exponential (where no prune follows), never by the user-facing product, whose pair order matches the reference's.
#include <rocblas/rocblas.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <type_traits>

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

template <typename T>
__global__ void dia_to_dense_kernel(const DiagMeta *__restrict__ meta, const T *__restrict__ re,
                                    const T *__restrict__ im, T *__restrict__ D, int64_t ld) {
    const DiagMeta m = meta[blockIdx.y];
    const int64_t len = m.rhi - m.rlo;
    for (int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; p < len;
         p += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t r = m.rlo + p, c = r + m.off;
        D[2 * (r * ld + c)] = re[m.rowbase + r];
        D[2 * (r * ld + c) + 1] = im[m.rowbase + r];
    }
}

template <typename T>
__global__ void dense_to_dia_kernel(const DiagMeta *__restrict__ meta, const T *__restrict__ D, int64_t ld,
                                    T *__restrict__ re, T *__restrict__ im, double divisor) {
    const DiagMeta m = meta[blockIdx.y];
    const int64_t len = m.rhi - m.rlo;
    for (int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; p < len;
         p += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t r = m.rlo + p, c = r + m.off;
        T vr = D[2 * (r * ld + c)], vi = D[2 * (r * ld + c) + 1];
        if (divisor != 1.0) {
            vr = static_cast<T>(vr / divisor);
            vi = static_cast<T>(vi / divisor);
        }
        re[m.rowbase + r] = vr;
        im[m.rowbase + r] = vi;
    }
}

namespace {
rocblas_handle blas_handle() {
    static std::mutex mu;
    static std::map<int, rocblas_handle> *handles = new std::map<int, rocblas_handle>;  // per device, kept
    int dev = 0;
    DQ_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto it = handles->find(dev);
    if (it != handles->end()) return it->second;
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) throw HipError("rocblas_create_handle failed");
    (*handles)[dev] = h;
    return h;
}

void blas_check(rocblas_status st, const char *what) {
    if (st != rocblas_status_success)
        throw HipError(std::string(what) + " failed: " + rocblas_status_to_string(st));
}

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
    auto C = make_devmat(A.rows, B.cols, res.data(), static_cast<int>(res.size()), A.precision, s);  // zeroed
    if (res.empty()) return C;
    const int64_t M = A.rows, K = A.cols, N = B.cols;
    const size_t eb = 2 * static_cast<size_t>(elem_bytes(A.precision));  // complex element
    DevBuf Ad(static_cast<size_t>(M * K) * eb), Bd(static_cast<size_t>(K * N) * eb), Cd(static_cast<size_t>(M * N) * eb);
    zero_bytes(Ad.p, Ad.bytes, s);
    zero_bytes(Bd.p, Bd.bytes, s);
    rocblas_handle h = blas_handle();
    blas_check(rocblas_set_stream(h, s), "rocblas_set_stream");
    auto run = [&](auto tag) {
        using T = decltype(tag);
        if (A.n())
            hipLaunchKernelGGL(dia_to_dense_kernel<T>, dim3(std::max<unsigned>(1, blocks_for(std::min(M, K), 256, 64)),
                               static_cast<unsigned>(A.n())), dim3(256), 0, s, A.d_meta.as<DiagMeta>(), A.re.as<T>(),
                               A.im.as<T>(), Ad.as<T>(), K);
        if (B.n())
            hipLaunchKernelGGL(dia_to_dense_kernel<T>, dim3(std::max<unsigned>(1, blocks_for(std::min(K, N), 256, 64)),
                               static_cast<unsigned>(B.n())), dim3(256), 0, s, B.d_meta.as<DiagMeta>(), B.re.as<T>(),
                               B.im.as<T>(), Bd.as<T>(), N);
        DQ_LAUNCH_CHECK();
        // row-major C = A B is column-major C^T = B^T A^T
        if constexpr (std::is_same<T, double>::value) {
            const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0};
            blas_check(rocblas_zgemm(h, rocblas_operation_none, rocblas_operation_none, static_cast<rocblas_int>(N),
                                     static_cast<rocblas_int>(M), static_cast<rocblas_int>(K), &one,
                                     Bd.as<rocblas_double_complex>(), static_cast<rocblas_int>(N),
                                     Ad.as<rocblas_double_complex>(), static_cast<rocblas_int>(K), &zero,
                                     Cd.as<rocblas_double_complex>(), static_cast<rocblas_int>(N)),
                       "rocblas_zgemm");
        } else {
            const rocblas_float_complex one = {1.0f, 0.0f}, zero = {0.0f, 0.0f};
            blas_check(rocblas_cgemm(h, rocblas_operation_none, rocblas_operation_none, static_cast<rocblas_int>(N),
                                     static_cast<rocblas_int>(M), static_cast<rocblas_int>(K), &one,
                                     Bd.as<rocblas_float_complex>(), static_cast<rocblas_int>(N),
                                     Ad.as<rocblas_float_complex>(), static_cast<rocblas_int>(K), &zero,
                                     Cd.as<rocblas_float_complex>(), static_cast<rocblas_int>(N)),
                       "rocblas_cgemm");
        }
        hipLaunchKernelGGL(dense_to_dia_kernel<T>, dim3(std::max<unsigned>(1, blocks_for(std::min(M, N), 256, 64)),
                           static_cast<unsigned>(C->n())), dim3(256), 0, s, C->d_meta.as<DiagMeta>(), Cd.as<T>(), N,
                           C->re.as<T>(), C->im.as<T>(), divisor);
        DQ_LAUNCH_CHECK();
    };
    if (A.precision == DQ_F64) run(double{});
    else run(float{});
    DQ_HIP(hipStreamSynchronize(s));  // before the dense buffers go back to the cache
    return C;
}

}  // namespace dq
