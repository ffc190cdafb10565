// SpME on gfx950: exp(A) and exp(-i t A) by Taylor series with scaling and squaring, all on the device.
//
// Reference: expm_taylor_scaled_cuda / expm_taylor_cuda / expm_negative_taylor_cuda /
// expm_negative_taylor_scaled_cuda (src/spME_gpu.cu:10-294) with the host helpers they call between device products
// (estimate_1norm, copy_and_scale_by_imag, divide_by_int, add, norm: src/diaQMatrix.hpp). The reference uploads both
// operands and downloads the product for every Taylor term and runs the helpers on the host; here every matrix stays
// in HBM: the 1/i of divide_by_int is applied in the product's epilogue, the union-add and the Frobenius norm are
// device kernels, and only the norm (one double per term) comes back to decide the stop. For an FP32 matrix the
// 1-norm is summed in FP64, the Taylor terms run in FP32 (compensated row sums) and the squarings in FP64.
#include <cmath>
#include <cstdio>

#include "dq_common.hpp"

namespace dq {

std::unique_ptr<dq_devmat> matrix_pow(const dq_devmat &A, int64_t n, hipStream_t s);

// One-diagonal fast path (src/spME_gpu.cu:99-124, :170-197): the identity with its first len entries replaced by
// exp(a + ib) (positive form) or e^{tb}(cos ta - i sin ta) (negative form), a + ib being the lone diagonal's values
// whatever its offset (the reference writes them onto the main diagonal).
template <typename T>
__global__ void single_diag_exp_kernel(int64_t n, const T *__restrict__ ar, const T *__restrict__ ai,
                                       T *__restrict__ ur, T *__restrict__ ui, double t, int negative) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double a = static_cast<double>(ar[i]), b = static_cast<double>(ai[i]);
        if (negative) {
            const double e = exp(t * b);
            ur[i] = static_cast<T>(e * cos(t * a));
            ui[i] = static_cast<T>(-e * sin(t * a));
        } else {
            const double e = exp(a);
            ur[i] = static_cast<T>(e * cos(b));
            ui[i] = static_cast<T>(e * sin(b));
        }
    }
}

static std::unique_ptr<dq_devmat> single_diag_exp(const dq_devmat &A, double t, bool negative, hipStream_t s) {
    const int32_t zero = 0;
    auto U = make_devmat(A.rows, A.rows, &zero, 1, A.precision, s);
    fill_identity(*U, s);
    const int64_t n = std::min<int64_t>(A.len[0], U->len[0]);
    if (n > 0) {
        const unsigned g = blocks_for(n, 256, 4096);
        const size_t eb = elem_bytes(A.precision);
        if (A.precision == DQ_F64)
            hipLaunchKernelGGL(single_diag_exp_kernel<double>, dim3(g), dim3(256), 0, s, n,
                               A.re.as<double>() + A.start[0], A.im.as<double>() + A.start[0],
                               U->re.as<double>() + U->start[0], U->im.as<double>() + U->start[0], t, negative ? 1 : 0);
        else
            hipLaunchKernelGGL(single_diag_exp_kernel<float>, dim3(g), dim3(256), 0, s, n,
                               A.re.as<float>() + A.start[0], A.im.as<float>() + A.start[0],
                               U->re.as<float>() + U->start[0], U->im.as<float>() + U->start[0], t, negative ? 1 : 0);
        (void)eb;
        DQ_LAUNCH_CHECK();
    }
    return U;
}

// result = I + sum_{i>=1} X^i / i!, stopping after the first term whose Frobenius norm is below tol (checked after
// adding it) or after max_iterations terms; products are not pruned (src/spME_gpu.cu:126-135, :199-209).
static std::unique_ptr<dq_devmat> taylor(const dq_devmat &X, int max_iterations, double tol, const char *who,
                                         int *terms, hipStream_t s) {
    const int32_t zero = 0;
    auto result = make_devmat(X.rows, X.rows, &zero, 1, X.precision, s);
    fill_identity(*result, s);
    auto term = make_devmat(X.rows, X.rows, &zero, 1, X.precision, s);
    fill_identity(*term, s);
    int i;
    for (i = 1; i <= max_iterations; ++i) {
        term = product(*term, X, static_cast<double>(i), s, true);
        // once the terms reach every diagonal the sum holds, the union is the sum's own layout: add in place
        if (same_layout(*result, *term))
            add_in_place(*result, *term, s);
        else
            result = union_add(*result, *term, s);
        if (frobenius(*term, s) < tol) break;
    }
    if (i == max_iterations) {  // the reference's check (one short of the loop bound), kept as is
        std::printf("Warning: %s reached max iterations (%d) without convergence.\n", who, max_iterations);
    }
    if (terms) *terms = std::min(i, max_iterations);
    return result;
}

// U^(2^sc) (src/spME_gpu.cu:282, :67 -> pow_cuda). An FP32 U -- Taylor terms in FP32 -- is squared in FP64: each
// squaring doubles the relative error it inherits, so FP32 rounding added at every one of sc squarings would grow as
// 2^sc. The result is rounded back to FP32 once, at the end.
static std::unique_ptr<dq_devmat> square_phase(std::unique_ptr<dq_devmat> U, int sc, hipStream_t s) {
    if (sc == 0) return U;  // pow_cuda(U, 1) is a copy of U
    if (U->precision == DQ_F64) return matrix_pow(*U, int64_t(1) << sc, s);
    auto U64 = convert_devmat(*U, DQ_F64, s);
    U.reset();
    auto P = matrix_pow(*U64, int64_t(1) << sc, s);
    U64.reset();
    return convert_devmat(*P, DQ_F32, s);
}

static int scaling_exponent(double scaled_norm, double cutoff) {
    return (scaled_norm > cutoff) ? static_cast<int>(std::ceil(std::log2(scaled_norm / cutoff))) : 0;
}

std::unique_ptr<dq_devmat> expm_negative(const dq_devmat &A, double t, int max_iterations, double cutoff,
                                         double tolerance, int *terms, int *s_out, hipStream_t s) {
    if (A.rows != A.cols) throw ArgError("expm_negative: the matrix must be square");
    if (A.n() == 1) {
        if (terms) *terms = 0;
        if (s_out) *s_out = 0;
        return single_diag_exp(A, t, true, s);
    }
    const double norm_est = estimate_1norm(A, s);
    const double scaled_norm = std::fabs(t) * norm_est + 1e-12;
    const int sc = scaling_exponent(scaled_norm, cutoff);
    if (sc > 10)
        std::printf("\033[93m[WARNING]\033[0m expm_negative_taylor_scaled_cuda: scaling exponent s = %d is large "
                    "(‖A‖₁ = %.3f)\n", sc, norm_est);
    const double scaled_t = std::ldexp(t, -sc);
    const double tol_scaled = std::max(1e-15, tolerance * std::ldexp(1.0, -sc));
    auto X = make_devmat(A.rows, A.cols, A.offs.data(), A.n(), A.precision, s);
    if (-1.0 * scaled_t != 0.0) times_i_scale_copy(A, *X, -1.0 * scaled_t, s);
    auto U = taylor(*X, max_iterations, tol_scaled, "expm_negative_taylor_cuda", terms, s);
    if (s_out) *s_out = sc;
    return square_phase(std::move(U), sc, s);
}

std::unique_ptr<dq_devmat> expm(const dq_devmat &A, int max_iterations, double cutoff, double tolerance,
                                hipStream_t s) {
    if (A.rows != A.cols) throw ArgError("expm: the matrix must be square");
    if (A.n() == 1) return single_diag_exp(A, 1.0, false, s);
    const double norm_est = estimate_1norm(A, s);
    const int sc = scaling_exponent(norm_est + 1e-12, cutoff);
    if (sc > 10)
        std::printf("\033[93m[WARNING]\033[0m expm_taylor_scaled_cuda: scaling exponent s = %d is large "
                    "(‖A‖₁ = %.3f)\n", sc, norm_est);
    const double inv_scale = std::ldexp(1.0, -sc);
    auto As = copy_devmat(A, s);
    scale_real(*As, inv_scale, s);
    const double tol_scaled = std::max(1e-15, tolerance * inv_scale);
    auto U = taylor(*As, max_iterations, tol_scaled, "expm_taylor_cuda", nullptr, s);
    return square_phase(std::move(U), sc, s);
}

}  // namespace dq

using namespace dq;

extern "C" int dq_expm_taylor_scaled(const dq_devmat *A, int max_iterations, double scale_cutoff, double tolerance,
                                     dq_devmat **out) {
    return guard([&] {
        if (!A || !out) throw ArgError("NULL argument");
        *out = expm(*A, max_iterations, scale_cutoff, tolerance, 0).release();
        DQ_HIP(hipStreamSynchronize(0));
    });
}

extern "C" int dq_expm_negative_taylor_scaled(const dq_devmat *A, double t, int max_iterations, double scale_cutoff,
                                              double tolerance, dq_devmat **out, int *terms_out, int *s_out) {
    return guard([&] {
        if (!A || !out) throw ArgError("NULL argument");
        *out = expm_negative(*A, t, max_iterations, scale_cutoff, tolerance, terms_out, s_out, 0).release();
        DQ_HIP(hipStreamSynchronize(0));
    });
}
