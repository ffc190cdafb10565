// Complex FP64 dense product on the matrix cores of gfx950: C = (A B) / divisor for row-major complex matrices stored
// as interleaved (re, im) doubles. It serves the near-dense products inside pow and the matrix exponential (the
// squarings of a propagator that fills every diagonal), after the DIA operands are expanded to dense.
//
// Tiling. A workgroup of 512 lanes (8 waves, 4 x 2) owns a 128 x 128 tile of C and walks K in steps of 8. Each step
// stages A's 128 x 8 and B's 8 x 128 complex panels in LDS, split into real and imaginary planes, k-major, so that
// the MFMA operand reads (16 consecutive rows or columns per k) are conflict-free; the next step's panels are
// loaded into registers while the current step computes (two LDS buffers, 66 KB). A wave owns 32 x 64 of C as
// 2 x 4 v_mfma_f64_16x16x4 tiles and keeps their real and imaginary accumulators in registers (64 doubles per lane):
// Re C += Re A Re B - Im A Im B and Im C += Re A Im B + Im A Re B, four MFMAs per tile and k-group of 4. The kernel
// needs 192 VGPRs and no AGPRs, so two waves share each SIMD: one wave alone cannot keep the FP64 matrix pipe busy
// (a 64 x 64 wave tile at one wave per SIMD ran at 46 TFLOP/s against 67 for this form at 4096^3).
//
// Operand layout of v_mfma_f64_16x16x4 (measured on gfx950): lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16];
// it holds D[l / 16 + 4 v][l % 16] in register v of its 4 accumulators.
//
// Order. Workgroups are numbered so that each XCD (8 of them, workgroup ids dealt round-robin) takes a contiguous
// range of tile rows: the A panels of those rows stay in that XCD's L2 while it sweeps the columns.
#include "dq_common.hpp"

namespace dq {

namespace {
typedef double d4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));

constexpr int GB_M = 128, GB_N = 128, GB_K = 8;
constexpr int G_THREADS = 512;
constexpr int G_PAD = 4;  // LDS row padding (doubles)
constexpr int G_LDA = GB_M + G_PAD, G_LDB = GB_N + G_PAD;

struct GemmSmem {
    double ar[2][GB_K][G_LDA], ai[2][GB_K][G_LDA];
    double br[2][GB_K][G_LDB], bi[2][GB_K][G_LDB];
};
}  // namespace

// One K step's panels, global -> registers: lane t loads A row t / 4, k columns 2 (t % 4) .. +2, and B row k = t / 64,
// columns 2 (t % 64) .. +2; out-of-range elements read as zero.
__device__ inline void gemm_load(const d2 *__restrict__ A, const d2 *__restrict__ B, int64_t M, int64_t N, int64_t K,
                                 int64_t m0, int64_t n0, int64_t k0, d2 (&ra)[2], d2 (&rb)[2]) {
    const int t = threadIdx.x;
    {
        const int64_t r = m0 + t / 4, k = k0 + 2 * (t % 4);
        const bool rok = r < M;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            ra[e] = d2{0.0, 0.0};
            if (rok && k + e < K) ra[e] = A[r * K + k + e];
        }
    }
    {
        const int64_t k = k0 + t / 64, c = n0 + 2 * (t % 64);
        const bool kok = k < K;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            rb[e] = d2{0.0, 0.0};
            if (kok && c + e < N) rb[e] = B[k * N + c + e];
        }
    }
}

__device__ inline void gemm_store_lds(GemmSmem &sm, int buf, const d2 (&ra)[2], const d2 (&rb)[2]) {
    const int t = threadIdx.x;
    {
        const int r = t / 4, k = 2 * (t % 4);
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            sm.ar[buf][k + e][r] = ra[e].x;
            sm.ai[buf][k + e][r] = ra[e].y;
        }
    }
    {
        const int k = t / 64, c = 2 * (t % 64);
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            sm.br[buf][k][c + e] = rb[e].x;
            sm.bi[buf][k][c + e] = rb[e].y;
        }
    }
}

__global__ __launch_bounds__(G_THREADS) __attribute__((amdgpu_waves_per_eu(2, 2))) void zgemm_mfma_kernel(
    int64_t M, int64_t N, int64_t K, const d2 *__restrict__ A, const d2 *__restrict__ B, d2 *__restrict__ C,
    double divisor, int tiles_m, int tiles_n) {
    __shared__ GemmSmem sm;
    // workgroup id -> tile: XCD x = id % 8 takes the x-th contiguous eighth of the row-major tile order
    const int ntiles = tiles_m * tiles_n;
    const int per_xcd = (ntiles + 7) / 8;
    const int id = blockIdx.x;
    const int tile = (id % 8) * per_xcd + id / 8;
    if (tile >= ntiles) return;  // uniform: the last XCD's share may be short
    const int64_t m0 = static_cast<int64_t>(tile / tiles_n) * GB_M, n0 = static_cast<int64_t>(tile % tiles_n) * GB_N;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 64;
    const int li = lane & 15, lk = lane >> 4;

    d4 cr[2][4], ci[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            cr[i][j] = d4{0.0, 0.0, 0.0, 0.0};
            ci[i][j] = d4{0.0, 0.0, 0.0, 0.0};
        }

    d2 ra[2], rb[2];
    gemm_load(A, B, M, N, K, m0, n0, 0, ra, rb);
    gemm_store_lds(sm, 0, ra, rb);
    __syncthreads();
    const int64_t nk = (K + GB_K - 1) / GB_K;
    for (int64_t ks = 0; ks < nk; ++ks) {
        const int buf = static_cast<int>(ks & 1);
        if (ks + 1 < nk) gemm_load(A, B, M, N, K, m0, n0, (ks + 1) * GB_K, ra, rb);
#pragma unroll
        for (int kk = 0; kk < GB_K; kk += 4) {
            double a_r[2], a_i[2], a_n[2], b_r[4], b_i[4];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                a_r[i] = sm.ar[buf][kk + lk][wm + 16 * i + li];
                a_i[i] = sm.ai[buf][kk + lk][wm + 16 * i + li];
                a_n[i] = -a_i[i];
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                b_r[j] = sm.br[buf][kk + lk][wn + 16 * j + li];
                b_i[j] = sm.bi[buf][kk + lk][wn + 16 * j + li];
            }
            // 16 independent MFMAs separate the two that update one accumulator
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_r[i], b_r[j], cr[i][j], 0, 0, 0);
                    ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_r[i], b_i[j], ci[i][j], 0, 0, 0);
                }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_n[i], b_i[j], cr[i][j], 0, 0, 0);
                    ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_i[i], b_r[j], ci[i][j], 0, 0, 0);
                }
        }
        if (ks + 1 < nk) {
            gemm_store_lds(sm, buf ^ 1, ra, rb);
            __syncthreads();
        }
    }
    // epilogue: lane (li, lk) of MFMA tile (i, j) holds rows lk + 4 v, column li
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t c = n0 + wn + 16 * j + li;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t r = m0 + wm + 16 * i + lk + 4 * v;
                if (r < M && c < N) {
                    double vr = cr[i][j][v], vi = ci[i][j][v];
                    if (divisor != 1.0) {
                        vr = vr / divisor;
                        vi = vi / divisor;
                    }
                    C[r * N + c] = d2{vr, vi};
                }
            }
        }
}

// C (M x N) = (A (M x K) B (K x N)) / divisor, complex interleaved row-major, on stream s.
void zgemm_dense(int64_t M, int64_t N, int64_t K, const double *A, const double *B, double *C, double divisor,
                 hipStream_t s) {
    if (M <= 0 || N <= 0) return;
    const int64_t tm = (M + GB_M - 1) / GB_M, tn = (N + GB_N - 1) / GB_N;
    if (tm * tn > (int64_t(1) << 30)) throw ArgError("dense product: too many tiles");
    const int tiles_m = static_cast<int>(tm), tiles_n = static_cast<int>(tn);
    const int per_xcd = (tiles_m * tiles_n + 7) / 8;
    hipLaunchKernelGGL(zgemm_mfma_kernel, dim3(static_cast<unsigned>(8 * per_xcd)), dim3(G_THREADS), 0, s, M, N, K,
                       reinterpret_cast<const d2 *>(A), reinterpret_cast<const d2 *>(B), reinterpret_cast<d2 *>(C),
                       divisor, tiles_m, tiles_n);
    DQ_LAUNCH_CHECK();
}

}  // namespace dq
