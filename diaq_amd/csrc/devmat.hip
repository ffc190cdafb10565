// Device-resident DIA matrices: layout, upload / download / copy, elementwise and segment kernels, reductions
// (Frobenius norm, 1-norm estimate, trace), union-add and prune.
//
// Reference counterparts: deviceDiaQMatrix (src/deviceMatrix_gpu.hpp:26-79, src/deviceMatrix_gpu.cu:10-113) and the
// host helpers the GPU SpME interleaves with device products (src/diaQMatrix.hpp:867-916, 971-988, 1181-1211,
// 1255-1358, 1543-1604), which run here on the GPU.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "dq_common.hpp"

namespace dq {

static thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }

// ------------------------------------------------------------------------------------------------------------------
// Device memory with a cache of freed blocks. A product allocates its result (gigabytes) and its work lists on every
// call; hipMalloc / hipFree of such blocks cost about a millisecond per product. dev_free waits for the device to be
// idle -- as hipFree does -- and keeps the block; dev_alloc hands out the smallest cached block of the same device
// that is large enough and at most a quarter larger (rounded sizes: 2 MiB granularity above 1 MiB), else allocates.
// The cache holds at most a budget (default an eighth of the device memory, dq_set_tuning "alloc_cache_mb"); beyond
// it, and when an allocation fails, blocks go back to the runtime. Cached blocks of a process live until it ends
// (never freed from a static destructor, when the runtime may already be gone), or until dq_release_cached_memory.
// ------------------------------------------------------------------------------------------------------------------
namespace {
struct AllocCache {
    std::mutex mu;
    std::unordered_map<void *, std::pair<int, size_t>> live;      // pointer -> (device, capacity)
    std::map<std::pair<int, size_t>, std::vector<void *>> free_;  // (device, capacity) -> cached blocks
    size_t cached = 0;
    int64_t budget = -1;  // bytes; -1: not yet set (an eighth of the device memory)
};
AllocCache &cache() {
    static AllocCache *c = new AllocCache;  // never destroyed: see above
    return *c;
}
size_t round_size(size_t n) {
    const size_t mib = size_t(1) << 20;
    if (n >= mib) return (n + 2 * mib - 1) / (2 * mib) * (2 * mib);
    return (n + 511) / 512 * 512;
}
void release_all_locked(AllocCache &c) {
    (void)hipDeviceSynchronize();
    for (auto &kv : c.free_)
        for (void *p : kv.second) (void)hipFree(p);
    c.free_.clear();
    c.cached = 0;
}
}  // namespace

void *dev_alloc(size_t bytes) {
    if (bytes == 0) return nullptr;
    AllocCache &c = cache();
    int dev = 0;
    (void)hipGetDevice(&dev);
    const size_t want = round_size(bytes);
    {
        std::lock_guard<std::mutex> lk(c.mu);
        auto it = c.free_.lower_bound({dev, want});
        if (it != c.free_.end() && it->first.first == dev && it->first.second <= want + want / 4) {
            void *p = it->second.back();
            it->second.pop_back();
            const size_t cap = it->first.second;
            if (it->second.empty()) c.free_.erase(it);
            c.cached -= cap;
            c.live[p] = {dev, cap};
            return p;
        }
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {  // give the cached blocks back and try once more
        (void)hipGetLastError();
        {
            std::lock_guard<std::mutex> lk(c.mu);
            release_all_locked(c);
        }
        e = hipMalloc(&p, want);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw AllocError("HIP error: " + std::string(hipGetErrorString(e)) + " allocating " +
                         std::to_string(bytes) + " bytes of device memory");
    }
    std::lock_guard<std::mutex> lk(c.mu);
    c.live[p] = {dev, want};
    return p;
}

void dev_free(void *p) {
    if (!p) return;
    AllocCache &c = cache();
    std::lock_guard<std::mutex> lk(c.mu);
    auto it = c.live.find(p);
    if (it == c.live.end()) {  // not ours (cannot happen through DevBuf): plain free
        (void)hipFree(p);
        return;
    }
    const auto [dev, cap] = it->second;
    c.live.erase(it);
    if (c.budget < 0) {
        size_t fr = 0, tot = 0;
        c.budget = hipMemGetInfo(&fr, &tot) == hipSuccess ? static_cast<int64_t>(tot / 8) : 0;
        (void)hipGetLastError();
    }
    (void)hipDeviceSynchronize();  // as hipFree: no kernel may still use the block when it is handed out again
    if (c.cached + cap > static_cast<size_t>(c.budget)) {
        (void)hipFree(p);
        return;
    }
    c.free_[{dev, cap}].push_back(p);
    c.cached += cap;
}

void set_alloc_cache_budget(int64_t mib) {
    AllocCache &c = cache();
    std::lock_guard<std::mutex> lk(c.mu);
    c.budget = mib < 0 ? -1 : mib * (int64_t(1) << 20);
    if (c.budget >= 0 && c.cached > static_cast<size_t>(c.budget)) release_all_locked(c);
}

void release_cached_memory() {
    AllocCache &c = cache();
    std::lock_guard<std::mutex> lk(c.mu);
    release_all_locked(c);
}

__global__ void zero_bytes_kernel(uint4 *__restrict__ p16, int64_t n16, uint8_t *__restrict__ tail, int ntail) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
        p16[i] = make_uint4(0u, 0u, 0u, 0u);
    if (blockIdx.x == 0 && static_cast<int>(threadIdx.x) < ntail) tail[threadIdx.x] = 0;
}

void zero_bytes(void *p, size_t bytes, hipStream_t s) {
    if (!bytes) return;
    if (reinterpret_cast<uintptr_t>(p) % 16 != 0) throw ArgError("zero_bytes: pointer not 16-byte aligned");
    const int64_t n16 = static_cast<int64_t>(bytes / 16);
    const int ntail = static_cast<int>(bytes % 16);
    hipLaunchKernelGGL(zero_bytes_kernel, dim3(std::max<unsigned>(1, blocks_for(n16, 256, 8192))), dim3(256), 0, s,
                       static_cast<uint4 *>(p), n16, static_cast<uint8_t *>(p) + n16 * 16, ntail);
    DQ_LAUNCH_CHECK();
}

// ------------------------------------------------------------------------------------------------------------------
// Layout
// ------------------------------------------------------------------------------------------------------------------
std::unique_ptr<dq_devmat> make_devmat(uint32_t rows, uint32_t cols, const int32_t *offsets, int n, int precision,
                                       hipStream_t stream, bool zero_fill) {
    if (n < 0) throw ArgError("negative number of diagonals");
    if (precision != DQ_F64 && precision != DQ_F32) throw ArgError("precision must be DQ_F64 or DQ_F32");
    if (n > 0 && !offsets) throw ArgError("offsets is NULL");
    auto M = std::make_unique<dq_devmat>();
    M->rows = rows;
    M->cols = cols;
    M->precision = precision;
    M->offs.assign(offsets, offsets + n);
    M->start.resize(n);
    M->len.resize(n);
    M->meta.resize(n);
    const int64_t align = DQ_ALIGN_BYTES / elem_bytes(precision);
    int64_t cursor = 0;
    for (int k = 0; k < n; ++k) {
        const int32_t d = offsets[k];
        if (d <= -(int64_t)rows || d >= (int64_t)cols) {
            if (!(rows == 0 || cols == 0)) throw ArgError("diagonal offset " + std::to_string(d) + " outside the matrix");
        }
        const int64_t len = diagonal_length(d, rows, cols);
        const int64_t rlo = d < 0 ? -(int64_t)d : 0;
        int64_t start = cursor + (((rlo - cursor) % align) + align) % align;  // rowbase = start - rlo aligned
        M->start[k] = start;
        M->len[k] = len;
        M->meta[k] = DiagMeta{start - rlo, rlo, rlo + len, d, 0};
        cursor = start + len;
    }
    M->total = ((cursor + align - 1) / align) * align;
    const size_t vb = static_cast<size_t>(M->total) * elem_bytes(precision);
    M->re = DevBuf(vb);
    M->im = DevBuf(vb);
    if (vb && zero_fill) {
        zero_bytes(M->re.p, vb, stream);
        zero_bytes(M->im.p, vb, stream);
    }
    if (n) {
        M->d_meta = DevBuf(sizeof(DiagMeta) * n);
        DQ_HIP(hipMemcpyAsync(M->d_meta.p, M->meta.data(), sizeof(DiagMeta) * n, hipMemcpyHostToDevice, stream));
    }
    return M;
}

// ------------------------------------------------------------------------------------------------------------------
// Segment kernels: dst[i] = A[a + i] + sign_b * B[b + i] over pieces of at most PIECE elements (an absent source is
// zero). One workgroup per piece; pieces are cut from every segment so long diagonals spread over many workgroups.
// ------------------------------------------------------------------------------------------------------------------
struct Piece {
    int64_t dst, a, b;  // element indices; a or b = -1: absent
    int32_t n;
    int32_t pad_;
};
constexpr int PIECE = 4096;

struct Seg {
    int64_t dst, a, b, len;
};

static std::vector<Piece> cut(const std::vector<Seg> &segs) {
    std::vector<Piece> out;
    for (const Seg &s : segs) {
        for (int64_t o = 0; o < s.len; o += PIECE) {
            const int32_t n = static_cast<int32_t>(std::min<int64_t>(PIECE, s.len - o));
            out.push_back(Piece{s.dst + o, s.a < 0 ? -1 : s.a + o, s.b < 0 ? -1 : s.b + o, n, 0});
        }
    }
    return out;
}

template <typename TD, typename TS>
__global__ __launch_bounds__(256) void piece_kernel(const Piece *__restrict__ pieces, TD *__restrict__ dr,
                                                    TD *__restrict__ di, const TS *__restrict__ ar,
                                                    const TS *__restrict__ ai, const TS *__restrict__ br,
                                                    const TS *__restrict__ bi, double sign_b) {
    const Piece p = pieces[blockIdx.x];
    for (int t = threadIdx.x; t < p.n; t += blockDim.x) {
        double vr = 0.0, vi = 0.0;
        if (p.a >= 0) { vr = static_cast<double>(ar[p.a + t]); vi = static_cast<double>(ai[p.a + t]); }
        if (p.b >= 0) {
            const double wr = static_cast<double>(br[p.b + t]), wi = static_cast<double>(bi[p.b + t]);
            if (p.a >= 0) { vr = vr + sign_b * wr; vi = vi + sign_b * wi; }
            else { vr = sign_b * wr; vi = sign_b * wi; }
        }
        dr[p.dst + t] = static_cast<TD>(vr);
        di[p.dst + t] = static_cast<TD>(vi);
    }
}

// Run segments from (A[, B]) into D; D, A, B may have different precisions only for the single-source copy.
static void run_segments(const std::vector<Seg> &segs, dq_devmat &D, const dq_devmat *A, const dq_devmat *B,
                         double sign_b, hipStream_t s) {
    const std::vector<Piece> pieces = cut(segs);
    if (pieces.empty()) return;
    DevBuf dp(sizeof(Piece) * pieces.size());
    DQ_HIP(hipMemcpyAsync(dp.p, pieces.data(), sizeof(Piece) * pieces.size(), hipMemcpyHostToDevice, s));
    const unsigned grid = static_cast<unsigned>(pieces.size());
    const dq_devmat *src = A ? A : B;
    const int sp = src->precision;
    if (B && A && A->precision != B->precision) throw ArgError("mixed precisions in one elementwise operation");
    auto srcp = [](const dq_devmat *M, bool imag) -> const void * {
        return M ? (imag ? M->im.p : M->re.p) : nullptr;
    };
    if (D.precision == DQ_F64 && sp == DQ_F64) {
        hipLaunchKernelGGL((piece_kernel<double, double>), dim3(grid), dim3(256), 0, s, dp.as<Piece>(),
                           D.re.as<double>(), D.im.as<double>(), (const double *)srcp(A, false),
                           (const double *)srcp(A, true), (const double *)srcp(B, false),
                           (const double *)srcp(B, true), sign_b);
    } else if (D.precision == DQ_F32 && sp == DQ_F32) {
        hipLaunchKernelGGL((piece_kernel<float, float>), dim3(grid), dim3(256), 0, s, dp.as<Piece>(),
                           D.re.as<float>(), D.im.as<float>(), (const float *)srcp(A, false),
                           (const float *)srcp(A, true), (const float *)srcp(B, false), (const float *)srcp(B, true),
                           sign_b);
    } else if (D.precision == DQ_F32 && sp == DQ_F64) {
        hipLaunchKernelGGL((piece_kernel<float, double>), dim3(grid), dim3(256), 0, s, dp.as<Piece>(),
                           D.re.as<float>(), D.im.as<float>(), (const double *)srcp(A, false),
                           (const double *)srcp(A, true), (const double *)srcp(B, false),
                           (const double *)srcp(B, true), sign_b);
    } else {
        hipLaunchKernelGGL((piece_kernel<double, float>), dim3(grid), dim3(256), 0, s, dp.as<Piece>(),
                           D.re.as<double>(), D.im.as<double>(), (const float *)srcp(A, false),
                           (const float *)srcp(A, true), (const float *)srcp(B, false), (const float *)srcp(B, true),
                           sign_b);
    }
    DQ_LAUNCH_CHECK();
    DQ_HIP(hipStreamSynchronize(s));  // the piece list is freed on return
}

std::unique_ptr<dq_devmat> copy_devmat(const dq_devmat &A, hipStream_t s) {
    auto C = make_devmat(A.rows, A.cols, A.offs.data(), A.n(), A.precision, s);
    const size_t vb = static_cast<size_t>(A.total) * elem_bytes(A.precision);
    if (vb) {
        DQ_HIP(hipMemcpyAsync(C->re.p, A.re.p, vb, hipMemcpyDeviceToDevice, s));
        DQ_HIP(hipMemcpyAsync(C->im.p, A.im.p, vb, hipMemcpyDeviceToDevice, s));
    }
    return C;
}

std::unique_ptr<dq_devmat> convert_devmat(const dq_devmat &A, int precision, hipStream_t s) {
    auto C = make_devmat(A.rows, A.cols, A.offs.data(), A.n(), precision, s);
    std::vector<Seg> segs;
    for (int k = 0; k < A.n(); ++k) segs.push_back(Seg{C->start[k], A.start[k], -1, A.len[k]});
    run_segments(segs, *C, &A, nullptr, 1.0, s);
    return C;
}

std::unique_ptr<dq_devmat> union_add(const dq_devmat &A, const dq_devmat &B, hipStream_t s) {
    if (A.rows != B.rows || A.cols != B.cols) throw ArgError("diaQMatrix::add: shape mismatch");
    std::vector<int32_t> u(A.offs);
    u.insert(u.end(), B.offs.begin(), B.offs.end());
    std::sort(u.begin(), u.end());
    u.erase(std::unique(u.begin(), u.end()), u.end());
    std::unordered_map<int32_t, int> sa, sb;
    for (int k = 0; k < A.n(); ++k) sa[A.offs[k]] = k;
    for (int k = 0; k < B.n(); ++k) sb[B.offs[k]] = k;
    auto C = make_devmat(A.rows, A.cols, u.data(), static_cast<int>(u.size()), A.precision, s);
    std::vector<Seg> segs;
    for (int k = 0; k < C->n(); ++k) {
        const int32_t d = u[k];
        auto ia = sa.find(d);
        auto ib = sb.find(d);
        segs.push_back(Seg{C->start[k], ia == sa.end() ? -1 : A.start[ia->second],
                           ib == sb.end() ? -1 : B.start[ib->second], C->len[k]});
    }
    run_segments(segs, *C, &A, &B, 1.0, s);
    return C;
}

// ------------------------------------------------------------------------------------------------------------------
// Flat elementwise kernels (padding is zero and stays zero)
// ------------------------------------------------------------------------------------------------------------------
template <typename T>
__global__ void scale_real_kernel(int64_t n, T *__restrict__ re, T *__restrict__ im, double f) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        re[i] = static_cast<T>(static_cast<double>(re[i]) * f);
        im[i] = static_cast<T>(static_cast<double>(im[i]) * f);
    }
}

void scale_real(dq_devmat &M, double f, hipStream_t s) {
    if (!M.total) return;
    const unsigned g = blocks_for(M.total, 256, 8192);
    if (M.precision == DQ_F64)
        hipLaunchKernelGGL(scale_real_kernel<double>, dim3(g), dim3(256), 0, s, M.total, M.re.as<double>(),
                           M.im.as<double>(), f);
    else
        hipLaunchKernelGGL(scale_real_kernel<float>, dim3(g), dim3(256), 0, s, M.total, M.re.as<float>(),
                           M.im.as<float>(), f);
    DQ_LAUNCH_CHECK();
}

// A += B for two matrices with one layout (same shape, precision and offsets in the same order): one flat pass over
// the value arrays, padding included (zero plus zero). The same sums as union_add's A + 1.0 * B.
template <typename T>
__global__ void add_flat_kernel(int64_t n, T *__restrict__ ar, T *__restrict__ ai, const T *__restrict__ br,
                                const T *__restrict__ bi) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        ar[i] = ar[i] + br[i];
        ai[i] = ai[i] + bi[i];
    }
}

bool same_layout(const dq_devmat &A, const dq_devmat &B) {
    return A.rows == B.rows && A.cols == B.cols && A.precision == B.precision && A.total == B.total &&
           A.offs == B.offs;
}

void add_in_place(dq_devmat &A, const dq_devmat &B, hipStream_t s) {
    if (!same_layout(A, B)) throw ArgError("add_in_place: layouts differ");
    if (!A.total) return;
    const unsigned g = blocks_for(A.total, 256, 8192);
    if (A.precision == DQ_F64)
        hipLaunchKernelGGL(add_flat_kernel<double>, dim3(g), dim3(256), 0, s, A.total, A.re.as<double>(),
                           A.im.as<double>(), B.re.as<double>(), B.im.as<double>());
    else
        hipLaunchKernelGGL(add_flat_kernel<float>, dim3(g), dim3(256), 0, s, A.total, A.re.as<float>(),
                           A.im.as<float>(), B.re.as<float>(), B.im.as<float>());
    DQ_LAUNCH_CHECK();
}

// C = (i * imag) A: (re, im) -> (-1.0 * im * imag, re * imag) (src/diaQMatrix.hpp:1181-1211)
template <typename T>
__global__ void times_i_kernel(int64_t n, const T *__restrict__ ar, const T *__restrict__ ai, T *__restrict__ cr,
                               T *__restrict__ ci, double imag) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double r = static_cast<double>(ar[i]), m = static_cast<double>(ai[i]);
        cr[i] = static_cast<T>(-1.0 * m * imag);
        ci[i] = static_cast<T>(r * imag);
    }
}

void times_i_scale_copy(const dq_devmat &A, dq_devmat &C, double imag, hipStream_t s) {
    if (C.total != A.total || C.precision != A.precision) throw ArgError("times_i_scale_copy: layouts differ");
    if (!A.total) return;
    const unsigned g = blocks_for(A.total, 256, 8192);
    if (A.precision == DQ_F64)
        hipLaunchKernelGGL(times_i_kernel<double>, dim3(g), dim3(256), 0, s, A.total, A.re.as<double>(),
                           A.im.as<double>(), C.re.as<double>(), C.im.as<double>(), imag);
    else
        hipLaunchKernelGGL(times_i_kernel<float>, dim3(g), dim3(256), 0, s, A.total, A.re.as<float>(),
                           A.im.as<float>(), C.re.as<float>(), C.im.as<float>(), imag);
    DQ_LAUNCH_CHECK();
}

template <typename T>
__global__ void fill_kernel(int64_t n, T *__restrict__ re, T *__restrict__ im, T vr, T vi) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        re[i] = vr;
        im[i] = vi;
    }
}

void fill_identity(dq_devmat &M, hipStream_t s) {
    if (M.n() != 1 || M.offs[0] != 0) throw ArgError("fill_identity: needs the single main diagonal");
    const int64_t n = M.len[0];
    if (!n) return;
    const unsigned g = blocks_for(n, 256, 8192);
    if (M.precision == DQ_F64)
        hipLaunchKernelGGL(fill_kernel<double>, dim3(g), dim3(256), 0, s, n, M.re.as<double>() + M.start[0],
                           M.im.as<double>() + M.start[0], 1.0, 0.0);
    else
        hipLaunchKernelGGL(fill_kernel<float>, dim3(g), dim3(256), 0, s, n, M.re.as<float>() + M.start[0],
                           M.im.as<float>() + M.start[0], 1.0f, 0.0f);
    DQ_LAUNCH_CHECK();
}

// ------------------------------------------------------------------------------------------------------------------
// Reductions: per-workgroup partials in FP64, then one workgroup folds them; wave64 shuffles inside a workgroup.
// ------------------------------------------------------------------------------------------------------------------
template <bool MAX>
__device__ inline double wave_reduce(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_xor(v, o, 64);
        v = MAX ? fmax(v, w) : v + w;
    }
    return v;
}

template <bool MAX>
__device__ inline double block_reduce(double v) {
    __shared__ double part[16];
    v = wave_reduce<MAX>(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) part[wid] = v;
    __syncthreads();
    const int nw = (blockDim.x + 63) >> 6;
    v = (threadIdx.x < nw) ? part[threadIdx.x] : (MAX ? 0.0 : 0.0);
    if (wid == 0) v = wave_reduce<MAX>(v);
    return v;  // valid in thread 0
}

template <bool MAX>
__global__ __launch_bounds__(1024) void fold_kernel(int n, const double *__restrict__ partial, double *out) {
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v = MAX ? fmax(v, partial[i]) : v + partial[i];
    v = block_reduce<MAX>(v);
    if (threadIdx.x == 0) *out = v;
}

template <bool MAX>
static double fold(DevBuf &partials, int n, hipStream_t s) {
    DevBuf out(sizeof(double));
    hipLaunchKernelGGL(fold_kernel<MAX>, dim3(1), dim3(1024), 0, s, n, partials.as<double>(), out.as<double>());
    DQ_LAUNCH_CHECK();
    double h = 0.0;
    DQ_HIP(hipMemcpyAsync(&h, out.p, sizeof(double), hipMemcpyDeviceToHost, s));
    DQ_HIP(hipStreamSynchronize(s));
    return h;
}

template <typename T>
__global__ __launch_bounds__(256) void sumsq_kernel(int64_t n, const T *__restrict__ re, const T *__restrict__ im,
                                                    double *partial) {
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double r = static_cast<double>(re[i]), m = static_cast<double>(im[i]);
        acc += r * r + m * m;
    }
    acc = block_reduce<false>(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

double frobenius(const dq_devmat &A, hipStream_t s) {
    if (!A.total) return 0.0;
    const unsigned g = blocks_for(A.total, 256 * 8, 1024);
    DevBuf partial(sizeof(double) * g);
    if (A.precision == DQ_F64)
        hipLaunchKernelGGL(sumsq_kernel<double>, dim3(g), dim3(256), 0, s, A.total, A.re.as<double>(),
                           A.im.as<double>(), partial.as<double>());
    else
        hipLaunchKernelGGL(sumsq_kernel<float>, dim3(g), dim3(256), 0, s, A.total, A.re.as<float>(),
                           A.im.as<float>(), partial.as<double>());
    DQ_LAUNCH_CHECK();
    return std::sqrt(fold<false>(partial, static_cast<int>(g), s));
}

// Column sums of |re| + |im| (one thread per column, every diagonal visited in slot order), then the maximum.
template <typename T>
__global__ __launch_bounds__(256) void colsum_kernel(int64_t cols, int n, const DiagMeta *__restrict__ meta,
                                                     const T *__restrict__ re, const T *__restrict__ im,
                                                     double *partial) {
    double best = 0.0;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < cols; c += (int64_t)gridDim.x * blockDim.x) {
        double sum = 0.0;
        for (int k = 0; k < n; ++k) {
            const DiagMeta m = meta[k];
            const int64_t r = c - m.off;
            if (r >= m.rlo && r < m.rhi)
                sum += fabs(static_cast<double>(re[m.rowbase + r])) + fabs(static_cast<double>(im[m.rowbase + r]));
        }
        best = fmax(best, sum);
    }
    best = block_reduce<true>(best);
    if (threadIdx.x == 0) partial[blockIdx.x] = best;
}

double estimate_1norm(const dq_devmat &A, hipStream_t s) {
    if (!A.cols || !A.n()) return 0.0;
    const unsigned g = blocks_for(A.cols, 256, 1024);
    DevBuf partial(sizeof(double) * g);
    if (A.precision == DQ_F64)
        hipLaunchKernelGGL(colsum_kernel<double>, dim3(g), dim3(256), 0, s, (int64_t)A.cols, A.n(),
                           A.d_meta.as<DiagMeta>(), A.re.as<double>(), A.im.as<double>(), partial.as<double>());
    else
        hipLaunchKernelGGL(colsum_kernel<float>, dim3(g), dim3(256), 0, s, (int64_t)A.cols, A.n(),
                           A.d_meta.as<DiagMeta>(), A.re.as<float>(), A.im.as<float>(), partial.as<double>());
    DQ_LAUNCH_CHECK();
    return fold<true>(partial, static_cast<int>(g), s);
}

template <typename T>
__global__ __launch_bounds__(256) void sum_kernel(int64_t n, const T *__restrict__ v, double *partial) {
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        acc += static_cast<double>(v[i]);
    acc = block_reduce<false>(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

static double sum_range(const void *p, int precision, int64_t n, hipStream_t s) {
    if (!n) return 0.0;
    const unsigned g = blocks_for(n, 256 * 8, 1024);
    DevBuf partial(sizeof(double) * g);
    if (precision == DQ_F64)
        hipLaunchKernelGGL(sum_kernel<double>, dim3(g), dim3(256), 0, s, n, (const double *)p, partial.as<double>());
    else
        hipLaunchKernelGGL(sum_kernel<float>, dim3(g), dim3(256), 0, s, n, (const float *)p, partial.as<double>());
    DQ_LAUNCH_CHECK();
    return fold<false>(partial, static_cast<int>(g), s);
}

// Sum of |A - B|^2 over pieces that pair the same diagonal of two matrices (precisions may differ).
template <typename TA, typename TB>
__global__ __launch_bounds__(256) void diffsq_kernel(const Piece *__restrict__ pieces, const TA *__restrict__ ar,
                                                     const TA *__restrict__ ai, const TB *__restrict__ br,
                                                     const TB *__restrict__ bi, double *partial) {
    const Piece p = pieces[blockIdx.x];
    double acc = 0.0;
    for (int t = threadIdx.x; t < p.n; t += blockDim.x) {
        const double dr = static_cast<double>(ar[p.a + t]) - static_cast<double>(br[p.b + t]);
        const double di = static_cast<double>(ai[p.a + t]) - static_cast<double>(bi[p.b + t]);
        acc += dr * dr + di * di;
    }
    acc = block_reduce<false>(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

double frobenius_diff(const dq_devmat &A, const dq_devmat &B, hipStream_t s) {
    if (A.rows != B.rows || A.cols != B.cols || A.offs != B.offs)
        throw ArgError("norm_diff: the matrices must have the same shape and diagonals");
    std::vector<Seg> segs;
    for (int k = 0; k < A.n(); ++k) segs.push_back(Seg{0, A.start[k], B.start[k], A.len[k]});
    const std::vector<Piece> pieces = cut(segs);
    if (pieces.empty()) return 0.0;
    DevBuf dp(sizeof(Piece) * pieces.size()), partial(sizeof(double) * pieces.size());
    DQ_HIP(hipMemcpyAsync(dp.p, pieces.data(), dp.bytes, hipMemcpyHostToDevice, s));
    const unsigned g = static_cast<unsigned>(pieces.size());
#define DQ_DIFF(TA, TB)                                                                                        \
    hipLaunchKernelGGL((diffsq_kernel<TA, TB>), dim3(g), dim3(256), 0, s, dp.as<Piece>(), A.re.as<TA>(),    \
                       A.im.as<TA>(), B.re.as<TB>(), B.im.as<TB>(), partial.as<double>())
    if (A.precision == DQ_F64 && B.precision == DQ_F64) DQ_DIFF(double, double);
    else if (A.precision == DQ_F64) DQ_DIFF(double, float);
    else if (B.precision == DQ_F64) DQ_DIFF(float, double);
    else DQ_DIFF(float, float);
#undef DQ_DIFF
    DQ_LAUNCH_CHECK();
    return std::sqrt(fold<false>(partial, static_cast<int>(g), s));
}

// "Any value nonzero" per diagonal: one flag per diagonal, set by the workgroups of its pieces.
template <typename T>
__global__ __launch_bounds__(256) void nonzero_kernel(const Piece *__restrict__ pieces, const int32_t *__restrict__ seg_of,
                                                      const T *__restrict__ re, const T *__restrict__ im,
                                                      int32_t *flags) {
    const Piece p = pieces[blockIdx.x];
    int nz = 0;
    for (int t = threadIdx.x; t < p.n; t += blockDim.x) nz |= (re[p.a + t] != T(0)) | (im[p.a + t] != T(0));
    nz = __syncthreads_or(nz);
    if (threadIdx.x == 0 && nz) flags[seg_of[blockIdx.x]] = 1;
}

std::unique_ptr<dq_devmat> prune(const dq_devmat &A, hipStream_t s) {
    std::vector<Seg> segs;
    for (int k = 0; k < A.n(); ++k) segs.push_back(Seg{0, A.start[k], -1, A.len[k]});
    std::vector<Piece> pieces;
    std::vector<int32_t> seg_of;
    for (int k = 0; k < A.n(); ++k) {
        std::vector<Piece> pk = cut({segs[k]});
        for (auto &p : pk) { pieces.push_back(p); seg_of.push_back(k); }
    }
    std::vector<int32_t> flags(A.n(), 0);
    if (!pieces.empty()) {
        DevBuf dp(sizeof(Piece) * pieces.size()), ds(sizeof(int32_t) * seg_of.size()), df(sizeof(int32_t) * A.n());
        DQ_HIP(hipMemcpyAsync(dp.p, pieces.data(), dp.bytes, hipMemcpyHostToDevice, s));
        DQ_HIP(hipMemcpyAsync(ds.p, seg_of.data(), ds.bytes, hipMemcpyHostToDevice, s));
        zero_bytes(df.p, df.bytes, s);
        const unsigned g = static_cast<unsigned>(pieces.size());
        if (A.precision == DQ_F64)
            hipLaunchKernelGGL(nonzero_kernel<double>, dim3(g), dim3(256), 0, s, dp.as<Piece>(), ds.as<int32_t>(),
                               A.re.as<double>(), A.im.as<double>(), df.as<int32_t>());
        else
            hipLaunchKernelGGL(nonzero_kernel<float>, dim3(g), dim3(256), 0, s, dp.as<Piece>(), ds.as<int32_t>(),
                               A.re.as<float>(), A.im.as<float>(), df.as<int32_t>());
        DQ_LAUNCH_CHECK();
        DQ_HIP(hipMemcpyAsync(flags.data(), df.p, df.bytes, hipMemcpyDeviceToHost, s));
        DQ_HIP(hipStreamSynchronize(s));
    }
    std::vector<int32_t> keep_offs;
    std::vector<int> keep;
    for (int k = 0; k < A.n(); ++k)
        if (flags[k]) { keep.push_back(k); keep_offs.push_back(A.offs[k]); }
    auto C = make_devmat(A.rows, A.cols, keep_offs.data(), static_cast<int>(keep.size()), A.precision, s);
    std::vector<Seg> copy;
    for (size_t j = 0; j < keep.size(); ++j) copy.push_back(Seg{C->start[j], A.start[keep[j]], -1, A.len[keep[j]]});
    run_segments(copy, *C, &A, nullptr, 1.0, s);
    return C;
}

}  // namespace dq

// ======================================================================================================================
// C ABI
// ======================================================================================================================
using namespace dq;

extern "C" const char *dq_last_error(void) { return g_last_error.c_str(); }

extern "C" const char *dq_version(void) { return "diaq-mi355x 0.1.0 (gfx950)"; }

extern "C" int dq_devmat_zeros(uint32_t rows, uint32_t cols, int32_t n, const int32_t *offsets, int precision,
                               dq_devmat **out) {
    return guard([&] {
        if (!out) throw ArgError("out is NULL");
        auto M = make_devmat(rows, cols, offsets, n, precision, 0);
        DQ_HIP(hipStreamSynchronize(0));
        *out = M.release();
    });
}

extern "C" int dq_devmat_upload(uint32_t rows, uint32_t cols, int32_t n, const int32_t *offsets,
                                const int64_t *starts, const double *re, const double *im, int precision,
                                dq_devmat **out) {
    return guard([&] {
        if (!out) throw ArgError("out is NULL");
        if (n > 0 && (!starts || !re || !im)) throw ArgError("dq_devmat_upload: NULL host array");
        auto M = make_devmat(rows, cols, offsets, n, precision, 0);
        if (M->total) {
            // stage the diagonals in the device layout (padding zero), then one copy per value array
            const size_t eb = elem_bytes(precision);
            std::vector<char> sr(static_cast<size_t>(M->total) * eb, 0), si(static_cast<size_t>(M->total) * eb, 0);
            for (int k = 0; k < n; ++k) {
                const double *r = re + starts[k], *i = im + starts[k];
                if (precision == DQ_F64) {
                    std::memcpy(sr.data() + M->start[k] * eb, r, M->len[k] * eb);
                    std::memcpy(si.data() + M->start[k] * eb, i, M->len[k] * eb);
                } else {
                    float *fr = reinterpret_cast<float *>(sr.data()) + M->start[k];
                    float *fi = reinterpret_cast<float *>(si.data()) + M->start[k];
                    for (int64_t t = 0; t < M->len[k]; ++t) { fr[t] = (float)r[t]; fi[t] = (float)i[t]; }
                }
            }
            DQ_HIP(hipMemcpy(M->re.p, sr.data(), sr.size(), hipMemcpyHostToDevice));
            DQ_HIP(hipMemcpy(M->im.p, si.data(), si.size(), hipMemcpyHostToDevice));
        }
        DQ_HIP(hipStreamSynchronize(0));
        *out = M.release();
    });
}

extern "C" int dq_devmat_set_packed_device(dq_devmat *A, const void *re_dev, const void *im_dev, uintptr_t stream) {
    return guard([&] {
        if (!A) throw ArgError("matrix is NULL");
        if (A->n() && (!re_dev || !im_dev)) throw ArgError("dq_devmat_set_packed_device: NULL device array");
        hipStream_t s = as_stream(stream);
        const size_t eb = elem_bytes(A->precision);
        int64_t src = 0;
        for (int k = 0; k < A->n(); ++k) {
            if (A->len[k]) {
                DQ_HIP(hipMemcpyAsync(static_cast<char *>(A->re.p) + A->start[k] * eb,
                                      static_cast<const char *>(re_dev) + src * eb, A->len[k] * eb,
                                      hipMemcpyDeviceToDevice, s));
                DQ_HIP(hipMemcpyAsync(static_cast<char *>(A->im.p) + A->start[k] * eb,
                                      static_cast<const char *>(im_dev) + src * eb, A->len[k] * eb,
                                      hipMemcpyDeviceToDevice, s));
            }
            src += A->len[k];
        }
        // the values changed: the zero-block map a tuning kept is recomputed for them, in its own buffer
        refresh_zero_map(*A, s);
    });
}

extern "C" int dq_devmat_copy(const dq_devmat *A, dq_devmat **out) {
    return guard([&] {
        if (!A || !out) throw ArgError("NULL argument");
        auto C = copy_devmat(*A, 0);
        DQ_HIP(hipStreamSynchronize(0));
        *out = C.release();
    });
}

extern "C" int dq_devmat_convert(const dq_devmat *A, int precision, dq_devmat **out) {
    return guard([&] {
        if (!A || !out) throw ArgError("NULL argument");
        if (precision != DQ_F64 && precision != DQ_F32) throw ArgError("precision must be DQ_F64 or DQ_F32");
        *out = convert_devmat(*A, precision, 0).release();
    });
}

// Diagonal k of the device layout (start[k], len[k]) into a double buffer at the host layout's hstart[k]; one
// workgroup row per diagonal (grid y), FP32 widened.
template <typename T>
__global__ void repack_to_host_layout_kernel(int n, const int64_t *__restrict__ dstart,
                                             const int64_t *__restrict__ hstart, const int64_t *__restrict__ len,
                                             const T *__restrict__ re, const T *__restrict__ im,
                                             double *__restrict__ ore, double *__restrict__ oim) {
    for (int k = blockIdx.y; k < n; k += gridDim.y) {
        const int64_t L = len[k], s = dstart[k], h = hstart[k];
        for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < L; i += (int64_t)gridDim.x * blockDim.x) {
            ore[h + i] = static_cast<double>(re[s + i]);
            oim[h + i] = static_cast<double>(im[s + i]);
        }
    }
}

extern "C" int dq_devmat_download(const dq_devmat *A, const int64_t *starts, double *re, double *im) {
    return guard([&] {
        if (!A) throw ArgError("matrix is NULL");
        if (A->n() && (!starts || !re || !im)) throw ArgError("dq_devmat_download: NULL host array");
        if (!A->total || !A->n()) return;
        // the host layout's extent; its gaps are written as zeros
        int64_t lo = INT64_MAX, hi = 0, maxlen = 0;
        for (int k = 0; k < A->n(); ++k) {
            if (starts[k] < 0) throw ArgError("dq_devmat_download: negative start");
            lo = std::min(lo, starts[k]);
            hi = std::max(hi, starts[k] + A->len[k]);
            maxlen = std::max(maxlen, A->len[k]);
        }
        if (hi <= lo) return;
        // repack on the device into the host layout (range [lo, hi)), then one copy per array into the caller's
        // arrays: no host-side staging buffer or per-diagonal copies
        const int64_t ext = hi - lo;
        DevBuf br(static_cast<size_t>(ext) * 8), bi(static_cast<size_t>(ext) * 8);
        zero_bytes(br.p, br.bytes, 0);
        zero_bytes(bi.p, bi.bytes, 0);
        std::vector<int64_t> rel(A->n());
        for (int k = 0; k < A->n(); ++k) rel[k] = starts[k] - lo;
        DevBuf dst(sizeof(int64_t) * A->n()), hst(sizeof(int64_t) * A->n()), dln(sizeof(int64_t) * A->n());
        DQ_HIP(hipMemcpy(dst.p, A->start.data(), dst.bytes, hipMemcpyHostToDevice));
        DQ_HIP(hipMemcpy(hst.p, rel.data(), hst.bytes, hipMemcpyHostToDevice));
        DQ_HIP(hipMemcpy(dln.p, A->len.data(), dln.bytes, hipMemcpyHostToDevice));
        const dim3 grid(std::max<unsigned>(1, blocks_for(maxlen, 256, 64)),
                        static_cast<unsigned>(std::min(A->n(), 65535)));
        if (A->precision == DQ_F64)
            hipLaunchKernelGGL(repack_to_host_layout_kernel<double>, grid, dim3(256), 0, 0, A->n(), dst.as<int64_t>(),
                               hst.as<int64_t>(), dln.as<int64_t>(), A->re.as<double>(), A->im.as<double>(),
                               br.as<double>(), bi.as<double>());
        else
            hipLaunchKernelGGL(repack_to_host_layout_kernel<float>, grid, dim3(256), 0, 0, A->n(), dst.as<int64_t>(),
                               hst.as<int64_t>(), dln.as<int64_t>(), A->re.as<float>(), A->im.as<float>(),
                               br.as<double>(), bi.as<double>());
        DQ_LAUNCH_CHECK();
        DQ_HIP(hipMemcpy(re + lo, br.p, br.bytes, hipMemcpyDeviceToHost));
        DQ_HIP(hipMemcpy(im + lo, bi.p, bi.bytes, hipMemcpyDeviceToHost));
    });
}

extern "C" int dq_devmat_get_diagonal(const dq_devmat *A, int32_t offset, double *re, double *im) {
    return guard([&] {
        if (!A) throw ArgError("matrix is NULL");
        int k = -1;
        for (int j = 0; j < A->n(); ++j)
            if (A->offs[j] == offset) { k = j; break; }
        if (k < 0) throw ArgError("diagonal index not present in dIndices.");
        const int64_t n = A->len[k];
        if (!n) return;
        if (!re || !im) throw ArgError("NULL host array");
        if (A->precision == DQ_F64) {
            DQ_HIP(hipMemcpy(re, A->re.as<double>() + A->start[k], n * 8, hipMemcpyDeviceToHost));
            DQ_HIP(hipMemcpy(im, A->im.as<double>() + A->start[k], n * 8, hipMemcpyDeviceToHost));
        } else {
            std::vector<float> a(n), b(n);
            DQ_HIP(hipMemcpy(a.data(), A->re.as<float>() + A->start[k], n * 4, hipMemcpyDeviceToHost));
            DQ_HIP(hipMemcpy(b.data(), A->im.as<float>() + A->start[k], n * 4, hipMemcpyDeviceToHost));
            std::copy(a.begin(), a.end(), re);
            std::copy(b.begin(), b.end(), im);
        }
    });
}

extern "C" void dq_devmat_destroy(dq_devmat *A) { delete A; }

extern "C" int dq_devmat_info(const dq_devmat *A, uint32_t *rows, uint32_t *cols, int32_t *n, int *precision,
                              int64_t *stored) {
    return guard([&] {
        if (!A) throw ArgError("matrix is NULL");
        if (rows) *rows = A->rows;
        if (cols) *cols = A->cols;
        if (n) *n = A->n();
        if (precision) *precision = A->precision;
        if (stored) *stored = A->stored();
    });
}

extern "C" int dq_devmat_offsets(const dq_devmat *A, int32_t *out) {
    return guard([&] {
        if (!A) throw ArgError("matrix is NULL");
        if (A->n() && !out) throw ArgError("out is NULL");
        std::copy(A->offs.begin(), A->offs.end(), out);
    });
}

extern "C" int dq_prune(const dq_devmat *A, dq_devmat **out) {
    return guard([&] {
        if (!A || !out) throw ArgError("NULL argument");
        *out = prune(*A, 0).release();
    });
}

extern "C" int dq_estimate_1norm(const dq_devmat *A, double *out) {
    return guard([&] {
        if (!A || !out) throw ArgError("NULL argument");
        *out = estimate_1norm(*A, 0);
    });
}

extern "C" int dq_frobenius_norm(const dq_devmat *A, double *out) {
    return guard([&] {
        if (!A || !out) throw ArgError("NULL argument");
        *out = frobenius(*A, 0);
    });
}

extern "C" int dq_frobenius_diff(const dq_devmat *A, const dq_devmat *B, double *out) {
    return guard([&] {
        if (!A || !B || !out) throw ArgError("NULL argument");
        *out = frobenius_diff(*A, *B, 0);
    });
}

extern "C" int dq_trace(const dq_devmat *A, double *re, double *im) {
    return guard([&] {
        if (!A || !re || !im) throw ArgError("NULL argument");
        *re = 0.0;
        *im = 0.0;
        for (int k = 0; k < A->n(); ++k) {
            if (A->offs[k] == 0) {
                const size_t eb = elem_bytes(A->precision);
                *re = sum_range(static_cast<const char *>(A->re.p) + A->start[k] * eb, A->precision, A->len[k], 0);
                *im = sum_range(static_cast<const char *>(A->im.p) + A->start[k] * eb, A->precision, A->len[k], 0);
                return;
            }
        }
    });
}

extern "C" int dq_release_cached_memory(void) {
    return guard([&] {
        dq::release_dryrun_cache();
        dq::release_cached_memory();
    });
}
