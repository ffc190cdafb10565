// Internals shared by the DiaQ HIP library (gfx950): error handling, the device DIA layout, allocation, launch helpers.
//
// Device layout of a DIA matrix (dq_devmat):
//   - values in split real / imaginary arrays (structure of arrays), FP64 or FP32;
//   - diagonal k (offset d) covers rows [rlo, rhi) with rlo = max(0, -d) and rhi = rlo + len, where len is the
//     reference's diagonal length (src/deviceMatrix_gpu.cu:16-28); its element for row r, A[r, r + d], lives at index
//     rowbase + r. Each diagonal's storage start is padded so that rowbase is a multiple of DQ_ALIGN elements: a wave
//     working on 64 consecutive rows then reads whole, aligned 128-byte lines of every diagonal. Padding is zero.
//   - metadata (offset, rowbase, row range) in a device array of DiagMeta, plus a host copy.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <atomic>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "diaq_hip.h"

namespace dq {

// ------------------------------------------------------------------------------------------------------------------
// Errors: every C entry point runs its body under guard(), which turns exceptions into a status code plus a
// thread-local message (dq_last_error).
// ------------------------------------------------------------------------------------------------------------------
struct ArgError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct AllocError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void set_last_error(const std::string &msg);

template <typename F>
int guard(F &&f) {
    try {
        f();
        return DQ_OK;
    } catch (const ArgError &e) {
        set_last_error(e.what());
        return DQ_ERR_ARG;
    } catch (const AllocError &e) {
        set_last_error(e.what());
        return DQ_ERR_ALLOC;
    } catch (const HipError &e) {
        set_last_error(e.what());
        return DQ_ERR_HIP;
    } catch (const std::bad_alloc &) {
        set_last_error("host allocation failed");
        return DQ_ERR_ALLOC;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return DQ_ERR_STATE;
    }
}

#define DQ_HIP(call)                                                                                               \
    do {                                                                                                           \
        hipError_t dq_err_ = (call);                                                                               \
        if (dq_err_ != hipSuccess) {                                                                               \
            (void)hipGetLastError();                                                                               \
            throw ::dq::HipError(std::string("HIP error: ") + hipGetErrorString(dq_err_) + " (" + __FILE__ + ":" + \
                                 std::to_string(__LINE__) + ")");                                                  \
        }                                                                                                          \
    } while (0)

#define DQ_LAUNCH_CHECK() DQ_HIP(hipGetLastError())

inline hipStream_t as_stream(uintptr_t s) { return reinterpret_cast<hipStream_t>(s); }

// ------------------------------------------------------------------------------------------------------------------
// Device memory
// ------------------------------------------------------------------------------------------------------------------
void *dev_alloc(size_t bytes);          // device memory (a cached block or hipMalloc) with a clear error; nullptr for 0
void dev_free(void *p);                 // back to the cache (after the device is idle) or hipFree; never throws
void set_alloc_cache_budget(int64_t mib);  // bytes the cache of freed blocks may hold; -1: an eighth of the device
void release_cached_memory();           // hipFree every cached block
void release_dryrun_cache();            // drop the product dry-runs kept for reuse (product.hip)

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n) : p(dev_alloc(n)), bytes(n) {}
    ~DevBuf() { dev_free(p); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { dev_free(p); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    template <typename T> T *as() const { return static_cast<T *>(p); }
};

// ------------------------------------------------------------------------------------------------------------------
// DIA layout
// ------------------------------------------------------------------------------------------------------------------
constexpr int64_t DQ_ALIGN_BYTES = 128;

struct DiagMeta {
    int64_t rowbase;  // element index of row 0 on this diagonal (row r at rowbase + r)
    int64_t rlo;      // first row holding an element
    int64_t rhi;      // one past the last row
    int32_t off;      // diagonal offset d: row r meets column r + d
    int32_t pad_;
};
static_assert(sizeof(DiagMeta) == 32, "DiagMeta is 32 bytes");

// Length of diagonal d of a rows x cols matrix (src/deviceMatrix_gpu.cu:16-28).
inline int64_t diagonal_length(int32_t d, int64_t rows, int64_t cols) {
    int64_t n = d < 0 ? rows + d : (d > 0 ? cols - d : (rows < cols ? rows : cols));
    if (n > rows) n = rows;
    if (n > cols) n = cols;
    return n < 0 ? 0 : n;
}

inline int elem_bytes(int precision) { return precision == DQ_F32 ? 4 : 8; }

}  // namespace dq

// The opaque handle of the C ABI.
struct dq_devmat {
    uint32_t rows = 0, cols = 0;
    int precision = DQ_F64;
    std::vector<int32_t> offs;        // diagonal offsets in slot order (the caller's order is kept)
    std::vector<int64_t> start;       // storage start of each diagonal (elements)
    std::vector<int64_t> len;         // stored elements of each diagonal
    std::vector<dq::DiagMeta> meta;   // host copy of the device metadata
    int64_t total = 0;                // allocated elements per value array (padding included)
    dq::DevBuf re, im;                // values
    dq::DevBuf d_meta;                // DiagMeta[n]
    // SpMV schedule chosen by dq_spmv_tune for this matrix (0 = the library default): rows per lane, non-temporal
    // matrix loads (-1 = default), load schedule (-1 = default, 0 = per-row predicated, 1 = batched), LDS x window
    // (-1 = default, 0 = off, 1 = on).
    int tuned_rpt = 0, tuned_nt = -1, tuned_pipe = -1, tuned_lds = -1;
    // Row-tile order chosen by dq_spmv_tune (-1 = default round-robin, 0 = round-robin, 1 = XCD-contiguous).
    int tuned_xcd = -1;
    // Fused-evolution term schedule, timed on the first evolution of this matrix (0 = not yet): rows per lane (2, 4,
    // 8) and LDS x window (0, 1). A speed choice only: every schedule gives the same bits. Atomic: evolutions of one
    // matrix may run from several threads, and the first ones may each time and store a choice.
    mutable std::atomic<int> term_rpt{0}, term_lds{0};
    // SpMV schedule for products that skip the matrix's all-zero blocks, chosen by dq_spmv_tune (0 / -1 = the
    // library default): rows per lane, non-temporal loads, load schedule, LDS x window.
    int skip_rpt = 0, skip_nt = -1, skip_pipe = -1, skip_lds = -1;
    // The zero-block map (256-row blocks, zmap_words words per diagonal) dq_spmv_tune keeps when the matrix has
    // all-zero blocks; a call that rewrites the values (dq_devmat_set_packed_device) recomputes it in place.
    dq::DevBuf zmap;
    int64_t zmap_words = 0;
    bool zmap_valid = false;

    int n() const { return static_cast<int>(offs.size()); }
    int64_t stored() const {
        int64_t s = 0;
        for (int64_t l : len) s += l;
        return s;
    }
};

namespace dq {

// Compute the layout of (rows, cols, offsets) and allocate the values (zero-filled unless zero_fill is false, for a
// caller that writes every element) and the device metadata (stream-ordered).
std::unique_ptr<dq_devmat> make_devmat(uint32_t rows, uint32_t cols, const int32_t *offsets, int n, int precision,
                                       hipStream_t stream, bool zero_fill = true);

// Zero the padding of M's value arrays (before each diagonal's first stored row and after the last), for a matrix
// made with zero_fill = false whose stored elements a kernel writes.
void zero_gaps(dq_devmat &M, hipStream_t s);

// bytes zero bytes at p (16-byte aligned) by a fill kernel on s: cheaper on the host than the runtime's memset, and
// safe to capture into a graph.
void zero_bytes(void *p, size_t bytes, hipStream_t s);

// Launch geometry helpers.
inline unsigned blocks_for(int64_t work, int per_block, int64_t cap = 1 << 20) {
    int64_t b = (work + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return static_cast<unsigned>(b);
}

// Elementwise kernels over the stored values of matrices that share one layout (same shape and offsets).
void fill_identity(dq_devmat &M, hipStream_t s);  // M must have the single diagonal 0
void scale_real(dq_devmat &M, double factor, hipStream_t s);
void times_i_scale_copy(const dq_devmat &A, dq_devmat &C, double imag, hipStream_t s);  // C = (i*imag) A
// A + B over the ascending union of their diagonals (src/diaQMatrix.hpp:867-916); shapes must match.
std::unique_ptr<dq_devmat> union_add(const dq_devmat &A, const dq_devmat &B, hipStream_t s);
// Whether A and B share one layout (shape, precision, offsets in the same order), and A += B for such a pair.
bool same_layout(const dq_devmat &A, const dq_devmat &B);
void add_in_place(dq_devmat &A, const dq_devmat &B, hipStream_t s);
double frobenius(const dq_devmat &A, hipStream_t s);
double estimate_1norm(const dq_devmat &A, hipStream_t s);
std::unique_ptr<dq_devmat> prune(const dq_devmat &A, hipStream_t s);
std::unique_ptr<dq_devmat> copy_devmat(const dq_devmat &A, hipStream_t s);
// The same matrix (shape, offsets, layout) with its values rounded or widened to `precision`.
std::unique_ptr<dq_devmat> convert_devmat(const dq_devmat &A, int precision, hipStream_t s);

// Product tuning keys ("product_interleave", "product_nt_store"); false if the key is not one of them.
bool product_set_tuning(const std::string &key, int value);
// spmm_dia tuning key ("spmm_tiled"); false if it is another key.
bool spmm_set_tuning(const std::string &key, int value);
// Batched-evolution tuning keys ("evo_rows_per_lane", "evo_xcd", "evo_lds", "evo_skip_zero_blocks"); false if the
// key is not one of them.
bool evo_set_tuning(const std::string &key, int value);

// C = (A B) / divisor with the structural dry-run (product.hip). allow_dense: a near-dense product may take the dense
// path (dense.hip) -- same offsets, values equal up to the summation order; for pow and the matrix exponential.
std::unique_ptr<dq_devmat> product(const dq_devmat &A, const dq_devmat &B, double divisor, hipStream_t s,
                                   bool allow_dense = false);
// Dense path (dense.hip): expand to dense, the complex GEMM of gemm.hip, gather C's diagonals.
std::unique_ptr<dq_devmat> product_dense(const dq_devmat &A, const dq_devmat &B, double divisor, hipStream_t s);
bool dense_product_pays(const dq_devmat &A, const dq_devmat &B);
// Dense side of the dense path (builders.hip): every element of the rows x cols interleaved complex double array
// (leading dimension ld) from M's diagonals (zeros elsewhere), and M's stored elements back from such an array.
void dia_to_dense(const dq_devmat &M, double *dense, int64_t ld, hipStream_t s);
void dense_to_dia(dq_devmat &M, const double *dense, int64_t ld, hipStream_t s);
// C = (A B) / divisor for interleaved complex double row-major dense matrices, on the matrix cores (gemm.hip).
void zgemm_dense(int64_t M, int64_t N, int64_t K, const double *A, const double *B, double *C, double divisor,
                 hipStream_t s);
// Dense-path tuning key ("product_dense"); false if it is another key.
bool dense_set_tuning(const std::string &key, int value);
// Dense <-> DIA conversion tuning key ("dense_convert"); false if it is another key.
bool convert_set_tuning(const std::string &key, int value);
// Gate-pass tuning key ("gates_tile_qubits"); false if it is another key.
bool gates_set_tuning(const std::string &key, int value);
// Fused-evolution tuning keys ("hamsim_skip_zero_blocks", "hamsim_rows_per_lane", "hamsim_lds"); false if the key is
// not one of them.
bool hamsim_set_tuning(const std::string &key, int value);
// y (+)= (I_{2^before} (x) G (x) I_{2^after}) x on device vectors of G's precision (kron.hip); G square, checked by
// the caller; x == y allowed when G is at most 16 x 16 with at most 64 diagonals.
void kron_apply(const dq_devmat &G, int before, int after, const void *xr, const void *xi, void *yr, void *yi,
                bool accumulate, hipStream_t s);

// y = A x for device vectors of the matrix precision; col band [col_lo, col_hi) and accumulate for multi-pass use.
void spmv(const dq_devmat &A, const void *xr, const void *xi, void *yr, void *yi, hipStream_t s);

// Recompute the zero-block map dq_spmv_tune keeps in A's handle (if any) after A's values were rewritten on s.
void refresh_zero_map(dq_devmat &A, hipStream_t s);

// Whether the LDS x window is the untuned default for A at rows-per-lane rpt (some off-diagonals within its halo).
bool lds_window_pays(const dq_devmat &A, int rpt);
// Whether row-tile kernels over A at rows-per-lane rpt stage the x window in LDS: the process-wide knob if set, else
// the tuned choice, else lds_window_pays.
bool lds_window_for(const dq_devmat &A, int rpt);
// Row-tile order of row-tile kernels over A (0 = round-robin over the XCDs, 1 = XCD-contiguous): the process-wide knob
// if set, else the tuned choice, else XCD-contiguous when an eighth of x fits an XCD's L2 and diagonals are near.
int xcd_tiles_for(const dq_devmat &A);

}  // namespace dq
