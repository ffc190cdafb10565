// DIA SpMV on gfx950: y = A x with A, x, y resident in HBM.
//
// Reference: spmv_single_diag_kernel / spmv_row_kernel / spmv_device (src/spMV_gpu.cu:17-168), spmv_host_vectors
// (:171-183), spMV_ntimes_cuda (src/spMV_ntimes_gpu.cu:61-263).
//
// One kernel serves every diagonal count (the reference's one-diagonal kernel gives the same values: y is zero
// except on the rows of the lone diagonal). A workgroup owns a tile of 256*RPT consecutive rows; lane l of the tile
// handles rows tile + l + 256*j. For each diagonal (a uniform loop, metadata through scalar loads) the wave reads 64
// consecutive values of the diagonal (one aligned 512-byte FP64 segment, row-aligned layout) and 64 consecutive
// values of the shifted x window, so every access is a full, coalesced line. Each row sums its diagonals in slot
// order in registers and y is written once: A is read once, y written once, and x is read once per diagonal (the
// compulsory bytes count it once; re-reads hit L2 or the Infinity Cache when the windows are near each other).
//
// Column bands: with n_bands > 1 the kernel runs once per band of columns [c0, c1), each pass adding only the
// elements whose column lies in the band (the first pass writes y, later passes add to it). A pass touches only
// 1/n_bands of x, so its x lines stay in the 256 MB Infinity Cache between the diagonals that share them, at the cost
// of one read and one write of y per extra pass.
//
// LDS x window: with lds on, each workgroup stages x[tile - TILE/2, tile + 3 TILE/2) in LDS and the diagonals with
// |offset| <= TILE/2 read x from there (bands, qubit Hamiltonians); scattered offsets gain nothing from it. The
// per-matrix tuner picks it with the other schedule choices.
//
// FP32: values and vectors are float; each row's diagonal sum is Kahan-compensated in FP32.
#include <algorithm>
#include <cstdlib>
#include <mutex>

#include "dia_rows.hpp"

namespace dq {

// Process-wide tuning knobs (set through dq_set_tuning).
static int g_spmv_rpt = 0;    // rows per lane; 0 = automatic
static int g_spmv_bands = 0;  // column bands; 0 = automatic
static int g_spmv_nt = -1;    // non-temporal matrix loads: -1 = automatic, 0 = off, 1 = on
static int g_spmv_pipe = -1;  // load schedule: -1 = automatic, 0 = per-row predicated, 1 = batched
static int g_spmv_lds = -1;   // LDS x window: -1 = automatic, 0 = off, 1 = on
static int g_spmv_xcd = -1;   // tile order: -1 = automatic, 0 = round-robin over the XCDs, 1 = XCD-contiguous

// ZM: zm is a zero-block map of A (256-row blocks, zwords words per diagonal) whose all-zero blocks are skipped: the
// map dq_spmv_tune keeps in the handle (dropped when the values are rewritten), or one spMV_ntimes makes for a call.
template <typename T, int RPT, bool NT, bool PIPE, bool LDS, bool ZM>
__global__ __launch_bounds__(256) void spmv_rows_kernel(int64_t rows, int64_t cols, int n,
                                                        const DiagMeta *__restrict__ meta, const T *__restrict__ are,
                                                        const T *__restrict__ aim, const T *__restrict__ xr,
                                                        const T *__restrict__ xi, T *__restrict__ yr,
                                                        T *__restrict__ yi, int64_t col_lo, int64_t col_hi,
                                                        int accumulate, int64_t ntiles, int xcd,
                                                        const uint32_t *__restrict__ zm, int64_t zwords) {
    constexpr int TILE = 256 * RPT;
    constexpr int HALO = LDS ? TILE / 2 : 0;
    constexpr int W = LDS ? TILE + 2 * HALO : 1;
    __shared__ T sr[W], si[W];
    const int64_t t = tile_of_block(blockIdx.x, ntiles, xcd);
    if (t >= ntiles) return;  // uniform: a surplus workgroup of the XCD-contiguous grid
    const int64_t tile = t * TILE;
    const int64_t tile_end = tile + TILE < rows ? tile + TILE : rows;
    if constexpr (LDS) stage_x_window<T, RPT, HALO>(tile, cols, xr, xi, sr, si);
    Acc<T> ar[RPT], ai[RPT];
    if (accumulate) {
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const int64_t r = tile + threadIdx.x + 256 * j;
            if (r < rows) { ar[j].add(yr[r]); ai[j].add(yi[r]); }
        }
    }
    dia_rows_tile<T, RPT, NT, PIPE, HALO, ZM>(tile, tile_end, n, meta, are, aim, xr, xi, col_lo, col_hi, ar, ai, sr,
                                              si, zm, zwords);
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        const int64_t r = tile + threadIdx.x + 256 * j;
        if (r < rows) { yr[r] = ar[j].get(); yi[r] = ai[j].get(); }
    }
}

template <typename T>
static void launch_rows(const dq_devmat &A, const T *xr, const T *xi, T *yr, T *yi, int rpt, bool nt, bool pipe,
                        bool lds, int xcd, int64_t c0, int64_t c1, int accumulate, hipStream_t s,
                        const uint32_t *zm = nullptr, int64_t zwords = 0) {
    const int64_t rows = A.rows, cols = A.cols;
    const DiagMeta *meta = A.d_meta.as<DiagMeta>();
    const T *are = static_cast<const T *>(A.re.p), *aim = static_cast<const T *>(A.im.p);
    const int64_t ntiles = (rows + 256 * rpt - 1) / (256 * rpt);
    const int64_t grid = grid_for_tiles(ntiles, xcd);
    if (grid > 0x7fffffffLL) throw ArgError("spmv: too many rows");
#define DQ_SPMV_LAUNCH_Z(R, NT, PIPE, LDS, Z)                                                                    \
    hipLaunchKernelGGL((spmv_rows_kernel<T, R, NT, PIPE, LDS, Z>), dim3(static_cast<unsigned>(grid)), dim3(256), 0, \
                       s, rows, cols, A.n(), meta, are, aim, xr, xi, yr, yi, c0, c1, accumulate, ntiles, xcd, zm,    \
                       zwords)
#define DQ_SPMV_LAUNCH(R, NT, PIPE, LDS)                           \
    do {                                                           \
        if (zm) { DQ_SPMV_LAUNCH_Z(R, NT, PIPE, LDS, true); }      \
        else { DQ_SPMV_LAUNCH_Z(R, NT, PIPE, LDS, false); }        \
    } while (0)
#define DQ_SPMV_SCHED(R, LDS)                                         \
    if (nt && pipe) DQ_SPMV_LAUNCH(R, true, true, LDS);               \
    else if (nt) DQ_SPMV_LAUNCH(R, true, false, LDS);                 \
    else if (pipe) DQ_SPMV_LAUNCH(R, false, true, LDS);               \
    else DQ_SPMV_LAUNCH(R, false, false, LDS);
#define DQ_SPMV_CASE(R)                                         \
    case R:                                                     \
        if (lds) { DQ_SPMV_SCHED(R, true) }                     \
        else { DQ_SPMV_SCHED(R, false) }                        \
        break;
    switch (rpt) {
        DQ_SPMV_CASE(1)
        DQ_SPMV_CASE(2)
        DQ_SPMV_CASE(4)
        DQ_SPMV_CASE(8)
        DQ_SPMV_CASE(16)
        default: throw ArgError("spmv: rows per lane must be 1, 2, 4, 8 or 16");
    }
#undef DQ_SPMV_LAUNCH
#undef DQ_SPMV_LAUNCH_Z
#undef DQ_SPMV_SCHED
#undef DQ_SPMV_CASE
    DQ_LAUNCH_CHECK();
}

template <typename T>
__global__ void zero2_kernel(int64_t n, T *__restrict__ a, T *__restrict__ b) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        a[i] = T(0);
        b[i] = T(0);
    }
}

static int auto_bands(const dq_devmat &A) {
    if (g_spmv_bands > 0) return g_spmv_bands;
    return 1;
}

static int auto_rpt(const dq_devmat &A) {
    if (g_spmv_rpt > 0) return g_spmv_rpt;
    if (A.tuned_rpt > 0) return A.tuned_rpt;
    return 8;
}

static bool auto_nt(const dq_devmat &A) {
    if (g_spmv_nt >= 0) return g_spmv_nt == 1;
    if (A.tuned_nt >= 0) return A.tuned_nt == 1;
    return true;
}

static bool auto_pipe(const dq_devmat &A) {
    if (g_spmv_pipe >= 0) return g_spmv_pipe == 1;
    if (A.tuned_pipe >= 0) return A.tuned_pipe == 1;
    return true;
}

// Untuned default for the LDS window: on when at least two off-diagonals lie within the halo of a tile.
bool lds_window_pays(const dq_devmat &A, int rpt) {
    const int halo = 256 * rpt / 2;
    int near = 0;
    for (int32_t d : A.offs)
        if (d != 0 && d >= -halo && d <= halo) ++near;
    return near >= 2;
}

// The schedule of SpMVs that skip A's zero blocks: fewer live diagonals per tile favour smaller tiles and
// per-row loads (measured on a 24-qubit TFIM: 4 rows per lane, per-row predicated non-temporal loads, LDS window).
static int skip_rpt_for(const dq_devmat &A) {
    if (g_spmv_rpt > 0) return g_spmv_rpt;
    return A.skip_rpt > 0 ? A.skip_rpt : 4;
}
static bool skip_nt_for(const dq_devmat &A) {
    if (g_spmv_nt >= 0) return g_spmv_nt == 1;
    return A.skip_nt >= 0 ? A.skip_nt == 1 : true;
}
static bool skip_pipe_for(const dq_devmat &A) {
    if (g_spmv_pipe >= 0) return g_spmv_pipe == 1;
    return A.skip_pipe >= 0 ? A.skip_pipe == 1 : false;
}
static bool skip_lds_for(const dq_devmat &A, int rpt) {
    if (g_spmv_lds >= 0) return g_spmv_lds == 1;
    return A.skip_lds >= 0 ? A.skip_lds == 1 : lds_window_pays(A, rpt);
}

int xcd_tiles_for(const dq_devmat &A) {
    if (g_spmv_xcd >= 0) return g_spmv_xcd;
    if (A.tuned_xcd >= 0) return A.tuned_xcd;
    // untuned: XCD-contiguous when an XCD's eighth of x fits its 4 MB L2 and some diagonals are near each other
    const int64_t x_bytes = static_cast<int64_t>(A.cols) * 2 * elem_bytes(A.precision);
    return (x_bytes / 8 <= (int64_t(4) << 20) && lds_window_pays(A, 8)) ? 1 : 0;
}

bool lds_window_for(const dq_devmat &A, int rpt) {
    if (g_spmv_lds >= 0) return g_spmv_lds == 1;
    if (A.tuned_lds >= 0) return A.tuned_lds == 1;
    return lds_window_pays(A, rpt);
}

// A complete SpMV schedule, for callers that choose one explicitly (the tuner's candidates).
struct SpmvSched {
    int rpt;
    bool nt, pipe, lds;
    int xcd;
};

// y = A x in n_bands column passes (0: automatic), skipping A's all-zero blocks when zm is given. The schedule is
// `force` if given, else the process-wide knobs, else A's tuned choice, else the defaults.
void spmv_banded(const dq_devmat &A, const void *xr, const void *xi, void *yr, void *yi, int n_bands,
                 hipStream_t s, const uint32_t *zm = nullptr, int64_t zwords = 0, const SpmvSched *force = nullptr) {
    if (A.rows == 0) return;
    if (A.n() == 0) {  // no diagonals: y = 0 (a fill kernel, safe inside graph capture)
        const unsigned g = blocks_for(A.rows, 256, 4096);
        if (A.precision == DQ_F64)
            hipLaunchKernelGGL(zero2_kernel<double>, dim3(g), dim3(256), 0, s, (int64_t)A.rows, (double *)yr,
                               (double *)yi);
        else
            hipLaunchKernelGGL(zero2_kernel<float>, dim3(g), dim3(256), 0, s, (int64_t)A.rows, (float *)yr,
                               (float *)yi);
        DQ_LAUNCH_CHECK();
        return;
    }
    if (n_bands <= 0) n_bands = auto_bands(A);
    if (n_bands > static_cast<int>(A.cols)) n_bands = std::max<int>(1, A.cols);
    const int rpt = force ? force->rpt : zm ? skip_rpt_for(A) : auto_rpt(A);
    const bool nt = force ? force->nt : zm ? skip_nt_for(A) : auto_nt(A);
    const bool pipe = force ? force->pipe : zm ? skip_pipe_for(A) : auto_pipe(A);
    const bool lds = force ? force->lds : zm ? skip_lds_for(A, rpt) : lds_window_for(A, rpt);
    const int xcd = force ? force->xcd : xcd_tiles_for(A);
    const int64_t cols = A.cols;
    for (int b = 0; b < n_bands; ++b) {
        const int64_t c0 = cols * b / n_bands, c1 = cols * (b + 1) / n_bands;
        if (A.precision == DQ_F64)
            launch_rows<double>(A, (const double *)xr, (const double *)xi, (double *)yr, (double *)yi, rpt, nt, pipe,
                                lds, xcd, c0, c1, b > 0, s, zm, zwords);
        else
            launch_rows<float>(A, (const float *)xr, (const float *)xi, (float *)yr, (float *)yi, rpt, nt, pipe, lds,
                               xcd, c0, c1, b > 0, s, zm, zwords);
    }
}

static int g_spmv_skip = 1;  // "spmv_skip_zero_blocks"

// Whether a sequence of n SpMVs on A's values should make its own zero-block map: not for a single product (the map
// would cost about as much as the product saves), not when the handle already holds a valid one, nor with the skip
// switched off.
static bool ntimes_uses_map(const dq_devmat &A, int n) {
    return g_spmv_skip && n >= 2 && A.n() > 0 && A.rows > 0 && !A.zmap_valid;
}

// The map a product with A may use without making one: the handle's, for FP64 matrices. (In FP64 a skipped block
// changes no bit; a compensated FP32 sum would change in its last bits, and a single FP32 SpMV keeps the same bits
// whether or not A was tuned.)
static const uint32_t *handle_map(const dq_devmat &A) {
    return (g_spmv_skip && A.zmap_valid && A.precision == DQ_F64) ? A.zmap.as<uint32_t>() : nullptr;
}

// The handle's map for a sequence of n products: a single product follows handle_map; two or more use the map in
// any precision (their results are not compared with a single untuned SpMV's).
static const uint32_t *sequence_map(const dq_devmat &A, int n) {
    if (n == 1) return handle_map(A);
    return (g_spmv_skip && A.zmap_valid) ? A.zmap.as<uint32_t>() : nullptr;
}

static size_t ntimes_map_bytes(const dq_devmat &A, int64_t &zwords) {
    zwords = zero_block_words(A.rows, 256);
    return sizeof(uint32_t) * static_cast<size_t>(zwords) * static_cast<size_t>(A.n());
}

// Whether the map z of A shows a 256-row block, inside some diagonal's rows, whose values are all zero.
static bool has_zero_blocks(const dq_devmat &A, const uint32_t *z, int64_t zwords, hipStream_t s) {
    std::vector<uint32_t> h(static_cast<size_t>(zwords) * A.n());
    DQ_HIP(hipMemcpyAsync(h.data(), z, h.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    DQ_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < A.n(); ++k) {
        const DiagMeta &m = A.meta[k];
        if (m.rhi <= m.rlo) continue;
        for (int64_t b = m.rlo / 256; b <= (m.rhi - 1) / 256; ++b)
            if (!((h[static_cast<size_t>(k) * zwords + (b >> 5)] >> (b & 31)) & 1u)) return true;
    }
    return false;
}

// A's zero-block map (256-row blocks) into z.
static void ntimes_zero_map(const dq_devmat &A, int64_t zwords, uint32_t *z, hipStream_t s) {
    const dim3 grid(static_cast<unsigned>(zwords), zero_map_grid_y(A.n()), 1);
    if (A.precision == DQ_F64)
        hipLaunchKernelGGL((zero_blocks_kernel<double, 256>), grid, dim3(256), 0, s, static_cast<int64_t>(A.rows),
                           A.n(), A.d_meta.as<DiagMeta>(), int64_t(0), A.re.as<double>(), A.im.as<double>(), zwords, z);
    else
        hipLaunchKernelGGL((zero_blocks_kernel<float, 256>), grid, dim3(256), 0, s, static_cast<int64_t>(A.rows),
                           A.n(), A.d_meta.as<DiagMeta>(), int64_t(0), A.re.as<float>(), A.im.as<float>(), zwords, z);
    DQ_LAUNCH_CHECK();
}

// After A's values were rewritten on stream s: recompute the zero-block map the handle keeps into the same buffer,
// ordered after the rewrite. A graph captured earlier that reads the handle's map then sees the new map on its next
// replay, and later SpMVs skip exactly the blocks that are zero now (none, if every block became nonzero).
void refresh_zero_map(dq_devmat &A, hipStream_t s) {
    if (!A.zmap_valid || !A.zmap.p) return;
    ntimes_zero_map(A, A.zmap_words, A.zmap.as<uint32_t>(), s);
}

// y = A x: the schedule of A, skipping its all-zero blocks when the handle holds their map (see handle_map).
void spmv(const dq_devmat &A, const void *xr, const void *xi, void *yr, void *yi, hipStream_t s) {
    const uint32_t *zm = handle_map(A);
    spmv_banded(A, xr, xi, yr, yi, 0, s, zm, zm ? A.zmap_words : 0);
}

}  // namespace dq

using namespace dq;

extern "C" int dq_set_tuning(const char *key, int value) {
    return guard([&] {
        if (!key) throw ArgError("key is NULL");
        const std::string k(key);
        if (k == "spmv_rows_per_lane") {
            if (value != 0 && value != 1 && value != 2 && value != 4 && value != 8 && value != 16)
                throw ArgError("spmv_rows_per_lane must be 0, 1, 2, 4, 8 or 16");
            g_spmv_rpt = value;
        } else if (k == "spmv_nontemporal") {
            if (value < -1 || value > 1) throw ArgError("spmv_nontemporal must be -1, 0 or 1");
            g_spmv_nt = value;
        } else if (k == "spmv_pipelined") {
            if (value < -1 || value > 1) throw ArgError("spmv_pipelined must be -1, 0 or 1");
            g_spmv_pipe = value;
        } else if (k == "spmv_lds") {
            if (value < -1 || value > 1) throw ArgError("spmv_lds must be -1, 0 or 1");
            g_spmv_lds = value;
        } else if (k == "spmv_xcd") {
            if (value < -1 || value > 1) throw ArgError("spmv_xcd must be -1, 0 or 1");
            g_spmv_xcd = value;
        } else if (k == "spmv_bands") {
            if (value < 0 || value > 4096) throw ArgError("spmv_bands must be in [0, 4096]");
            g_spmv_bands = value;
        } else if (k == "alloc_cache_mb") {
            if (value < -1) throw ArgError("alloc_cache_mb must be -1 (an eighth of the device memory) or >= 0");
            set_alloc_cache_budget(value);
        } else if (k == "spmv_skip_zero_blocks") {
            if (value != 0 && value != 1) throw ArgError("spmv_skip_zero_blocks must be 0 or 1");
            g_spmv_skip = value;
        } else if (!product_set_tuning(k, value) && !evo_set_tuning(k, value) && !spmm_set_tuning(k, value) &&
                   !dense_set_tuning(k, value) && !convert_set_tuning(k, value) && !gates_set_tuning(k, value) &&
                   !hamsim_set_tuning(k, value)) {
            throw ArgError("unknown tuning key: " + k);
        }
    });
}

// Time every (load schedule, rows per lane, non-temporal, LDS window) candidate on scratch vectors and keep the fastest
// in the matrix handle; later SpMVs with this matrix use it (unless a process-wide knob overrides it). The LDS window
// is only tried when some off-diagonal lies within its halo. The candidates are passed to the kernels explicitly, so
// SpMVs of other matrices meanwhile keep their own schedules; calls with this matrix must not overlap its tuning
// (the handle's schedule and zero-block map change during it).
extern "C" int dq_spmv_tune(dq_devmat *A, int repeats, uintptr_t stream, float *best_ms) {
    return guard([&] {
        if (!A) throw ArgError("matrix is NULL");
        if (repeats <= 0) repeats = 3;
        hipStream_t s = as_stream(stream);
        if (A->rows == 0 || A->n() == 0) return;
        const size_t eb = elem_bytes(A->precision);
        DevBuf xr(A->cols * eb), xi(A->cols * eb), yr(A->rows * eb), yi(A->rows * eb);
        zero_bytes(xr.p, xr.bytes, s);
        zero_bytes(xi.p, xi.bytes, s);
        hipEvent_t e0, e1;
        DQ_HIP(hipEventCreate(&e0));
        DQ_HIP(hipEventCreate(&e1));
        struct Events {
            hipEvent_t a, b;
            ~Events() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
        } ev{e0, e1};
        const uint32_t *tz = nullptr;  // the zero-block map while the skip schedule is timed
        int64_t tzw = 0;
        auto time_it = [&](const SpmvSched &c) {
            spmv_banded(*A, xr.p, xi.p, yr.p, yi.p, 1, s, tz, tzw, &c);  // warm
            DQ_HIP(hipEventRecord(e0, s));
            for (int r = 0; r < repeats; ++r) spmv_banded(*A, xr.p, xi.p, yr.p, yi.p, 1, s, tz, tzw, &c);
            DQ_HIP(hipEventRecord(e1, s));
            DQ_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            DQ_HIP(hipEventElapsedTime(&ms, e0, e1));
            return ms / repeats;
        };
        float best = 1e30f;
        SpmvSched win{8, true, true, false, 0};
        for (int lds = 0; lds <= 1; ++lds)
            for (int pipe = 0; pipe <= 1; ++pipe)
                for (int rpt : {2, 4, 8, 16})
                    for (int nt = 0; nt <= 1; ++nt) {
                        if (lds && !lds_window_pays(*A, rpt) && !lds_window_pays(*A, 16)) continue;
                        const SpmvSched c{rpt, nt == 1, pipe == 1, lds == 1, 0};
                        const float ms = time_it(c);
                        if (ms < best) { best = ms; win = c; }
                    }
        // second stage: the winner with the XCD-contiguous tile order
        SpmvSched wx = win;
        wx.xcd = 1;
        const float msx = time_it(wx);
        if (msx < best) { best = msx; win = wx; }
        A->tuned_rpt = win.rpt; A->tuned_nt = win.nt; A->tuned_pipe = win.pipe; A->tuned_lds = win.lds;
        A->tuned_xcd = win.xcd;
        // third stage: when A has all-zero blocks, keep their map in the handle and time the schedules that skip
        // them (used by A's FP64 SpMVs and by spMV_ntimes)
        A->skip_rpt = win.rpt; A->skip_nt = win.nt; A->skip_pipe = win.pipe; A->skip_lds = win.lds;
        A->zmap_valid = false;
        if (g_spmv_skip) {
            const size_t zb = ntimes_map_bytes(*A, tzw);
            if (A->zmap.bytes != zb) A->zmap = DevBuf(zb);
            A->zmap_words = tzw;
            ntimes_zero_map(*A, tzw, A->zmap.as<uint32_t>(), s);
            if (has_zero_blocks(*A, A->zmap.as<uint32_t>(), tzw, s)) {
                tz = A->zmap.as<uint32_t>();
                float zbest = 1e30f;
                SpmvSched zw{4, true, false, false, win.xcd};
                for (int lds = 0; lds <= 1; ++lds)
                    for (int pipe = 0; pipe <= 1; ++pipe)
                        for (int rpt : {1, 2, 4, 8, 16})
                            for (int nt = 0; nt <= 1; ++nt) {
                                if (lds && !lds_window_pays(*A, rpt)) continue;
                                const SpmvSched c{rpt, nt == 1, pipe == 1, lds == 1, win.xcd};
                                const float zms = time_it(c);
                                if (zms < zbest) { zbest = zms; zw = c; }
                            }
                tz = nullptr;
                A->skip_rpt = zw.rpt; A->skip_nt = zw.nt; A->skip_pipe = zw.pipe; A->skip_lds = zw.lds;
                A->zmap_valid = true;
            }
        }
        if (best_ms) *best_ms = best;
    });
}

extern "C" int dq_spmv_tuned_config(const dq_devmat *A, int *rows_per_lane, int *nontemporal, int *pipelined,
                                    int *lds_window, int *xcd_tiles) {
    return guard([&] {
        if (!A) throw ArgError("matrix is NULL");
        if (rows_per_lane) *rows_per_lane = auto_rpt(*A);
        if (nontemporal) *nontemporal = auto_nt(*A) ? 1 : 0;
        if (pipelined) *pipelined = auto_pipe(*A) ? 1 : 0;
        if (lds_window) *lds_window = lds_window_for(*A, auto_rpt(*A)) ? 1 : 0;
        if (xcd_tiles) *xcd_tiles = xcd_tiles_for(*A);
    });
}

extern "C" int dq_spmv_device(const dq_devmat *A, const void *xr, const void *xi, void *yr, void *yi,
                              uintptr_t stream) {
    return guard([&] {
        if (!A) throw ArgError("matrix is NULL");
        if (A->rows && (!yr || !yi)) throw ArgError("spmv_device: NULL y");
        if (A->n() && (!xr || !xi)) throw ArgError("spmv_device: NULL x");
        spmv(*A, xr, xi, yr, yi, as_stream(stream));
    });
}

extern "C" int dq_spmv_device_banded(const dq_devmat *A, const void *xr, const void *xi, void *yr, void *yi,
                                     int n_bands, uintptr_t stream) {
    return guard([&] {
        if (!A) throw ArgError("matrix is NULL");
        if (n_bands < 0) throw ArgError("n_bands must be >= 0");
        if (A->rows && (!yr || !yi)) throw ArgError("spmv_device: NULL y");
        if (A->n() && (!xr || !xi)) throw ArgError("spmv_device: NULL x");
        spmv_banded(*A, xr, xi, yr, yi, n_bands, as_stream(stream));
    });
}

// Host vectors of doubles through device buffers of the matrix precision.
namespace {
struct DevVec {
    DevBuf r, i;
    int64_t n;
    int prec;
    DevVec(int64_t n_, int prec_) : r(n_ * elem_bytes(prec_)), i(n_ * elem_bytes(prec_)), n(n_), prec(prec_) {}
    void put(const double *hr, const double *hi) {
        if (!n) return;
        if (prec == DQ_F64) {
            DQ_HIP(hipMemcpy(r.p, hr, n * 8, hipMemcpyHostToDevice));
            DQ_HIP(hipMemcpy(i.p, hi, n * 8, hipMemcpyHostToDevice));
        } else {
            std::vector<float> a(hr, hr + n), b(hi, hi + n);
            DQ_HIP(hipMemcpy(r.p, a.data(), n * 4, hipMemcpyHostToDevice));
            DQ_HIP(hipMemcpy(i.p, b.data(), n * 4, hipMemcpyHostToDevice));
        }
    }
    void get(double *hr, double *hi) const {
        if (!n) return;
        if (prec == DQ_F64) {
            DQ_HIP(hipMemcpy(hr, r.p, n * 8, hipMemcpyDeviceToHost));
            DQ_HIP(hipMemcpy(hi, i.p, n * 8, hipMemcpyDeviceToHost));
        } else {
            std::vector<float> a(n), b(n);
            DQ_HIP(hipMemcpy(a.data(), r.p, n * 4, hipMemcpyDeviceToHost));
            DQ_HIP(hipMemcpy(b.data(), i.p, n * 4, hipMemcpyDeviceToHost));
            std::copy(a.begin(), a.end(), hr);
            std::copy(b.begin(), b.end(), hi);
        }
    }
};
}  // namespace

extern "C" int dq_spmv_host(const dq_devmat *A, const double *xr, const double *xi, double *yr, double *yi) {
    return guard([&] {
        if (!A) throw ArgError("matrix is NULL");
        if ((A->cols && (!xr || !xi)) || (A->rows && (!yr || !yi))) throw ArgError("spmv_cuda: NULL vector");
        DevVec x(A->cols, A->precision), y(A->rows, A->precision);
        x.put(xr, xi);
        spmv(*A, x.r.p, x.i.p, y.r.p, y.i.p, 0);
        DQ_LAUNCH_CHECK();
        y.get(yr, yi);
    });
}

// y = A^n x on device vectors (src/spMV_ntimes_gpu.cu:61-263 on resident data). `fixed_map` (may be null) is
// device memory of fixed_words words the call may use for A's zero-block map instead of allocating one: the
// allocation-free entry point passes its static scratch array; then the map is made even while the stream is being
// captured (the array outlives every replay), and a map that would not fit is an error.
static void ntimes_device(const dq_devmat *A, const void *xr, const void *xi, void *yr, void *yi, int n, void *work_r,
                          void *work_i, hipStream_t s, uint32_t *fixed_map, int64_t fixed_words) {
    if (!A) throw ArgError("matrix is NULL");
    if (n < 0) throw ArgError("spmv_ntimes_cuda: num_times must be >= 0");
    if (!xr || !xi || !yr || !yi) throw ArgError("spmv_ntimes_device: NULL vector");
    const int64_t M = A->rows, N = A->cols;
    const size_t eb = elem_bytes(A->precision);
    if (n == 0) {
        if (M != N) throw ArgError("spmv_ntimes_cuda: num_times==0 requires square matrix (M==N) to copy x -> y");
        if (xr != yr) DQ_HIP(hipMemcpyAsync(yr, xr, M * eb, hipMemcpyDeviceToDevice, s));
        if (xi != yi) DQ_HIP(hipMemcpyAsync(yi, xi, M * eb, hipMemcpyDeviceToDevice, s));
        return;
    }
    if (A->n() == 0) {  // no diagonals: y = 0 for any num_times >= 1 (src/spMV_ntimes_gpu.cu:98-105)
        spmv_banded(*A, xr, xi, yr, yi, 0, s);
        return;
    }
    if (n > 1 && M != N)
        throw ArgError("spmv_ntimes_cuda: repeated SpMV requires square matrix (M==N) for num_times > 1");
    const bool in_place = xr == yr || xi == yi;
    if (in_place && (xr != yr || xi != yi)) throw ArgError("spmv_ntimes_device: x and y overlap in one part only");
    if ((n > 1 || in_place) && (!work_r || !work_i))
        throw ArgError("spmv_ntimes_device: work vectors are needed for num_times > 1 or x == y");
    // destinations alternate between y and the work vectors so that the last product lands in y; with x == y and an
    // odd count the first product must not overwrite x, so the order flips and a copy ends the sequence
    const bool flip = in_place && (n % 2 == 1);
    // the zero blocks of A, skipped by every product of the sequence (A cannot change during the call): the map
    // dq_spmv_tune kept in the handle, or else, for two or more products, one made here -- in the caller's fixed
    // memory, or in stream-ordered memory so the call stays asynchronous. Without fixed memory no map is made while
    // the stream is being captured into a graph (a replay must not read memory this call has given back).
    int64_t zwords = 0;
    uint32_t *zmap = nullptr;
    const uint32_t *zuse = sequence_map(*A, n);
    if (zuse) zwords = A->zmap_words;
    if (!zuse && ntimes_uses_map(*A, n)) {
        const size_t bytes = ntimes_map_bytes(*A, zwords);
        if (fixed_map) {
            if (static_cast<int64_t>(bytes / sizeof(uint32_t)) > fixed_words)
                throw ArgError("spmv_ntimes_device_noalloc: the zero-block map of this matrix needs " +
                               std::to_string(bytes) + " bytes, more than the " +
                               std::to_string(fixed_words * sizeof(uint32_t)) +
                               " bytes of the library's static scratch (tune the matrix with dq_spmv_tune to keep "
                               "its map in the handle, or use dq_spmv_ntimes_device)");
            ntimes_zero_map(*A, zwords, fixed_map, s);
            zuse = fixed_map;
        } else {
            hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
            DQ_HIP(hipStreamIsCapturing(s, &cap));
            if (cap == hipStreamCaptureStatusNone) {
                if (hipMallocAsync(reinterpret_cast<void **>(&zmap), bytes, s) != hipSuccess) {
                    (void)hipGetLastError();  // no stream-ordered memory: run without the map
                    zmap = nullptr;
                } else {
                    ntimes_zero_map(*A, zwords, zmap, s);
                    zuse = zmap;
                }
            }
        }
    }
    const void *sr = xr, *si = xi;
    for (int t = 0; t < n; ++t) {
        const bool to_y = (((n - 1 - t) % 2) == 0) != flip;
        void *dr = to_y ? yr : work_r, *di = to_y ? yi : work_i;
        spmv_banded(*A, sr, si, dr, di, 0, s, zuse, zuse ? zwords : 0);
        sr = dr;
        si = di;
    }
    if (zmap) DQ_HIP(hipFreeAsync(zmap, s));
    if (flip) {
        DQ_HIP(hipMemcpyAsync(yr, work_r, M * eb, hipMemcpyDeviceToDevice, s));
        DQ_HIP(hipMemcpyAsync(yi, work_i, M * eb, hipMemcpyDeviceToDevice, s));
    }
}

extern "C" int dq_spmv_ntimes_device(const dq_devmat *A, const void *xr, const void *xi, void *yr, void *yi, int n,
                                     void *work_r, void *work_i, uintptr_t stream) {
    return guard([&] { ntimes_device(A, xr, xi, yr, yi, n, work_r, work_i, as_stream(stream), nullptr, 0); });
}

// The allocation-free form's scratch: room for the zero-block map of a sequence (256-row blocks, one bit each, for
// every diagonal), 16 MiB per GPU, held from the library's load until the process ends. In host code `scratch` only
// stands in for the symbol; its device address comes from hipGetSymbolAddress, once per device.
constexpr int64_t NOALLOC_SCRATCH_WORDS = int64_t(1) << 22;
static __device__ __align__(256) uint32_t scratch[NOALLOC_SCRATCH_WORDS];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static uint32_t *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// This device's copy of `scratch`, looked up on the first call on the device (with that device current).
static uint32_t *noalloc_scratch() {
    int dev = 0;
    DQ_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) throw ArgError("spmv_ntimes_device_noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        DQ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = static_cast<uint32_t *>(p);
    }
    return g_scratch_ptr[dev];
}

extern "C" int dq_spmv_ntimes_device_noalloc(const dq_devmat *A, const void *xr, const void *xi, void *yr, void *yi,
                                             int n, void *work_r, void *work_i, uintptr_t stream) {
    return guard([&] {
        ntimes_device(A, xr, xi, yr, yi, n, work_r, work_i, as_stream(stream), noalloc_scratch(),
                      NOALLOC_SCRATCH_WORDS);
    });
}

extern "C" int dq_spmv_ntimes_host(const dq_devmat *A, const double *xr, const double *xi, double *yr, double *yi,
                                   int n) {
    return guard([&] {
        if (!A) throw ArgError("matrix is NULL");
        if (n < 0) throw ArgError("spmv_ntimes_cuda: num_times must be >= 0");
        const int64_t M = A->rows, N = A->cols;
        if (n == 0) {
            if (M != N)
                throw ArgError("spmv_ntimes_cuda: num_times==0 requires square matrix (M==N) to copy x -> y");
            std::copy(xr, xr + M, yr);
            std::copy(xi, xi + M, yi);
            return;
        }
        if (A->n() == 0) {
            std::fill(yr, yr + M, 0.0);
            std::fill(yi, yi + M, 0.0);
            return;
        }
        if (n > 1 && M != N)
            throw ArgError("spmv_ntimes_cuda: repeated SpMV requires square matrix (M==N) for num_times > 1");
        DevVec a(N, A->precision), b(M, A->precision);
        a.put(xr, xi);
        int64_t zwords = 0;
        DevBuf zmap;
        const uint32_t *zuse = sequence_map(*A, n);
        if (zuse) zwords = A->zmap_words;
        if (!zuse && ntimes_uses_map(*A, n)) {
            zmap = DevBuf(ntimes_map_bytes(*A, zwords));
            ntimes_zero_map(*A, zwords, zmap.as<uint32_t>(), 0);
            zuse = zmap.as<uint32_t>();
        }
        DevVec *in = &a, *out = &b;
        for (int t = 0; t < n; ++t) {
            spmv_banded(*A, in->r.p, in->i.p, out->r.p, out->i.p, 0, 0, zuse, zuse ? zwords : 0);
            if (t != n - 1) std::swap(in, out);
        }
        out->get(yr, yi);
    });
}
