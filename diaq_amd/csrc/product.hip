// DIA SpGEMM on gfx950: C = A B by diagonal-pair enumeration, with the structural dry-run that sizes C up front.
//
// Reference: product_kernel / product_device / product_cuda (src/spGEMM_gpu.cu:25-161), pow_cuda
// (src/spPow_gpu.cu:21-56).
//
// Dry-run (host, integer only): every pair (slot a, slot b) with dC = dA + dB in [-(rows-1), cols-1] contributes to
// result diagonal dC; result offsets ascend and each dC keeps its pairs in A-slot order (then B-slot order), so the
// sums run in the reference's order. The pairs are bucketed by dC with a counting sort (two passes over the slot
// pairs, in slot order, so each bucket keeps the enumeration order) when the offset range is at most four times the
// pairs; fewer pairs over a wider range take a stable sort.
//
// Kernel: in row space the pair (a, b) adds A_a[i] * B_b[i + dA] to C_c[i] (row i of A, row i + dA of B), where a
// diagonal's element for row r sits at rowbase + r; rows outside a diagonal's range contribute nothing, exactly the
// reference's position bounds. The work is cut into items of (row chunk of 256*RPT rows, result diagonal) --
// fixed-size items, so result diagonals of length 1 and of length N balance. Items are ordered by row chunk: the
// result diagonals meeting a chunk form a contiguous range of the ascending result offsets, so the host only sizes
// each chunk's range and a device kernel writes the item list. The item index is remapped so the workgroups of one
// row chunk, which read the same rows of A's diagonals, land on the same XCD (workgroup b runs on XCD b mod 8) and
// share its L2. A workgroup walks its diagonal's pairs in batches of 256: the lanes load the batch's pair metadata
// in parallel and leave each pair's row range and base pointers in LDS, then the pair loop reads them from there
// (no chain of dependent metadata loads per pair).
//
// Zero blocks: for large operands the product first maps which 256-row blocks of each diagonal of A and B hold a
// nonzero value (zero_blocks_kernel), and a pair whose A rows, or whose B rows, in the item's chunk are all zero is
// dropped when its metadata is staged. Its products are all +-0, and a sum that starts at +0 is left unchanged by
// adding +-0, so the result has the same bits; only a NaN or Inf facing a zero block is not propagated. A qubit
// Hamiltonian's +-2^k flip diagonals are zero on alternate runs of 2^k rows, so for k at least the chunk's log2 size
// half of a far pair's chunks read nothing -- and far pairs' B rows are the reads no cache catches.
#include <algorithm>
#include <atomic>
#include <memory>
#include <numeric>

#include "dia_rows.hpp"

namespace dq {

struct DryRun {
    std::vector<int32_t> res;        // ascending result offsets
    std::vector<int64_t> pair_start; // res.size() + 1
    std::vector<int32_t> sa, sb;     // pairs, grouped by result diagonal, A-slot order
};

static DryRun dryrun(uint32_t rows, uint32_t cols, const std::vector<int32_t> &oa, const std::vector<int32_t> &ob) {
    const int64_t lo = -static_cast<int64_t>(rows) + 1, hi = static_cast<int64_t>(cols) - 1;
    DryRun r;
    const int64_t range = hi - lo + 1;
    const int64_t n_ab = static_cast<int64_t>(oa.size()) * static_cast<int64_t>(ob.size());
    if (range <= 0 || n_ab == 0) {
        r.pair_start.push_back(0);
        return r;
    }
    // counting sort by dC when the offset range is at most a few times the pairs (a Taylor term's 8191 x 25 offsets,
    // range 8191); a few pairs over a wide range (TFIM-20 H.H: 1681 pairs, range 2^21) sort faster than they count
    if (range <= 4 * n_ab) {
        std::vector<int64_t> cnt(static_cast<size_t>(range), 0);
        for (size_t a = 0; a < oa.size(); ++a)
            for (size_t b = 0; b < ob.size(); ++b) {
                const int64_t dc = static_cast<int64_t>(oa[a]) + ob[b];
                if (dc >= lo && dc <= hi) ++cnt[static_cast<size_t>(dc - lo)];
            }
        int64_t total = 0;
        for (int64_t i = 0; i < range; ++i) {
            if (!cnt[i]) continue;
            r.res.push_back(static_cast<int32_t>(i + lo));
            r.pair_start.push_back(total);
            const int64_t c = cnt[i];
            cnt[i] = total;  // becomes the bucket's fill position
            total += c;
        }
        r.pair_start.push_back(total);
        r.sa.resize(static_cast<size_t>(total));
        r.sb.resize(static_cast<size_t>(total));
        for (size_t a = 0; a < oa.size(); ++a)
            for (size_t b = 0; b < ob.size(); ++b) {
                const int64_t dc = static_cast<int64_t>(oa[a]) + ob[b];
                if (dc < lo || dc > hi) continue;
                const int64_t at = cnt[static_cast<size_t>(dc - lo)]++;
                r.sa[at] = static_cast<int32_t>(a);
                r.sb[at] = static_cast<int32_t>(b);
            }
        return r;
    }
    struct P {
        int64_t dc;
        int32_t a, b;
    };
    std::vector<P> pairs;
    pairs.reserve(static_cast<size_t>(n_ab));
    for (size_t a = 0; a < oa.size(); ++a)
        for (size_t b = 0; b < ob.size(); ++b) {
            const int64_t dc = static_cast<int64_t>(oa[a]) + ob[b];
            if (dc < lo || dc > hi) continue;
            pairs.push_back(P{dc, static_cast<int32_t>(a), static_cast<int32_t>(b)});
        }
    // stable by construction order (a, then b): sort by dc only
    std::stable_sort(pairs.begin(), pairs.end(), [](const P &x, const P &y) { return x.dc < y.dc; });
    r.sa.reserve(pairs.size());
    r.sb.reserve(pairs.size());
    for (size_t p = 0; p < pairs.size(); ++p) {
        if (p == 0 || pairs[p].dc != pairs[p - 1].dc) {
            r.res.push_back(static_cast<int32_t>(pairs[p].dc));
            r.pair_start.push_back(static_cast<int64_t>(p));
        }
        r.sa.push_back(pairs[p].a);
        r.sb.push_back(pairs[p].b);
    }
    r.pair_start.push_back(static_cast<int64_t>(pairs.size()));
    return r;
}

// The dry-run depends only on the shape and the two offset lists. A Taylor series multiplies operands with the same
// offsets term after term once the terms reach every diagonal they can (TFIM-12: terms 8 to 14, 8191 x 25 offsets,
// ~1.2 ms of host time each), so each thread keeps its last dry-run and reuses it when the inputs are equal. Only
// dry-runs of at most DRYRUN_CACHE_PAIRS slot pairs are kept (their lists take ~16 B per pair), and
// dq_release_cached_memory drops the kept ones: the calling thread's at once, other threads' at their next product.
constexpr int64_t DRYRUN_CACHE_PAIRS = int64_t(1) << 22;
static std::atomic<uint64_t> g_dryrun_generation{0};

struct DryRunCache {
    uint64_t generation = 0;
    uint32_t rows = 0, cols = 0;
    std::vector<int32_t> oa, ob;
    std::shared_ptr<const DryRun> d;
};

static DryRunCache &dryrun_cache() {
    thread_local DryRunCache c;
    return c;
}

void release_dryrun_cache() {
    g_dryrun_generation.fetch_add(1);
    DryRunCache &c = dryrun_cache();
    c.d.reset();
    c.oa = std::vector<int32_t>();
    c.ob = std::vector<int32_t>();
}

static std::shared_ptr<const DryRun> dryrun_reused(uint32_t rows, uint32_t cols, const std::vector<int32_t> &oa,
                                                   const std::vector<int32_t> &ob) {
    DryRunCache &c = dryrun_cache();
    const uint64_t gen = g_dryrun_generation.load();
    if (c.d && c.generation == gen && c.rows == rows && c.cols == cols && c.oa == oa && c.ob == ob) return c.d;
    auto d = std::make_shared<const DryRun>(dryrun(rows, cols, oa, ob));
    const bool keep = static_cast<int64_t>(oa.size()) * static_cast<int64_t>(ob.size()) <= DRYRUN_CACHE_PAIRS;
    c.d = keep ? d : nullptr;
    c.generation = gen;
    c.rows = rows;
    c.cols = cols;
    c.oa = keep ? oa : std::vector<int32_t>();
    c.ob = keep ? ob : std::vector<int32_t>();
    return d;
}

constexpr int PROD_BATCH = 256;

// Process-wide knobs (dq_set_tuning): chunk order across XCDs and non-temporal stores of C.
static int g_prod_interleave = 1;
static int g_prod_nt_store = 1;
// rows per lane: row chunks of 256 * rpt rows; 0 = automatic: 4 below 2^16 rows, where more and shorter items ran a
// TFIM-12 expm's 4096-row Taylor products faster (DeviceMatrix expm_negative 31.1 -> 28.8 ms), else 8 (TFIM-20 H.H:
// 3.96 ms against 4.02 ms at 4 and 4.16 ms at 16)
static int g_prod_rpt = 0;
static int g_prod_skip = 1; // drop pairs whose rows in a chunk are all zero, for operands of >= 2^16 rows

// Item list: the i-th chunk in processing order (row chunk q_of[i]) owns items [first[i], first[i + 1]), result
// diagonals c_lo[i], c_lo[i] + 1, ...
__global__ void product_items_kernel(int64_t n_chunks, const int64_t *__restrict__ first,
                                     const int32_t *__restrict__ c_lo, const int32_t *__restrict__ q_of,
                                     int2 *__restrict__ items) {
    for (int64_t i = blockIdx.x; i < n_chunks; i += gridDim.x) {
        const int64_t f = first[i], n = first[i + 1] - f;
        const int32_t c0 = c_lo[i], q = q_of[i];
        for (int64_t j = threadIdx.x; j < n; j += blockDim.x) items[f + j] = make_int2(q, c0 + static_cast<int>(j));
    }
}

// Zero the storage outside every diagonal's rows: the alignment gaps between diagonals and the tail.
template <typename T>
__global__ void zero_gaps_kernel(int n, const DiagMeta *__restrict__ meta, int64_t total, T *__restrict__ re,
                                 T *__restrict__ im) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > n) return;
    const int64_t b = k == 0 ? 0 : meta[k - 1].rowbase + meta[k - 1].rhi;
    const int64_t e = k == n ? total : meta[k].rowbase + meta[k].rlo;
    for (int64_t i = b; i < e; ++i) {
        re[i] = T(0);
        im[i] = T(0);
    }
}

void zero_gaps(dq_devmat &M, hipStream_t s) {
    const int n = M.n();
    if (n == 0 || M.total == 0) return;
    const unsigned g = blocks_for(n + 1, 256, 1 << 20);
    if (M.precision == DQ_F64)
        hipLaunchKernelGGL(zero_gaps_kernel<double>, dim3(g), dim3(256), 0, s, n, M.d_meta.as<DiagMeta>(), M.total,
                           M.re.as<double>(), M.im.as<double>());
    else
        hipLaunchKernelGGL(zero_gaps_kernel<float>, dim3(g), dim3(256), 0, s, n, M.d_meta.as<DiagMeta>(), M.total,
                           M.re.as<float>(), M.im.as<float>());
    DQ_LAUNCH_CHECK();
}

__device__ inline int64_t uniform64(int64_t v) {
    const int lo = __builtin_amdgcn_readfirstlane(static_cast<int>(static_cast<uint64_t>(v) & 0xffffffffu));
    const int hi = __builtin_amdgcn_readfirstlane(static_cast<int>(static_cast<uint64_t>(v) >> 32));
    return static_cast<int64_t>((static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) | static_cast<uint32_t>(lo));
}

// Whether a 256-row block of the map row z (one diagonal; bit b of word w covers rows 256 (32 w + b) ..) that meets
// rows [lo, hi) holds a nonzero value.
__device__ inline bool blocks_live(const uint32_t *__restrict__ z, int64_t lo, int64_t hi) {
    const int64_t b0 = lo >> 8, b1 = (hi - 1) >> 8;
    for (int64_t w = b0 >> 5; w <= (b1 >> 5); ++w) {
        uint32_t m = z[w];
        if (w == (b0 >> 5)) m &= 0xffffffffu << (b0 & 31);
        if (w == (b1 >> 5)) m &= 0xffffffffu >> (31 - (b1 & 31));
        if (m) return true;
    }
    return false;
}

template <typename T, bool NTS, int PROD_RPT>
__global__ __launch_bounds__(256) void product_kernel(const int2 *__restrict__ items, const int64_t *__restrict__ xcd_first,
                                                      const int64_t *__restrict__ pair_start,
                                                      const int32_t *__restrict__ pa, const int32_t *__restrict__ pb,
                                                      const DiagMeta *__restrict__ ma, const DiagMeta *__restrict__ mb,
                                                      const DiagMeta *__restrict__ mc, const T *__restrict__ are,
                                                      const T *__restrict__ aim, const T *__restrict__ bre,
                                                      const T *__restrict__ bim, T *__restrict__ cre,
                                                      T *__restrict__ cim, double divisor,
                                                      const uint32_t *__restrict__ za, int64_t zwa,
                                                      const uint32_t *__restrict__ zb, int64_t zwb) {
    constexpr int PROD_CHUNK = 256 * PROD_RPT;
    __shared__ int64_t s_abase[PROD_BATCH], s_bbase[PROD_BATCH];
    __shared__ int32_t s_lo[PROD_BATCH], s_hi[PROD_BATCH];
    // XCD-aware remap: the workgroups of XCD x (x = b mod 8) take the item range [xcd_first[x], xcd_first[x + 1]).
    const int64_t b = blockIdx.x;
    const int64_t w = xcd_first[b & 7] + (b >> 3);
    if (w >= xcd_first[(b & 7) + 1]) return;  // uniform
    const int2 it = items[w];
    const DiagMeta m_c = mc[it.y];
    const int64_t q0 = static_cast<int64_t>(it.x) * PROD_CHUNK;
    const int64_t r0 = max(m_c.rlo, q0), r1 = min(m_c.rhi, q0 + PROD_CHUNK);
    const int64_t p0 = pair_start[it.y], p1 = pair_start[it.y + 1];
    T acc_r[PROD_RPT], acc_i[PROD_RPT];
#pragma unroll
    for (int j = 0; j < PROD_RPT; ++j) { acc_r[j] = T(0); acc_i[j] = T(0); }
    for (int64_t pbat = p0; pbat < p1; pbat += PROD_BATCH) {
        const int nb = static_cast<int>(min<int64_t>(PROD_BATCH, p1 - pbat));
        if (static_cast<int>(threadIdx.x) < nb) {
            const int64_t p = pbat + threadIdx.x;
            const int32_t sa = pa[p], sb = pb[p];
            const DiagMeta ta = ma[sa];
            const DiagMeta tb = mb[sb];
            // rows r of this item with r on A's diagonal and r + dA on B's: [lo, hi), as offsets from q0
            const int64_t lo = max(max(ta.rlo, tb.rlo - ta.off), r0);
            const int64_t hi = min(min(ta.rhi, tb.rhi - ta.off), r1);
            bool live = hi > lo;
            if (live && za) live = blocks_live(za + static_cast<int64_t>(sa) * zwa, lo, hi);
            if (live && zb) live = blocks_live(zb + static_cast<int64_t>(sb) * zwb, lo + ta.off, hi + ta.off);
            s_abase[threadIdx.x] = ta.rowbase;
            s_bbase[threadIdx.x] = tb.rowbase + ta.off;
            s_lo[threadIdx.x] = live ? static_cast<int32_t>(lo - q0) : 0;
            s_hi[threadIdx.x] = live ? static_cast<int32_t>(hi - q0) : 0;
        }
        __syncthreads();
        for (int k = 0; k < nb; ++k) {
            const int64_t lo = q0 + __builtin_amdgcn_readfirstlane(s_lo[k]);
            const int64_t hi = q0 + __builtin_amdgcn_readfirstlane(s_hi[k]);
            if (hi <= lo) continue;  // uniform
            const int64_t abase = uniform64(s_abase[k]), bbase = uniform64(s_bbase[k]);
            const T *apr = are + abase, *api = aim + abase;
            const T *bpr = bre + bbase, *bpi = bim + bbase;
            T va_r[PROD_RPT], va_i[PROD_RPT], vb_r[PROD_RPT], vb_i[PROD_RPT];
            if (lo == q0 && hi == q0 + PROD_CHUNK) {  // uniform: the whole chunk, no per-row checks
#pragma unroll
                for (int j = 0; j < PROD_RPT; ++j) {
                    const int64_t r = q0 + threadIdx.x + 256 * j;
                    va_r[j] = apr[r]; va_i[j] = api[r]; vb_r[j] = bpr[r]; vb_i[j] = bpi[r];
                }
            } else {  // clamp to a valid row, zero both operands of the rows outside [lo, hi)
#pragma unroll
                for (int j = 0; j < PROD_RPT; ++j) {
                    const int64_t r = q0 + threadIdx.x + 256 * j;
                    const int64_t rc = r < lo ? lo : (r >= hi ? hi - 1 : r);
                    va_r[j] = apr[rc]; va_i[j] = api[rc]; vb_r[j] = bpr[rc]; vb_i[j] = bpi[rc];
                }
#pragma unroll
                for (int j = 0; j < PROD_RPT; ++j) {
                    const int64_t r = q0 + threadIdx.x + 256 * j;
                    if (r < lo || r >= hi) { va_r[j] = T(0); va_i[j] = T(0); vb_r[j] = T(0); vb_i[j] = T(0); }
                }
            }
#pragma unroll
            for (int j = 0; j < PROD_RPT; ++j) {
                T re, im;
                cmul(va_r[j], va_i[j], vb_r[j], vb_i[j], re, im);
                acc_r[j] += re;
                acc_i[j] += im;
            }
        }
        __syncthreads();  // the next batch overwrites the staged pairs
    }
#pragma unroll
    for (int j = 0; j < PROD_RPT; ++j) {
        const int64_t r = q0 + threadIdx.x + 256 * j;
        if (r >= r0 && r < r1) {
            T vr = acc_r[j], vi = acc_i[j];
            if (divisor != 1.0) {
                vr = static_cast<T>(acc_r[j] / divisor);
                vi = static_cast<T>(acc_i[j] / divisor);
            }
            if constexpr (NTS) {  // C is written once and not read here: keep it out of the caches
                __builtin_nontemporal_store(vr, cre + m_c.rowbase + r);
                __builtin_nontemporal_store(vi, cim + m_c.rowbase + r);
            } else {
                cre[m_c.rowbase + r] = vr;
                cim[m_c.rowbase + r] = vi;
            }
        }
    }
}

std::unique_ptr<dq_devmat> product(const dq_devmat &A, const dq_devmat &B, double divisor, hipStream_t s,
                                   bool allow_dense) {
    if (A.cols != B.rows) throw ArgError("diaQMatrix product_cuda: dimension mismatch.");
    if (A.precision != B.precision) throw ArgError("product: operands have different precisions");
    if (allow_dense && dense_product_pays(A, B)) return product_dense(A, B, divisor, s);
    const std::shared_ptr<const DryRun> dp = dryrun_reused(A.rows, B.cols, A.offs, B.offs);
    const DryRun &d = *dp;
    // every stored element of C is written by the kernel; only the gaps between diagonals are zeroed
    auto C = make_devmat(A.rows, B.cols, d.res.data(), static_cast<int>(d.res.size()), A.precision, s, false);
    const int nc = C->n();
    if (nc == 0) return C;
    zero_gaps(*C, s);
    // per row chunk, the contiguous range of result diagonals that meet it (rlo and rhi do not increase with c).
    // Chunk order: the chunks q with q mod 8 == 0 first, then those with q mod 8 == 1, ... (interleave: the XCDs work
    // on neighbouring chunks at the same time, so a far pair's B rows, which are another chunk's near rows, are read
    // from the Infinity Cache), or plain ascending order.
    int64_t max_rhi = 0;
    for (int c = 0; c < nc; ++c) max_rhi = std::max(max_rhi, C->meta[c].rhi);
    const int prod_rpt = g_prod_rpt > 0 ? g_prod_rpt : (A.rows < (1u << 16) ? 4 : 8);
    const int64_t PROD_CHUNK = 256 * static_cast<int64_t>(prod_rpt);
    const int64_t n_chunks = (max_rhi + PROD_CHUNK - 1) / PROD_CHUNK;
    std::vector<int64_t> order;
    order.reserve(static_cast<size_t>(n_chunks));
    std::vector<int64_t> xcd_first(9, 0);
    for (int x = 0; x < 8; ++x) {
        if (g_prod_interleave)
            for (int64_t q = x; q < n_chunks; q += 8) order.push_back(q);
        else
            for (int64_t q = n_chunks * x / 8; q < n_chunks * (x + 1) / 8; ++q) order.push_back(q);
    }
    std::vector<int64_t> first(static_cast<size_t>(n_chunks) + 1, 0);
    std::vector<int32_t> c_lo(static_cast<size_t>(std::max<int64_t>(1, n_chunks)), 0), q_of(c_lo.size(), 0);
    for (int64_t i = 0; i < n_chunks; ++i) {
        const int64_t q = order[i];
        const int64_t Q0 = q * PROD_CHUNK, Q1 = Q0 + PROD_CHUNK;
        int lo = 0, hi = nc;  // first c with rlo < Q1
        while (lo < hi) {
            const int mid = (lo + hi) / 2;
            if (C->meta[mid].rlo >= Q1) lo = mid + 1; else hi = mid;
        }
        const int c1 = lo;
        lo = c1; hi = nc;     // first c (from c1) with rhi <= Q0
        while (lo < hi) {
            const int mid = (lo + hi) / 2;
            if (C->meta[mid].rhi > Q0) lo = mid + 1; else hi = mid;
        }
        const int c2 = lo;
        c_lo[i] = c1;
        q_of[i] = static_cast<int32_t>(q);
        first[i + 1] = first[i] + std::max(0, c2 - c1);
    }
    const int64_t n_work = first[n_chunks];
    if (n_work == 0) return C;
    // the item list (in chunk order) is split evenly over the XCDs: with many chunks each XCD works through whole
    // chunks in the order above; with fewer chunks than XCDs a chunk's items spread over several
    for (int x = 0; x <= 8; ++x) xcd_first[x] = n_work * x / 8;
    int64_t per_xcd = 0;
    for (int x = 0; x < 8; ++x) per_xcd = std::max(per_xcd, xcd_first[x + 1] - xcd_first[x]);
    const int64_t grid = per_xcd * 8;
    if (grid > 0x7fffffffLL) throw ArgError("product: too many work items");
    DevBuf dfirst(sizeof(int64_t) * first.size()), dclo(sizeof(int32_t) * c_lo.size()),
        dqof(sizeof(int32_t) * q_of.size()), dxcd(sizeof(int64_t) * xcd_first.size()),
        ditems(sizeof(int2) * static_cast<size_t>(n_work)), dps(sizeof(int64_t) * d.pair_start.size()),
        dsa(sizeof(int32_t) * std::max<size_t>(1, d.sa.size())), dsb(sizeof(int32_t) * std::max<size_t>(1, d.sb.size()));
    DQ_HIP(hipMemcpyAsync(dfirst.p, first.data(), dfirst.bytes, hipMemcpyHostToDevice, s));
    DQ_HIP(hipMemcpyAsync(dclo.p, c_lo.data(), dclo.bytes, hipMemcpyHostToDevice, s));
    DQ_HIP(hipMemcpyAsync(dqof.p, q_of.data(), dqof.bytes, hipMemcpyHostToDevice, s));
    DQ_HIP(hipMemcpyAsync(dxcd.p, xcd_first.data(), dxcd.bytes, hipMemcpyHostToDevice, s));
    DQ_HIP(hipMemcpyAsync(dps.p, d.pair_start.data(), dps.bytes, hipMemcpyHostToDevice, s));
    DQ_HIP(hipMemcpyAsync(dsa.p, d.sa.data(), sizeof(int32_t) * d.sa.size(), hipMemcpyHostToDevice, s));
    DQ_HIP(hipMemcpyAsync(dsb.p, d.sb.data(), sizeof(int32_t) * d.sb.size(), hipMemcpyHostToDevice, s));
    // zero-block maps of A and B (a squaring maps its operand once; a map the handle keeps is used as it is)
    DevBuf zab, zbb;
    const uint32_t *za = nullptr, *zb = nullptr;
    const int64_t zwa = zero_block_words(A.rows, 256), zwb = zero_block_words(B.rows, 256);
    if (g_prod_skip && A.rows >= (int64_t(1) << 16) && A.n() > 0 && B.n() > 0) {
        auto map_of = [&](const dq_devmat &M, int64_t zw, DevBuf &buf) -> const uint32_t * {
            if (M.zmap_valid && M.zmap.p && M.zmap_words == zw) return M.zmap.as<uint32_t>();
            buf = DevBuf(sizeof(uint32_t) * static_cast<size_t>(zw) * static_cast<size_t>(M.n()));
            const dim3 g(static_cast<unsigned>(zw), zero_map_grid_y(M.n()), 1);
            if (M.precision == DQ_F64)
                hipLaunchKernelGGL((zero_blocks_kernel<double, 256>), g, dim3(256), 0, s, static_cast<int64_t>(M.rows),
                                   M.n(), M.d_meta.as<DiagMeta>(), int64_t(0), M.re.as<double>(), M.im.as<double>(),
                                   zw, buf.as<uint32_t>());
            else
                hipLaunchKernelGGL((zero_blocks_kernel<float, 256>), g, dim3(256), 0, s, static_cast<int64_t>(M.rows),
                                   M.n(), M.d_meta.as<DiagMeta>(), int64_t(0), M.re.as<float>(), M.im.as<float>(), zw,
                                   buf.as<uint32_t>());
            DQ_LAUNCH_CHECK();
            return buf.as<uint32_t>();
        };
        za = map_of(A, zwa, zab);
        zb = (&A == &B) ? za : map_of(B, zwb, zbb);
    }
    hipLaunchKernelGGL(product_items_kernel, dim3(static_cast<unsigned>(std::min<int64_t>(n_chunks, 65536))),
                       dim3(256), 0, s, n_chunks, dfirst.as<int64_t>(), dclo.as<int32_t>(), dqof.as<int32_t>(),
                       ditems.as<int2>());
    DQ_LAUNCH_CHECK();
#define DQ_PROD_LAUNCH_R(T, NTS, R)                                                                                \
    hipLaunchKernelGGL((product_kernel<T, NTS, R>), dim3(static_cast<unsigned>(grid)), dim3(256), 0, s,           \
                       ditems.as<int2>(), dxcd.as<int64_t>(), dps.as<int64_t>(), dsa.as<int32_t>(),               \
                       dsb.as<int32_t>(), A.d_meta.as<DiagMeta>(), B.d_meta.as<DiagMeta>(), C->d_meta.as<DiagMeta>(), \
                       A.re.as<T>(), A.im.as<T>(), B.re.as<T>(), B.im.as<T>(), C->re.as<T>(), C->im.as<T>(), divisor, \
                       za, zwa, zb, zwb)
#define DQ_PROD_LAUNCH(T, NTS)                                \
    do {                                                      \
        if (prod_rpt == 4) { DQ_PROD_LAUNCH_R(T, NTS, 4); }   \
        else if (prod_rpt == 16) { DQ_PROD_LAUNCH_R(T, NTS, 16); } \
        else { DQ_PROD_LAUNCH_R(T, NTS, 8); }                 \
    } while (0)
    if (A.precision == DQ_F64) {
        if (g_prod_nt_store) DQ_PROD_LAUNCH(double, true);
        else DQ_PROD_LAUNCH(double, false);
    } else {
        if (g_prod_nt_store) DQ_PROD_LAUNCH(float, true);
        else DQ_PROD_LAUNCH(float, false);
    }
#undef DQ_PROD_LAUNCH
#undef DQ_PROD_LAUNCH_R
    DQ_LAUNCH_CHECK();
    DQ_HIP(hipStreamSynchronize(s));  // before the work list is freed
    return C;
}

// A^n as pow_cuda (src/spPow_gpu.cu:21-56): n = 0 gives I, n = 1 a copy, 2^s by s squarings, otherwise binary
// exponentiation starting from the first odd power.
std::unique_ptr<dq_devmat> matrix_pow(const dq_devmat &A, int64_t n, hipStream_t s) {
    if (n < 0) throw ArgError("pow_cuda: negative exponent not supported");
    if (n == 0) {
        const int32_t zero = 0;
        auto I = make_devmat(A.rows, A.rows, &zero, 1, A.precision, s);
        fill_identity(*I, s);
        return I;
    }
    if (n == 1) return copy_devmat(A, s);
    uint64_t e = static_cast<uint64_t>(n);
    std::unique_ptr<dq_devmat> base = copy_devmat(A, s);
    if ((e & (e - 1)) == 0) {
        while (e > 1) {
            base = product(*base, *base, 1.0, s, true);
            e >>= 1;
        }
        return base;
    }
    std::unique_ptr<dq_devmat> result;
    while (e > 0) {
        if (e & 1ULL) result = result ? product(*result, *base, 1.0, s, true) : copy_devmat(*base, s);
        e >>= 1ULL;
        if (e) base = product(*base, *base, 1.0, s, true);
    }
    return result;
}

}  // namespace dq

using namespace dq;

extern "C" int dq_product_dryrun(uint32_t rows, uint32_t cols, int32_t n_a, const int32_t *offs_a, int32_t n_b,
                                 const int32_t *offs_b, int32_t *n_res, int64_t *n_pairs, int32_t *res_offsets,
                                 int64_t *res_pair_start, int32_t *pair_slot_a, int32_t *pair_slot_b) {
    return guard([&] {
        if (n_a < 0 || n_b < 0) throw ArgError("negative diagonal count");
        if ((n_a && !offs_a) || (n_b && !offs_b)) throw ArgError("NULL offsets");
        const DryRun d = dryrun(rows, cols, std::vector<int32_t>(offs_a, offs_a + n_a),
                                std::vector<int32_t>(offs_b, offs_b + n_b));
        if (n_res) *n_res = static_cast<int32_t>(d.res.size());
        if (n_pairs) *n_pairs = static_cast<int64_t>(d.sa.size());
        if (res_offsets) {
            std::copy(d.res.begin(), d.res.end(), res_offsets);
            if (res_pair_start) std::copy(d.pair_start.begin(), d.pair_start.end(), res_pair_start);
            if (pair_slot_a) std::copy(d.sa.begin(), d.sa.end(), pair_slot_a);
            if (pair_slot_b) std::copy(d.sb.begin(), d.sb.end(), pair_slot_b);
        }
    });
}

namespace dq {
// Called by dq_set_tuning for the "product_*" keys.
bool product_set_tuning(const std::string &key, int value) {
    if (key == "product_interleave") {
        if (value < 0 || value > 1) throw ArgError("product_interleave must be 0 or 1");
        g_prod_interleave = value;
        return true;
    }
    if (key == "product_rows_per_lane") {
        if (value != 0 && value != 4 && value != 8 && value != 16)
            throw ArgError("product_rows_per_lane must be 0 (automatic), 4, 8 or 16");
        g_prod_rpt = value;
        return true;
    }
    if (key == "product_skip_zero_blocks") {
        if (value < 0 || value > 1) throw ArgError("product_skip_zero_blocks must be 0 or 1");
        g_prod_skip = value;
        return true;
    }
    if (key == "product_nt_store") {
        if (value < 0 || value > 1) throw ArgError("product_nt_store must be 0 or 1");
        g_prod_nt_store = value;
        return true;
    }
    return false;
}
}  // namespace dq

extern "C" int dq_product_device(const dq_devmat *A, const dq_devmat *B, dq_devmat **out) {
    return guard([&] {
        if (!A || !B || !out) throw ArgError("NULL argument");
        *out = product(*A, *B, 1.0, 0).release();
    });
}

extern "C" int dq_pow(const dq_devmat *A, int64_t n, dq_devmat **out) {
    return guard([&] {
        if (!A || !out) throw ArgError("NULL argument");
        *out = matrix_pow(*A, n, 0).release();
        DQ_HIP(hipStreamSynchronize(0));
    });
}
