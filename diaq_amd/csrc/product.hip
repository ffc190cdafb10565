// DIA SpGEMM on gfx950: C = A B by diagonal-pair enumeration, with the structural dry-run that sizes C up front.
//
// Reference: product_kernel / product_device / product_cuda (src/spGEMM_gpu.cu:25-161), pow_cuda
// (src/spPow_gpu.cu:21-56).
//
// Dry-run (host, integer only): every pair (slot a, slot b) with dC = dA + dB in [-(rows-1), cols-1] contributes to
// result diagonal dC; result offsets ascend and each dC keeps its pairs in A-slot order (then B-slot order), so the
// sums run in the reference's order.
//
// Kernel: in row space the pair (a, b) adds A_a[i] * B_b[i + dA] to C_c[i] (row i of A, row i + dA of B), where a
// diagonal's element for row r sits at rowbase + r; rows outside a diagonal's range contribute nothing, exactly the
// reference's position bounds. The work is cut into items of (result diagonal, chunk of 256*RPT rows) -- fixed-size
// items, so result diagonals of length 1 and of length N balance -- and items are ordered by row chunk so the
// workgroups of one row chunk, which read the same rows of A's diagonals, run together; the item index is remapped
// so those workgroups land on the same XCD (workgroup b runs on XCD b mod 8) and share its L2.
#include <algorithm>
#include <numeric>

#include "dq_common.hpp"

namespace dq {

struct DryRun {
    std::vector<int32_t> res;        // ascending result offsets
    std::vector<int64_t> pair_start; // res.size() + 1
    std::vector<int32_t> sa, sb;     // pairs, grouped by result diagonal, A-slot order
};

static DryRun dryrun(uint32_t rows, uint32_t cols, const std::vector<int32_t> &oa, const std::vector<int32_t> &ob) {
    const int64_t lo = -static_cast<int64_t>(rows) + 1, hi = static_cast<int64_t>(cols) - 1;
    struct P {
        int64_t dc;
        int32_t a, b;
    };
    std::vector<P> pairs;
    pairs.reserve(oa.size() * ob.size());
    for (size_t a = 0; a < oa.size(); ++a)
        for (size_t b = 0; b < ob.size(); ++b) {
            const int64_t dc = static_cast<int64_t>(oa[a]) + ob[b];
            if (dc < lo || dc > hi) continue;
            pairs.push_back(P{dc, static_cast<int32_t>(a), static_cast<int32_t>(b)});
        }
    // stable by construction order (a, then b): sort by dc only
    std::stable_sort(pairs.begin(), pairs.end(), [](const P &x, const P &y) { return x.dc < y.dc; });
    DryRun r;
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

struct WorkItem {
    int64_t r0;     // first row
    int32_t c;      // result slot
    int32_t nrows;  // rows in this item
};

constexpr int PROD_RPT = 4;
constexpr int PROD_CHUNK = 256 * PROD_RPT;

template <typename T>
__global__ __launch_bounds__(256) void product_kernel(const WorkItem *__restrict__ work, int64_t n_work,
                                                      int64_t per_xcd, const int64_t *__restrict__ pair_start,
                                                      const int32_t *__restrict__ pa, const int32_t *__restrict__ pb,
                                                      const DiagMeta *__restrict__ ma, const DiagMeta *__restrict__ mb,
                                                      const DiagMeta *__restrict__ mc, const T *__restrict__ are,
                                                      const T *__restrict__ aim, const T *__restrict__ bre,
                                                      const T *__restrict__ bim, T *__restrict__ cre,
                                                      T *__restrict__ cim, double divisor) {
    // XCD-aware remap: the workgroups of XCD x (x = b mod 8) take the contiguous item range [x*per_xcd, ...).
    const int64_t b = blockIdx.x;
    const int64_t w = (b & 7) * per_xcd + (b >> 3);
    if (w >= n_work) return;
    const WorkItem it = work[w];
    const DiagMeta m_c = mc[it.c];
    const int64_t p0 = pair_start[it.c], p1 = pair_start[it.c + 1];
    T acc_r[PROD_RPT], acc_i[PROD_RPT];
#pragma unroll
    for (int j = 0; j < PROD_RPT; ++j) { acc_r[j] = T(0); acc_i[j] = T(0); }
    const int64_t r_end = it.r0 + it.nrows;
    for (int64_t p = p0; p < p1; ++p) {
        const DiagMeta ta = ma[pa[p]];
        const DiagMeta tb = mb[pb[p]];
        // rows r of this item with r on A's diagonal and r + dA on B's: [lo, hi)
        const int64_t lo = max(max(ta.rlo, tb.rlo - ta.off), it.r0);
        const int64_t hi = min(min(ta.rhi, tb.rhi - ta.off), r_end);
        if (hi <= lo) continue;  // uniform
        const T *apr = are + ta.rowbase, *api = aim + ta.rowbase;
        const T *bpr = bre + tb.rowbase + ta.off, *bpi = bim + tb.rowbase + ta.off;
        T va_r[PROD_RPT], va_i[PROD_RPT], vb_r[PROD_RPT], vb_i[PROD_RPT];
        if (lo == it.r0 && hi == it.r0 + PROD_CHUNK) {  // uniform: the whole item, no per-row checks
#pragma unroll
            for (int j = 0; j < PROD_RPT; ++j) {
                const int64_t r = it.r0 + threadIdx.x + 256 * j;
                va_r[j] = apr[r]; va_i[j] = api[r]; vb_r[j] = bpr[r]; vb_i[j] = bpi[r];
            }
        } else {  // clamp to a valid row, zero both operands of the rows outside [lo, hi)
#pragma unroll
            for (int j = 0; j < PROD_RPT; ++j) {
                const int64_t r = it.r0 + threadIdx.x + 256 * j;
                const int64_t rc = r < lo ? lo : (r >= hi ? hi - 1 : r);
                va_r[j] = apr[rc]; va_i[j] = api[rc]; vb_r[j] = bpr[rc]; vb_i[j] = bpi[rc];
            }
#pragma unroll
            for (int j = 0; j < PROD_RPT; ++j) {
                const int64_t r = it.r0 + threadIdx.x + 256 * j;
                if (r < lo || r >= hi) { va_r[j] = T(0); va_i[j] = T(0); vb_r[j] = T(0); vb_i[j] = T(0); }
            }
        }
#pragma unroll
        for (int j = 0; j < PROD_RPT; ++j) {
            acc_r[j] += va_r[j] * vb_r[j] - va_i[j] * vb_i[j];
            acc_i[j] += va_r[j] * vb_i[j] + va_i[j] * vb_r[j];
        }
    }
#pragma unroll
    for (int j = 0; j < PROD_RPT; ++j) {
        const int64_t r = it.r0 + threadIdx.x + 256 * j;
        if (r < it.r0 + it.nrows) {
            if (divisor != 1.0) {
                cre[m_c.rowbase + r] = static_cast<T>(acc_r[j] / divisor);
                cim[m_c.rowbase + r] = static_cast<T>(acc_i[j] / divisor);
            } else {
                cre[m_c.rowbase + r] = acc_r[j];
                cim[m_c.rowbase + r] = acc_i[j];
            }
        }
    }
}

std::unique_ptr<dq_devmat> product(const dq_devmat &A, const dq_devmat &B, double divisor, hipStream_t s) {
    if (A.cols != B.rows) throw ArgError("diaQMatrix product_cuda: dimension mismatch.");
    if (A.precision != B.precision) throw ArgError("product: operands have different precisions");
    const DryRun d = dryrun(A.rows, B.cols, A.offs, B.offs);
    auto C = make_devmat(A.rows, B.cols, d.res.data(), static_cast<int>(d.res.size()), A.precision, s);
    if (d.res.empty()) return C;
    // work items, row-chunk major (chunks aligned to global multiples of PROD_CHUNK)
    std::vector<WorkItem> work;
    int64_t max_q = 0;
    for (int c = 0; c < C->n(); ++c) max_q = std::max(max_q, (C->meta[c].rhi + PROD_CHUNK - 1) / PROD_CHUNK);
    std::vector<std::vector<WorkItem>> by_q(static_cast<size_t>(max_q));
    for (int c = 0; c < C->n(); ++c) {
        const DiagMeta &m = C->meta[c];
        if (m.rhi <= m.rlo) continue;
        for (int64_t q = m.rlo / PROD_CHUNK; q * PROD_CHUNK < m.rhi; ++q) {
            const int64_t r0 = std::max(m.rlo, q * PROD_CHUNK);
            const int64_t r1 = std::min(m.rhi, (q + 1) * PROD_CHUNK);
            by_q[q].push_back(WorkItem{r0, c, static_cast<int32_t>(r1 - r0)});
        }
    }
    for (auto &v : by_q) work.insert(work.end(), v.begin(), v.end());
    if (work.empty()) return C;
    const int64_t n_work = static_cast<int64_t>(work.size());
    const int64_t per_xcd = (n_work + 7) / 8;
    const int64_t grid = per_xcd * 8;
    DevBuf dw(sizeof(WorkItem) * work.size()), dps(sizeof(int64_t) * d.pair_start.size()),
        dsa(sizeof(int32_t) * std::max<size_t>(1, d.sa.size())), dsb(sizeof(int32_t) * std::max<size_t>(1, d.sb.size()));
    DQ_HIP(hipMemcpyAsync(dw.p, work.data(), dw.bytes, hipMemcpyHostToDevice, s));
    DQ_HIP(hipMemcpyAsync(dps.p, d.pair_start.data(), dps.bytes, hipMemcpyHostToDevice, s));
    DQ_HIP(hipMemcpyAsync(dsa.p, d.sa.data(), sizeof(int32_t) * d.sa.size(), hipMemcpyHostToDevice, s));
    DQ_HIP(hipMemcpyAsync(dsb.p, d.sb.data(), sizeof(int32_t) * d.sb.size(), hipMemcpyHostToDevice, s));
    if (grid > 0x7fffffffLL) throw ArgError("product: too many work items");
    if (A.precision == DQ_F64)
        hipLaunchKernelGGL(product_kernel<double>, dim3(static_cast<unsigned>(grid)), dim3(256), 0, s,
                           dw.as<WorkItem>(), n_work, per_xcd, dps.as<int64_t>(), dsa.as<int32_t>(),
                           dsb.as<int32_t>(), A.d_meta.as<DiagMeta>(), B.d_meta.as<DiagMeta>(),
                           C->d_meta.as<DiagMeta>(), A.re.as<double>(), A.im.as<double>(), B.re.as<double>(),
                           B.im.as<double>(), C->re.as<double>(), C->im.as<double>(), divisor);
    else
        hipLaunchKernelGGL(product_kernel<float>, dim3(static_cast<unsigned>(grid)), dim3(256), 0, s,
                           dw.as<WorkItem>(), n_work, per_xcd, dps.as<int64_t>(), dsa.as<int32_t>(),
                           dsb.as<int32_t>(), A.d_meta.as<DiagMeta>(), B.d_meta.as<DiagMeta>(),
                           C->d_meta.as<DiagMeta>(), A.re.as<float>(), A.im.as<float>(), B.re.as<float>(),
                           B.im.as<float>(), C->re.as<float>(), C->im.as<float>(), divisor);
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
            base = product(*base, *base, 1.0, s);
            e >>= 1;
        }
        return base;
    }
    std::unique_ptr<dq_devmat> result;
    while (e > 0) {
        if (e & 1ULL) result = result ? product(*result, *base, 1.0, s) : copy_devmat(*base, s);
        e >>= 1ULL;
        if (e) base = product(*base, *base, 1.0, s);
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
