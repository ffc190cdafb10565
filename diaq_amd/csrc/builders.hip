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
// Each element is written once. The host lists, for every diagonal, the terms that can land on it (a term with w
// flipped qubits lands on one of its 2^w diagonals in each row, chosen by the row's bits at those qubits), in term
// order.
// A lane owns rows and, diagonal by diagonal, adds the listed terms whose entry in its row falls on that diagonal,
// then stores the sum: the stores of a wave are coalesced, every stored element of every diagonal is written exactly
// once (so the layout needs no zero fill beforehand), and each element sums its terms in term order.
#include <algorithm>
#include <vector>

#include "dq_common.hpp"

namespace {
typedef double d2v __attribute__((ext_vector_type(2)));  // one 16-byte access for a complex value
}

namespace dq {

constexpr int PAULI_RPT = 4;  // rows per lane

__global__ __launch_bounds__(256) void pauli_diag_kernel(int64_t N, int n_diags, const DiagMeta *__restrict__ meta,
                                                         const int32_t *__restrict__ list_start,
                                                         const int32_t *__restrict__ list_term,
                                                         const uint64_t *__restrict__ xm,
                                                         const uint64_t *__restrict__ zm,
                                                         const double *__restrict__ cre,
                                                         const double *__restrict__ cim, double *__restrict__ re,
                                                         double *__restrict__ im) {
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * (256 * PAULI_RPT) + threadIdx.x;
    for (int k = blockIdx.y; k < n_diags; k += gridDim.y) {  // uniform
        const DiagMeta md = meta[k];
        const int t0 = list_start[k], t1 = list_start[k + 1];
        double ar[PAULI_RPT], ai[PAULI_RPT];
#pragma unroll
        for (int q = 0; q < PAULI_RPT; ++q) { ar[q] = 0.0; ai[q] = 0.0; }
        for (int t = t0; t < t1; ++t) {  // uniform
            const int j = list_term[t];
            const uint64_t x = xm[j], z = zm[j];
            const int ny = __popcll(x & z) & 3;  // i^ny from the term's Y factors
            const double pr0 = ny == 0 ? 1.0 : (ny == 2 ? -1.0 : 0.0);
            const double pi0 = ny == 1 ? 1.0 : (ny == 3 ? -1.0 : 0.0);
            const double vr = cre[j] * pr0 - cim[j] * pi0, vi = cre[j] * pi0 + cim[j] * pr0;
#pragma unroll
            for (int q = 0; q < PAULI_RPT; ++q) {
                const int64_t r = r0 + 256 * q;
                const int64_t c = static_cast<int64_t>(static_cast<uint64_t>(r) ^ x);
                if (r < N && c - r == md.off) {
                    const bool neg = __popcll(static_cast<uint64_t>(c) & z) & 1;
                    ar[q] += neg ? -vr : vr;
                    ai[q] += neg ? -vi : vi;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < PAULI_RPT; ++q) {
            const int64_t r = r0 + 256 * q;
            if (r >= md.rlo && r < md.rhi) {
                re[md.rowbase + r] = ar[q];
                im[md.rowbase + r] = ai[q];
            }
        }
    }
}

// The diagonals of term j: every +-2^k combination over its flipped qubits (ascending order not implied).
static void term_offsets(int n_qubits, uint64_t x, std::vector<int64_t> &out) {
    std::vector<int> bits;
    for (int k = 0; k < n_qubits; ++k)
        if ((x >> k) & 1) bits.push_back(k);
    if (bits.size() > 20) throw ArgError("pauli_sum_device: a term flips more than 20 qubits (2^20+ diagonals)");
    const uint64_t combos = uint64_t(1) << bits.size();
    for (uint64_t s = 0; s < combos; ++s) {
        int64_t d = 0;
        for (size_t b = 0; b < bits.size(); ++b) d += ((s >> b) & 1) ? -(int64_t(1) << bits[b]) : (int64_t(1) << bits[b]);
        out.push_back(d);
    }
}

// The structural diagonal set of the sum: every +-2^k combination over each term's flipped qubits.
static std::vector<int32_t> pauli_offsets(int n_qubits, int n_terms, const uint64_t *xm) {
    std::vector<int64_t> all;
    for (int j = 0; j < n_terms; ++j) {
        term_offsets(n_qubits, xm[j], all);
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

// Per diagonal slot, the terms that can land on it, in term order: list_start (n_diags + 1) and list_term.
static void pauli_lists(int n_qubits, int n_terms, const uint64_t *xm, const std::vector<int32_t> &offs,
                        std::vector<int32_t> &list_start, std::vector<int32_t> &list_term) {
    std::vector<std::vector<int32_t>> slots_of(static_cast<size_t>(n_terms));
    std::vector<int64_t> cnt(offs.size() + 1, 0);
    std::vector<int64_t> ds;
    for (int j = 0; j < n_terms; ++j) {
        ds.clear();
        term_offsets(n_qubits, xm[j], ds);
        for (int64_t d : ds) {
            const auto it = std::lower_bound(offs.begin(), offs.end(), static_cast<int32_t>(d));
            const int32_t k = static_cast<int32_t>(it - offs.begin());
            slots_of[j].push_back(k);
            ++cnt[static_cast<size_t>(k) + 1];
        }
    }
    for (size_t k = 0; k < offs.size(); ++k) cnt[k + 1] += cnt[k];
    if (cnt.back() > 0x7fffffff) throw ArgError("pauli_sum_device: too many (term, diagonal) pairs");
    list_start.assign(cnt.begin(), cnt.end());
    list_term.assign(static_cast<size_t>(cnt.back()), 0);
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    for (int j = 0; j < n_terms; ++j)  // terms in order, so each slot's list ascends
        for (int32_t k : slots_of[j]) list_term[static_cast<size_t>(fill[k]++)] = j;
}

// ------------------------------------------------------------------------------------------------------------------
// Dense <-> DIA on the device (src/dense_to_diagonals.hpp:20-58 and src/conversions.cpp:70-101 do it on the host).
//
// The dense matrix is row-major complex128 (re, im interleaved, row stride ld complex elements), the layout of a
// contiguous torch.complex128 or numpy complex array. Both directions walk it in tiles of CT_R x CT_C elements read
// or written row by row (coalesced 16-byte accesses). A tile meets CT_R + CT_C - 1 diagonals; the diagonal side of
// each tile goes through LDS, so the DIA side is also read or written in runs of consecutive rows.
// ------------------------------------------------------------------------------------------------------------------
constexpr int CT_R = 32, CT_C = 64, CT_D = CT_R + CT_C - 1;

__device__ inline int64_t tile_row0(int64_t t, int64_t tiles_c) { return (t / tiles_c) * CT_R; }
__device__ inline int64_t tile_col0(int64_t t, int64_t tiles_c) { return (t % tiles_c) * CT_C; }

// Pass 1 of dense -> DIA: flag[d + rows - 1] = 1 for every diagonal d holding an element with a nonzero real or
// imaginary part (-0.0 counts as zero, NaN as nonzero: the comparison v != 0.0 of the reference). The tile's flags
// are gathered in LDS first, so at most CT_D bytes go out per tile. Every write stores 1: the races between tiles
// are benign.
__global__ __launch_bounds__(256) void dense_nonzero_diags_kernel(int64_t rows, int64_t cols,
                                                                  const double *__restrict__ D, int64_t ld,
                                                                  int64_t tiles_c, int64_t ntiles,
                                                                  unsigned char *__restrict__ flags) {
    __shared__ int f[CT_D];
    const int j = threadIdx.x & (CT_C - 1);
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int64_t r0 = tile_row0(t, tiles_c), c0 = tile_col0(t, tiles_c);
        for (int k = threadIdx.x; k < CT_D; k += blockDim.x) f[k] = 0;
        __syncthreads();
        const int64_t c = c0 + j;
#pragma unroll
        for (int i = threadIdx.x / CT_C; i < CT_R; i += 256 / CT_C) {
            const int64_t r = r0 + i;
            if (r < rows && c < cols) {
                const d2v v = __builtin_nontemporal_load(reinterpret_cast<const d2v *>(D + 2 * (r * ld + c)));
                if (v.x != 0.0 || v.y != 0.0) f[j - i + CT_R - 1] = 1;
            }
        }
        __syncthreads();
        for (int k = threadIdx.x; k < CT_D; k += blockDim.x)
            if (f[k]) flags[c0 - r0 + k - (CT_R - 1) + rows - 1] = 1;
        __syncthreads();
    }
}

// The tiled kernels of the dense side (diagonal sets that fill a good part of the matrix) take the tile shape R x C
// as template arguments; C divides 256.
template <int R, int C>
struct Tile {
    static constexpr int D = R + C - 1;  // diagonals met by a tile
    __device__ static int64_t row0(int64_t t, int64_t tiles_c) { return (t / tiles_c) * R; }
    __device__ static int64_t col0(int64_t t, int64_t tiles_c) { return (t % tiles_c) * C; }
    // Whether the tile meets a stored diagonal: bit g of `bits` is set when diagonal g - (rows - 1) is stored, and the
    // tile meets g in [c0 - r0 - (R - 1), c0 - r0 + C - 1] + rows - 1.
    __device__ static bool meets_stored(int64_t r0, int64_t c0, int64_t rows, int64_t cols,
                                        const uint32_t *__restrict__ bits) {
        int64_t lo = c0 - r0 - (R - 1) + rows - 1, hi = lo + D - 1;
        if (lo < 0) lo = 0;
        if (hi > rows + cols - 2) hi = rows + cols - 2;
        for (int64_t w = lo >> 5; w <= (hi >> 5); ++w) {
            uint32_t m = bits[w];
            if (w == (lo >> 5)) m &= ~0u << (lo & 31);
            if (w == (hi >> 5)) m &= ~0u >> (31 - (hi & 31));
            if (m) return true;
        }
        return false;
    }
    // The rowbase of every diagonal that the tile meets (NOT_STORED if it is not stored), into LDS; rowbase_of has
    // rows + cols - 1 entries, rowbase_of[d + rows - 1].
    __device__ static void stage_bases(int64_t r0, int64_t c0, int64_t rows, int64_t cols,
                                       const int64_t *__restrict__ rowbase_of, int64_t *base);
};

// A block of the tile kernels owns the tiles blockIdx.x + q * gridDim.x, q < CT_TPB. It tests all of them against
// the bitmap at once (one lane each), so the tiles that meet no stored diagonal cost no dependent loads later.
constexpr int CT_TPB = 32;
constexpr int64_t NOT_STORED = INT64_MIN;

template <int R, int C>
__device__ void Tile<R, C>::stage_bases(int64_t r0, int64_t c0, int64_t rows, int64_t cols,
                                        const int64_t *__restrict__ rowbase_of, int64_t *base) {
    for (int k = threadIdx.x; k < D; k += blockDim.x) {
        const int64_t g = c0 - r0 + k - (R - 1) + rows - 1;
        base[k] = (g >= 0 && g < rows + cols - 1) ? rowbase_of[g] : NOT_STORED;
    }
    __syncthreads();
}

template <int R, int C>
__device__ inline void classify_block_tiles(int64_t rows, int64_t cols, int64_t tiles_c, int64_t ntiles,
                                            const uint32_t *__restrict__ bits, unsigned char *has) {
    if (threadIdx.x < CT_TPB) {
        const int64_t t = blockIdx.x + static_cast<int64_t>(threadIdx.x) * gridDim.x;
        has[threadIdx.x] = t < ntiles && Tile<R, C>::meets_stored(Tile<R, C>::row0(t, tiles_c),
                                                                   Tile<R, C>::col0(t, tiles_c), rows, cols, bits);
    }
    __syncthreads();
}

// Pass 2 of dense -> DIA: each tile that meets a stored diagonal loads the elements on stored diagonals and writes
// them out diagonal by diagonal: R consecutive lanes cover R consecutive rows of one diagonal.
template <typename TD, int R, int C>
__global__ __launch_bounds__(256) void dense_to_dia_tiles_kernel(int64_t rows, int64_t cols,
                                                                 const double *__restrict__ D, int64_t ld,
                                                                 int64_t tiles_c, int64_t ntiles,
                                                                 const int64_t *__restrict__ rowbase_of,
                                                                 const uint32_t *__restrict__ bits,
                                                                 TD *__restrict__ re, TD *__restrict__ im) {
    using TL = Tile<R, C>;
    __shared__ double tre[R][C], tim[R][C];
    __shared__ int64_t base[TL::D];
    __shared__ unsigned char has[CT_TPB];
    classify_block_tiles<R, C>(rows, cols, tiles_c, ntiles, bits, has);
    const int j = threadIdx.x % C;
    for (int q = 0; q < CT_TPB; ++q) {
        const int64_t t = blockIdx.x + static_cast<int64_t>(q) * gridDim.x;
        if (t >= ntiles) break;
        if (!has[q]) continue;
        const int64_t r0 = TL::row0(t, tiles_c), c0 = TL::col0(t, tiles_c);
        TL::stage_bases(r0, c0, rows, cols, rowbase_of, base);
        const int64_t c = c0 + j;
#pragma unroll
        for (int i = threadIdx.x / C; i < R; i += 256 / C) {
            const int64_t r = r0 + i;
            if (r < rows && c < cols && base[j - i + R - 1] != NOT_STORED) {
                const d2v v = __builtin_nontemporal_load(reinterpret_cast<const d2v *>(D + 2 * (r * ld + c)));
                tre[i][j] = v.x;
                tim[i][j] = v.y;
            }
        }
        __syncthreads();
        for (int p = threadIdx.x; p < R * TL::D; p += blockDim.x) {
            const int k = p / R, i = p % R, jj = i + k - (R - 1);
            const int64_t r = r0 + i, cc = c0 + jj;
            if (base[k] != NOT_STORED && jj >= 0 && jj < C && r < rows && cc < cols) {
                re[base[k] + r] = static_cast<TD>(tre[i][jj]);
                im[base[k] + r] = static_cast<TD>(tim[i][jj]);
            }
        }
        __syncthreads();
    }
}

// DIA -> dense: each tile that meets a stored diagonal reads the diagonals' runs into LDS; every tile writes all of
// its elements (zeros off the stored diagonals) row by row.
template <typename TS, int R, int C>
__global__ __launch_bounds__(256) void dia_to_dense_tiles_kernel(int64_t rows, int64_t cols,
                                                                 const int64_t *__restrict__ rowbase_of,
                                                                 const uint32_t *__restrict__ bits,
                                                                 const TS *__restrict__ re, const TS *__restrict__ im,
                                                                 double *__restrict__ D, int64_t ld, int64_t tiles_c,
                                                                 int64_t ntiles) {
    using TL = Tile<R, C>;
    __shared__ double tre[R][C], tim[R][C];
    __shared__ int64_t base[TL::D];
    __shared__ unsigned char has[CT_TPB];
    classify_block_tiles<R, C>(rows, cols, tiles_c, ntiles, bits, has);
    const int j = threadIdx.x % C;
    for (int q = 0; q < CT_TPB; ++q) {
        const int64_t t = blockIdx.x + static_cast<int64_t>(q) * gridDim.x;
        if (t >= ntiles) break;
        const bool h = has[q] != 0;
        const int64_t r0 = TL::row0(t, tiles_c), c0 = TL::col0(t, tiles_c);
        if (h) {
            TL::stage_bases(r0, c0, rows, cols, rowbase_of, base);
            for (int p = threadIdx.x; p < R * TL::D; p += blockDim.x) {
                const int k = p / R, i = p % R, jj = i + k - (R - 1);
                const int64_t r = r0 + i, cc = c0 + jj;
                if (base[k] != NOT_STORED && jj >= 0 && jj < C && r < rows && cc < cols) {
                    tre[i][jj] = static_cast<double>(re[base[k] + r]);
                    tim[i][jj] = static_cast<double>(im[base[k] + r]);
                }
            }
            __syncthreads();
        }
        const int64_t c = c0 + j;
#pragma unroll
        for (int i = threadIdx.x / C; i < R; i += 256 / C) {
            const int64_t r = r0 + i;
            if (r < rows && c < cols) {
                d2v v;
                v.x = 0.0;
                v.y = 0.0;
                if (h && base[j - i + R - 1] != NOT_STORED) {
                    v.x = tre[i][j];
                    v.y = tim[i][j];
                }
                __builtin_nontemporal_store(v, reinterpret_cast<d2v *>(D + 2 * (r * ld + c)));
            }
        }
        if (h) __syncthreads();  // the LDS tile and bases are rewritten by the next tile
    }
}

// Sparse diagonal sets (the stored elements are at most 1/16 of the dense matrix): per row, a group of RG_DIAGS
// diagonals. Lanes own consecutive rows, so the DIA side is read or written in full runs; the dense side is touched
// only at the stored elements (one 16-byte access each), instead of tile by tile.
constexpr int RG_DIAGS = 8;

static int g_convert_path = 0;  // dq_set_tuning "dense_convert": 0 = automatic, 1 = per-row kernels, 2 = tiled

inline bool sparse_dense_path(const dq_devmat &M) {
    const int64_t groups = (M.n() + RG_DIAGS - 1) / RG_DIAGS;
    if (groups > 65535 || g_convert_path == 2) return false;
    return g_convert_path == 1 || M.stored() * 16 <= static_cast<int64_t>(M.rows) * M.cols;
}

// dq_set_tuning "dense_convert_tile": the tiled kernels' tile, 0 = 32 x 64, 1 = 64 x 32, 2 = 64 x 64, 3 = 128 x 32
static int g_convert_tile_knob = 0;

bool convert_set_tuning(const std::string &key, int value) {
    if (key == "dense_convert_tile") {
        if (value < 0 || value > 3) throw ArgError("dense_convert_tile must be 0, 1, 2 or 3");
        g_convert_tile_knob = value;
        return true;
    }
    if (key != "dense_convert") return false;
    if (value < 0 || value > 2) throw ArgError("dense_convert must be 0, 1 or 2");
    g_convert_path = value;
    return true;
}

template <typename TD>
__global__ __launch_bounds__(256) void dense_to_dia_rows_kernel(int64_t rows, const double *__restrict__ D, int64_t ld,
                                                                int n, const DiagMeta *__restrict__ meta,
                                                                TD *__restrict__ re, TD *__restrict__ im) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const int k1 = min(n, static_cast<int>(blockIdx.y + 1) * RG_DIAGS);
    for (int k = blockIdx.y * RG_DIAGS; k < k1; ++k) {
        const DiagMeta m = meta[k];
        if (r >= m.rlo && r < m.rhi) {
            const d2v v = __builtin_nontemporal_load(reinterpret_cast<const d2v *>(D + 2 * (r * ld + r + m.off)));
            re[m.rowbase + r] = static_cast<TD>(v.x);
            im[m.rowbase + r] = static_cast<TD>(v.y);
        }
    }
}

// Zero the rows x cols block of the dense matrix (row stride ld): one linear sweep when the rows are contiguous.
__global__ __launch_bounds__(256) void zero_dense_kernel(int64_t rows, int64_t cols, double *__restrict__ D,
                                                         int64_t ld) {
    d2v z;
    z.x = 0.0;
    z.y = 0.0;
    if (ld == cols) {
        const int64_t n = rows * cols;
        for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
             e += static_cast<int64_t>(gridDim.x) * blockDim.x)
            *reinterpret_cast<d2v *>(D + 2 * e) = z;
        return;
    }
    for (int64_t r = blockIdx.x; r < rows; r += gridDim.x)
        for (int64_t c = threadIdx.x; c < cols; c += blockDim.x) *reinterpret_cast<d2v *>(D + 2 * (r * ld + c)) = z;
}

template <typename TS>
__global__ __launch_bounds__(256) void dia_to_dense_rows_kernel(int64_t rows, int n,
                                                                const DiagMeta *__restrict__ meta,
                                                                const TS *__restrict__ re, const TS *__restrict__ im,
                                                                double *__restrict__ D, int64_t ld) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const int k1 = min(n, static_cast<int>(blockIdx.y + 1) * RG_DIAGS);
    for (int k = blockIdx.y * RG_DIAGS; k < k1; ++k) {
        const DiagMeta m = meta[k];
        if (r >= m.rlo && r < m.rhi) {
            d2v v;
            v.x = static_cast<double>(re[m.rowbase + r]);
            v.y = static_cast<double>(im[m.rowbase + r]);
            __builtin_nontemporal_store(v, reinterpret_cast<d2v *>(D + 2 * (r * ld + r + m.off)));
        }
    }
}

static void check_dense_args(uint32_t rows, uint32_t cols, const void *dense, int64_t ld) {
    if (rows == 0 || cols == 0) return;
    if (!dense) throw ArgError("dense is NULL");
    if (ld < static_cast<int64_t>(cols)) throw ArgError("ld must be >= cols");
    if (reinterpret_cast<uintptr_t>(dense) % 16 != 0) throw ArgError("dense must be 16-byte aligned");
}

static unsigned tile_grid(int64_t ntiles) { return blocks_for(ntiles, 1, 1 << 16); }

// The per-diagonal tables of the tile kernels in one device buffer: rowbase_of[g] (the rowbase of diagonal
// g - (rows - 1), NOT_STORED if it is not stored) for g < nd, then the bitmap of stored diagonals (bit g of word g / 32).
struct DiagTable {
    DevBuf buf;
    int64_t nd = 0;
    DiagTable(const dq_devmat &M, hipStream_t s) : nd(static_cast<int64_t>(M.rows) + M.cols - 1) {
        const int64_t nw = (nd + 31) / 32;
        std::vector<int64_t> host(nd + (nw + 1) / 2, NOT_STORED);
        uint32_t *words = reinterpret_cast<uint32_t *>(host.data() + nd);
        std::fill(words, words + 2 * ((nw + 1) / 2), 0u);
        for (int k = 0; k < M.n(); ++k) {  // a repeated offset: the last slot
            const int64_t g = M.offs[k] + static_cast<int64_t>(M.rows) - 1;
            host[g] = M.meta[k].rowbase;
            words[g / 32] |= 1u << (g % 32);
        }
        buf = DevBuf(sizeof(int64_t) * host.size());
        DQ_HIP(hipMemcpyAsync(buf.p, host.data(), buf.bytes, hipMemcpyHostToDevice, s));
        DQ_HIP(hipStreamSynchronize(s));  // the host staging vector dies on return
    }
    const int64_t *rowbase() const { return buf.as<int64_t>(); }
    const uint32_t *bits() const { return reinterpret_cast<const uint32_t *>(buf.as<int64_t>() + nd); }
};

static unsigned tile_blocks(int64_t ntiles) { return static_cast<unsigned>((ntiles + CT_TPB - 1) / CT_TPB); }


template <typename T, int R, int C>
void launch_tiles(bool to_dia, const dq_devmat &M, const double *src, double *dst, int64_t ld, const DiagTable &tab,
                  hipStream_t s) {
    const int64_t rows = M.rows, cols = M.cols;
    const int64_t tiles_c = (cols + C - 1) / C, ntiles = ((rows + R - 1) / R) * tiles_c;
    if (to_dia)
        hipLaunchKernelGGL((dense_to_dia_tiles_kernel<T, R, C>), dim3(tile_blocks(ntiles)), dim3(256), 0, s, rows,
                           cols, src, ld, tiles_c, ntiles, tab.rowbase(), tab.bits(), M.re.as<T>(), M.im.as<T>());
    else
        hipLaunchKernelGGL((dia_to_dense_tiles_kernel<T, R, C>), dim3(tile_blocks(ntiles)), dim3(256), 0, s, rows,
                           cols, tab.rowbase(), tab.bits(), M.re.as<T>(), M.im.as<T>(), dst, ld, tiles_c, ntiles);
    DQ_LAUNCH_CHECK();
}

// The tiled kernels in one direction: dense `src` -> M's values (to_dia) or M -> dense `dst`.
template <typename T>
void run_tiles(bool to_dia, const dq_devmat &M, const double *src, double *dst, int64_t ld, hipStream_t s) {
    const DiagTable tab(M, s);
    switch (g_convert_tile_knob) {
        case 1: launch_tiles<T, 64, 32>(to_dia, M, src, dst, ld, tab, s); break;
        case 2: launch_tiles<T, 64, 64>(to_dia, M, src, dst, ld, tab, s); break;
        case 3: launch_tiles<T, 128, 32>(to_dia, M, src, dst, ld, tab, s); break;
        default: launch_tiles<T, 32, 64>(to_dia, M, src, dst, ld, tab, s); break;
    }
    DQ_HIP(hipStreamSynchronize(s));  // the tables are freed on return
}

void dia_to_dense(const dq_devmat &M, double *dense, int64_t ld, hipStream_t s) {
    const int64_t rows = M.rows, cols = M.cols;
    if (rows == 0 || cols == 0) return;
    if (sparse_dense_path(M)) {
        hipLaunchKernelGGL(zero_dense_kernel, dim3(blocks_for(rows * cols, 256, 1 << 16)), dim3(256), 0, s, rows,
                           cols, dense, ld);
        DQ_LAUNCH_CHECK();
        if (M.n()) {
            const dim3 grid(blocks_for(rows, 256, 1LL << 30), (M.n() + RG_DIAGS - 1) / RG_DIAGS);
            if (M.precision == DQ_F64)
                hipLaunchKernelGGL(dia_to_dense_rows_kernel<double>, grid, dim3(256), 0, s, rows, M.n(),
                                   M.d_meta.as<DiagMeta>(), M.re.as<double>(), M.im.as<double>(), dense, ld);
            else
                hipLaunchKernelGGL(dia_to_dense_rows_kernel<float>, grid, dim3(256), 0, s, rows, M.n(),
                                   M.d_meta.as<DiagMeta>(), M.re.as<float>(), M.im.as<float>(), dense, ld);
            DQ_LAUNCH_CHECK();
        }
        return;
    }
    if (M.precision == DQ_F64) run_tiles<double>(false, M, nullptr, dense, ld, s);
    else run_tiles<float>(false, M, nullptr, dense, ld, s);
}

void dense_to_dia(dq_devmat &M, const double *dense, int64_t ld, hipStream_t s) {
    if (M.rows == 0 || M.cols == 0 || M.n() == 0) return;
    if (M.precision == DQ_F64) run_tiles<double>(true, M, dense, nullptr, ld, s);
    else run_tiles<float>(true, M, dense, nullptr, ld, s);
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
        // every stored element is written by pauli_diag_kernel; only the alignment gaps are zeroed
        auto M = make_devmat(static_cast<uint32_t>(N), static_cast<uint32_t>(N), offs.data(),
                             static_cast<int>(offs.size()), DQ_F64, s, false);
        zero_gaps(*M, s);
        if (!offs.empty()) {
            std::vector<int32_t> list_start, list_term;
            pauli_lists(n_qubits, n_terms, x_masks, offs, list_start, list_term);
            const int nt = std::max(1, n_terms);
            DevBuf dx(8 * nt), dz(8 * nt), dr(8 * nt), di(8 * nt), dls(4 * list_start.size()),
                dlt(4 * std::max<size_t>(1, list_term.size()));
            if (n_terms) {
                DQ_HIP(hipMemcpyAsync(dx.p, x_masks, 8 * n_terms, hipMemcpyHostToDevice, s));
                DQ_HIP(hipMemcpyAsync(dz.p, z_masks, 8 * n_terms, hipMemcpyHostToDevice, s));
                DQ_HIP(hipMemcpyAsync(dr.p, coef_re, 8 * n_terms, hipMemcpyHostToDevice, s));
                DQ_HIP(hipMemcpyAsync(di.p, coef_im, 8 * n_terms, hipMemcpyHostToDevice, s));
            }
            DQ_HIP(hipMemcpyAsync(dls.p, list_start.data(), dls.bytes, hipMemcpyHostToDevice, s));
            if (!list_term.empty())
                DQ_HIP(hipMemcpyAsync(dlt.p, list_term.data(), 4 * list_term.size(), hipMemcpyHostToDevice, s));
            const int64_t chunks = (N + 256 * PAULI_RPT - 1) / (256 * PAULI_RPT);
            if (chunks > 0x7fffffff) throw ArgError("pauli_sum_device: too many rows");
            // enough workgroups to fill the GPU: slots are spread over the grid's y when the rows are few
            const int64_t gy = std::min<int64_t>(M->n(), std::max<int64_t>(1, 4096 / chunks));
            hipLaunchKernelGGL(pauli_diag_kernel, dim3(static_cast<unsigned>(chunks), static_cast<unsigned>(gy)),
                               dim3(256), 0, s, N, M->n(), M->d_meta.as<DiagMeta>(), dls.as<int32_t>(),
                               dlt.as<int32_t>(), dx.as<uint64_t>(), dz.as<uint64_t>(), dr.as<double>(),
                               di.as<double>(), M->re.as<double>(), M->im.as<double>());
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

extern "C" int dq_devmat_from_dense_device(uint32_t rows, uint32_t cols, const double *dense, int64_t ld,
                                           int precision, uintptr_t stream, dq_devmat **out) {
    return guard([&] {
        if (!out) throw ArgError("out is NULL");
        if (precision != DQ_F64 && precision != DQ_F32) throw ArgError("precision must be DQ_F64 or DQ_F32");
        check_dense_args(rows, cols, dense, ld);
        hipStream_t s = as_stream(stream);
        if (rows == 0 || cols == 0) {
            *out = make_devmat(rows, cols, nullptr, 0, precision, s).release();
            return;
        }
        const int64_t nd = static_cast<int64_t>(rows) + cols - 1;
        const int64_t tiles_c = (cols + CT_C - 1) / CT_C, ntiles = ((rows + CT_R - 1) / CT_R) * tiles_c;
        // pass 1: which diagonals hold a nonzero (read back: the diagonal set sizes the result)
        std::vector<unsigned char> flags(nd);
        {
            DevBuf dflags(nd);
            zero_bytes(dflags.p, nd, s);
            hipLaunchKernelGGL(dense_nonzero_diags_kernel, dim3(tile_grid(ntiles)), dim3(256), 0, s,
                               static_cast<int64_t>(rows), static_cast<int64_t>(cols), dense, ld, tiles_c, ntiles,
                               dflags.as<unsigned char>());
            DQ_LAUNCH_CHECK();
            DQ_HIP(hipMemcpyAsync(flags.data(), dflags.p, nd, hipMemcpyDeviceToHost, s));
            DQ_HIP(hipStreamSynchronize(s));
        }
        std::vector<int32_t> offs;  // ascending, as the reference stores them
        for (int64_t g = 0; g < nd; ++g)
            if (flags[g]) offs.push_back(static_cast<int32_t>(g - (static_cast<int64_t>(rows) - 1)));
        auto M = make_devmat(rows, cols, offs.data(), static_cast<int>(offs.size()), precision, s, false);
        if (!offs.empty() && sparse_dense_path(*M)) {
            zero_gaps(*M, s);
            const dim3 grid(blocks_for(rows, 256, 1LL << 30), (M->n() + RG_DIAGS - 1) / RG_DIAGS);
            if (precision == DQ_F64)
                hipLaunchKernelGGL(dense_to_dia_rows_kernel<double>, grid, dim3(256), 0, s, static_cast<int64_t>(rows),
                                   dense, ld, M->n(), M->d_meta.as<DiagMeta>(), M->re.as<double>(),
                                   M->im.as<double>());
            else
                hipLaunchKernelGGL(dense_to_dia_rows_kernel<float>, grid, dim3(256), 0, s, static_cast<int64_t>(rows),
                                   dense, ld, M->n(), M->d_meta.as<DiagMeta>(), M->re.as<float>(), M->im.as<float>());
            DQ_LAUNCH_CHECK();
        } else if (!offs.empty()) {
            zero_gaps(*M, s);
            if (precision == DQ_F64) run_tiles<double>(true, *M, dense, nullptr, ld, s);
            else run_tiles<float>(true, *M, dense, nullptr, ld, s);
        }
        *out = M.release();
    });
}

extern "C" int dq_devmat_to_dense_device(const dq_devmat *A, double *dense, int64_t ld, uintptr_t stream) {
    return guard([&] {
        if (!A) throw ArgError("A is NULL");
        check_dense_args(A->rows, A->cols, dense, ld);
        if (A->rows == 0 || A->cols == 0) return;
        hipStream_t s = as_stream(stream);
        const int64_t rows = A->rows, cols = A->cols;
        if (sparse_dense_path(*A)) {
            hipLaunchKernelGGL(zero_dense_kernel, dim3(blocks_for(rows * cols, 256, 1 << 16)), dim3(256), 0, s, rows,
                               cols, dense, ld);
            DQ_LAUNCH_CHECK();
            if (A->n()) {
                const dim3 grid(blocks_for(rows, 256, 1LL << 30), (A->n() + RG_DIAGS - 1) / RG_DIAGS);
                if (A->precision == DQ_F64)
                    hipLaunchKernelGGL(dia_to_dense_rows_kernel<double>, grid, dim3(256), 0, s, rows, A->n(),
                                       A->d_meta.as<DiagMeta>(), A->re.as<double>(), A->im.as<double>(), dense, ld);
                else
                    hipLaunchKernelGGL(dia_to_dense_rows_kernel<float>, grid, dim3(256), 0, s, rows, A->n(),
                                       A->d_meta.as<DiagMeta>(), A->re.as<float>(), A->im.as<float>(), dense, ld);
                DQ_LAUNCH_CHECK();
            }
            DQ_HIP(hipStreamSynchronize(s));
            return;
        }
        if (A->precision == DQ_F64) run_tiles<double>(false, *A, nullptr, dense, ld, s);
        else run_tiles<float>(false, *A, nullptr, dense, ld, s);
    });
}
