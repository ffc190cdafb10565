// Sequences of gates on a state vector (a Trotter step is one), applied in fused passes over HBM.
//
// Applying the gates one at a time (dq_kron_apply_device) streams the whole state through HBM once per gate. Here the
// sequence is cut into passes. A pass loads tiles of 2^L amplitudes into LDS, applies every gate of its stretch of
// the sequence to the tile in order, and writes the tile back: one HBM round trip for the whole stretch.
//
// A tile is the set of amplitudes that agree on every qubit outside the pass's tile qubits T (|T| = L). T holds the
// qubits of the stretch's non-diagonal gates, filled up with the lowest remaining qubits; the planner keeps qubits
// 0..low-1 in T (low = 5 by default), so each tile is made of runs of at least 2^low contiguous amplitudes. A gate
// whose only diagonal is the main one joins any pass: it multiplies each amplitude by the entry that the amplitude's
// own index selects, wherever its qubits lie.
//
// Each gate's arithmetic on each amplitude is the single-gate kernel's (each row adds its terms in the order of the
// gate's diagonal slots, with the rounding sequence of dia_rows.hpp), and the order of the gates is kept, so the
// result is bitwise equal to applying the gates one by one with dq_kron_apply_device (accumulate = 0, in place). The
// plan copies the gates' values: the gates may be freed after planning.
#include <algorithm>
#include <cmath>
#include <vector>

#include "dia_rows.hpp"
#include "dq_common.hpp"

namespace dq {

constexpr int GATES_MAX_K = 4;  // gates on up to four qubits (16 x 16)
constexpr int PASS_MAX_GATES = 64, PASS_MAX_ROWPTR = 512, PASS_MAX_TERMS = 512;  // per pass, staged in LDS

// A gate as the kernel sees it. Its terms are listed row by row: row r of the gate adds, in the order of the gate's
// diagonal slots, value * x[c] for each term (c, value) in [rowptr[rp + r], rowptr[rp + r + 1]) (indexes local to the
// pass). kind 0: a one- or two-qubit general gate on tile bits p .. p + k - 1; 1: main diagonal only, applied per
// amplitude; 2: main diagonal only on qubits outside the tile, so one factor per tile. Non-diagonal gates on three or
// four qubits get a pass of their own, run by the single-gate kernel (the fused kernel then needs no registers for
// 8- and 16-row columns).
struct GateOp {
    int k, q, p, kind, rp;
};

struct PassHdr {
    int n_gates, n_rowptr, n_terms, nc;
    int single, single_q;      // single >= 0: this pass is gate `single` of the plan's own copies, on qubits from single_q
    int tpos[16];              // tile bit i -> global qubit (ascending)
    int cpos[48];              // tile index bit m -> global qubit, for the qubits outside the tile (ascending)
    int64_t gop, rowptr, term;  // first GateOp / rowptr entry / term of this pass in the plan's arrays
};

template <typename T>
struct GateTerm {
    T re, im;
    int c, pad_;
};

// A general gate on tile bits [p, p + k): columns of the tile as in the single-gate kernel, with 2^p in place of
// 2^after. A lane takes CPT columns at a time so that each term read from LDS serves all of them; the rows' sums stay
// in registers until every row of the column is done (the column is updated in place).
template <typename T, int L, int GM>
__device__ inline void tile_gate(const GateOp &G, const int *rowptr, const GateTerm<T> *terms, T *sr, T *si) {
    constexpr int KG = GM == 2 ? 1 : GM == 4 ? 2 : GM == 8 ? 3 : 4;
    constexpr int NCOL = 1 << (L - KG);
    constexpr int CPT = NCOL >= 1024 ? 4 : NCOL >= 512 ? 2 : 1;
    const int p = G.p;
    for (int u0 = threadIdx.x; u0 < NCOL; u0 += 256 * CPT) {
        int base[CPT];
#pragma unroll
        for (int i = 0; i < CPT; ++i) {
            const int u = u0 + 256 * i;
            base[i] = ((u >> p) << (p + KG)) | (u & ((1 << p) - 1));
        }
        Acc<T> ar[CPT][GM], ai[CPT][GM];
#pragma unroll
        for (int r = 0; r < GM; ++r) {
            const int t1 = rowptr[G.rp + r + 1];
            for (int t = rowptr[G.rp + r]; t < t1; ++t) {
                const GateTerm<T> w = terms[t];
#pragma unroll
                for (int i = 0; i < CPT; ++i) {
                    if (u0 + 256 * i < NCOL) {
                        const int a = base[i] + (w.c << p);
                        cmac(ar[i][r], ai[i][r], w.re, w.im, sr[a], si[a]);
                    }
                }
            }
        }
#pragma unroll
        for (int i = 0; i < CPT; ++i) {
            if (u0 + 256 * i < NCOL) {
#pragma unroll
                for (int r = 0; r < GM; ++r) {
                    sr[base[i] + (r << p)] = ar[i][r].get();
                    si[base[i] + (r << p)] = ai[i][r].get();
                }
            }
        }
    }
}

template <typename T, int L>
__global__ __launch_bounds__(256) void gates_pass_kernel(const PassHdr *__restrict__ hdr,
                                                         const GateOp *__restrict__ gops,
                                                         const int *__restrict__ rowptrs,
                                                         const GateTerm<T> *__restrict__ terms_g, T *re, T *im,
                                                         int64_t ntiles) {
    constexpr int E = (1 << L) / 256;  // amplitudes per lane
    __shared__ T sr[1 << L], si[1 << L];
    __shared__ GateOp sg[PASS_MAX_GATES];
    __shared__ int srp[PASS_MAX_ROWPTR];
    __shared__ GateTerm<T> st[PASS_MAX_TERMS];
    const PassHdr &P = *hdr;
    const int ng = P.n_gates;
    for (int i = threadIdx.x; i < ng; i += blockDim.x) sg[i] = gops[P.gop + i];
    for (int i = threadIdx.x; i < P.n_rowptr; i += blockDim.x) srp[i] = rowptrs[P.rowptr + i];
    for (int i = threadIdx.x; i < P.n_terms; i += blockDim.x) st[i] = terms_g[P.term + i];
    // global offset of tile element e = deposit of e's bits at the tile qubits
    int64_t off[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int e = threadIdx.x + 256 * m;
        int64_t v = 0;
#pragma unroll
        for (int i = 0; i < L; ++i) v |= static_cast<int64_t>((e >> i) & 1) << P.tpos[i];
        off[m] = v;
    }
    __syncthreads();
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        int64_t base = 0;
        for (int i = 0; i < P.nc; ++i) base |= ((t >> i) & 1) << P.cpos[i];
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int e = threadIdx.x + 256 * m;
            sr[e] = re[base + off[m]];
            si[e] = im[base + off[m]];
        }
        bool own = true;  // the LDS entries this lane last wrote are its own elements e = tid + 256 m
        __syncthreads();
        for (int j = 0; j < ng; ++j) {
            const GateOp G = sg[j];
            if (G.kind == 0) {
                __syncthreads();
                if (G.k == 1) tile_gate<T, L, 2>(G, srp, st, sr, si);
                else tile_gate<T, L, 4>(G, srp, st, sr, si);
                own = false;
            } else {
                // a run of diagonal gates: each amplitude goes through all of them in registers, gate by gate with
                // the same rounding as one at a time (kind 2, qubits outside the tile: the index bits come from
                // `base` alone, so c is the same for the whole tile)
                if (!own) __syncthreads();
                own = true;
                int j2 = j;
                while (j2 + 1 < ng && (sg[j2 + 1].kind == 1 || sg[j2 + 1].kind == 2)) ++j2;
                T xr[E], xi[E];
#pragma unroll
                for (int m = 0; m < E; ++m) {
                    xr[m] = sr[threadIdx.x + 256 * m];
                    xi[m] = si[threadIdx.x + 256 * m];
                }
#pragma unroll 1
                for (int jj = j; jj <= j2; ++jj) {
                    const GateOp D = sg[jj];
                    const GateTerm<T> *row = st + srp[D.rp];  // one term per row: row c's term is row[c]
                    const int mask = (1 << D.k) - 1;
#pragma unroll
                    for (int m = 0; m < E; ++m) {
                        const GateTerm<T> w = row[static_cast<int>(((base + off[m]) >> D.q) & mask)];
                        Acc<T> ar, ai;
                        cmac(ar, ai, w.re, w.im, xr[m], xi[m]);
                        xr[m] = ar.get();
                        xi[m] = ai.get();
                    }
                }
#pragma unroll
                for (int m = 0; m < E; ++m) {
                    sr[threadIdx.x + 256 * m] = xr[m];
                    si[threadIdx.x + 256 * m] = xi[m];
                }
                j = j2;
            }
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int e = threadIdx.x + 256 * m;
            re[base + off[m]] = sr[e];
            im[base + off[m]] = si[e];
        }
        __syncthreads();  // the next tile reuses the LDS
    }
}

// ------------------------------------------------------------------------------------------------------------------
// Register layout: a lane keeps its E amplitudes of the tile (tile element e = threadIdx.x + 256 m) in registers for
// the whole pass. Tile bits 0..5 are lane bits within a wave, 6..7 the wave, 8.. the register index m. A gate on lane
// bits takes its partners by cross-lane shuffles, one on register bits from the lane's own registers, and only a
// gate touching a wave bit exchanges through LDS. Each lane computes the rows of the amplitudes it owns, with the
// same terms in the same order as the single-gate kernel, so the bits are the same.
// ------------------------------------------------------------------------------------------------------------------
// Row r of a gate over operands v[0..GM): the terms of the row in slot order.
template <typename T, int GM>
__device__ inline void gate_row(const int *rowptr, const GateTerm<T> *terms, int rp, int r, const T (&vr)[GM],
                                const T (&vi)[GM], T &out_r, T &out_i) {
    Acc<T> ar, ai;
    const int t1 = rowptr[rp + r + 1];
    for (int t = rowptr[rp + r]; t < t1; ++t) {
        const GateTerm<T> w = terms[t];
        T xr = vr[0], xi = vi[0];
#pragma unroll
        for (int c = 1; c < GM; ++c) {
            if (w.c == c) {
                xr = vr[c];
                xi = vi[c];
            }
        }
        cmac(ar, ai, w.re, w.im, xr, xi);
    }
    out_r = ar.get();
    out_i = ai.get();
}

// A gate of 2^K rows on tile bits [p, p + K) inside the lane bits (p + K <= 6): partners by shuffles.
template <typename T, int E, int K>
__device__ inline void regs_gate_lanes(const GateOp &G, const int *rowptr, const GateTerm<T> *terms, T (&xr)[E],
                                       T (&xi)[E]) {
    constexpr int GM = 1 << K;
    const int lane = threadIdx.x & 63;
    const int r = (lane >> G.p) & (GM - 1);
#pragma unroll
    for (int m = 0; m < E; ++m) {
        // partner d (d = c xor r, 1 <= d < GM) sits in lane lane ^ (d << p): one exchange per partner, the same
        // masks for every lane
        T wr[GM], wi[GM];
        wr[0] = xr[m];
        wi[0] = xi[m];
#pragma unroll
        for (int dd = 1; dd < GM; ++dd) {
            wr[dd] = __shfl_xor(xr[m], dd << G.p);
            wi[dd] = __shfl_xor(xi[m], dd << G.p);
        }
        T vr[GM], vi[GM];
#pragma unroll
        for (int c = 0; c < GM; ++c) {
            vr[c] = wr[0];
            vi[c] = wi[0];
#pragma unroll
            for (int dd = 1; dd < GM; ++dd) {
                if ((c ^ r) == dd) {
                    vr[c] = wr[dd];
                    vi[c] = wi[dd];
                }
            }
        }
        gate_row<T, GM>(rowptr, terms, G.rp, r, vr, vi, xr[m], xi[m]);
    }
}

// A gate of 2^K rows on register bits [RB, RB + K) of m: the partners are the lane's own registers.
template <typename T, int E, int K, int RB>
__device__ inline void regs_gate_regs(const GateOp &G, const int *rowptr, const GateTerm<T> *terms, T (&xr)[E],
                                      T (&xi)[E]) {
    constexpr int GM = 1 << K;
    if constexpr (RB + K <= 31 && (E >> (RB + K)) >= 1) {
#pragma unroll
        for (int m0 = 0; m0 < E; ++m0) {
            if (m0 & ((GM - 1) << RB)) continue;
            T vr[GM], vi[GM], yr[GM], yi[GM];
#pragma unroll
            for (int c = 0; c < GM; ++c) {
                vr[c] = xr[m0 | (c << RB)];
                vi[c] = xi[m0 | (c << RB)];
            }
#pragma unroll
            for (int r = 0; r < GM; ++r) gate_row<T, GM>(rowptr, terms, G.rp, r, vr, vi, yr[r], yi[r]);
#pragma unroll
            for (int c = 0; c < GM; ++c) {
                xr[m0 | (c << RB)] = yr[c];
                xi[m0 | (c << RB)] = yi[c];
            }
        }
    }
}

template <typename T, int E, int K, int RB>
__device__ inline void regs_gate_regs_dispatch(int rb, const GateOp &G, const int *rowptr, const GateTerm<T> *terms,
                                               T (&xr)[E], T (&xi)[E]) {
    if constexpr ((1 << (RB + K)) <= E) {
        if (rb == RB) regs_gate_regs<T, E, K, RB>(G, rowptr, terms, xr, xi);
        else regs_gate_regs_dispatch<T, E, K, RB + 1>(rb, G, rowptr, terms, xr, xi);
    }
}

template <typename T, int L>
__global__ __launch_bounds__(256) void gates_regs_kernel(const PassHdr *__restrict__ hdr,
                                                         const GateOp *__restrict__ gops,
                                                         const int *__restrict__ rowptrs,
                                                         const GateTerm<T> *__restrict__ terms_g, T *re, T *im,
                                                         int64_t ntiles) {
    constexpr int E = (1 << L) / 256;  // amplitudes per lane
    __shared__ T sr[1 << L], si[1 << L];  // only for gates that touch the wave bits 6 and 7
    __shared__ GateOp sg[PASS_MAX_GATES];
    __shared__ int srp[PASS_MAX_ROWPTR];
    __shared__ GateTerm<T> st[PASS_MAX_TERMS];
    const PassHdr &P = *hdr;
    const int ng = P.n_gates;
    for (int i = threadIdx.x; i < ng; i += blockDim.x) sg[i] = gops[P.gop + i];
    for (int i = threadIdx.x; i < P.n_rowptr; i += blockDim.x) srp[i] = rowptrs[P.rowptr + i];
    for (int i = threadIdx.x; i < P.n_terms; i += blockDim.x) st[i] = terms_g[P.term + i];
    int64_t off[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int e = threadIdx.x + 256 * m;
        int64_t v = 0;
#pragma unroll
        for (int i = 0; i < L; ++i) v |= static_cast<int64_t>((e >> i) & 1) << P.tpos[i];
        off[m] = v;
    }
    __syncthreads();
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        int64_t base = 0;
        for (int i = 0; i < P.nc; ++i) base |= ((t >> i) & 1) << P.cpos[i];
        T xr[E], xi[E];
#pragma unroll
        for (int m = 0; m < E; ++m) {
            xr[m] = re[base + off[m]];
            xi[m] = im[base + off[m]];
        }
        for (int j = 0; j < ng; ++j) {
            const GateOp G = sg[j];
            if (G.kind != 0) {  // diagonal: per amplitude, from its own index
                const GateTerm<T> *row = st + srp[G.rp];
                const int mask = (1 << G.k) - 1;
#pragma unroll
                for (int m = 0; m < E; ++m) {
                    const GateTerm<T> w = row[static_cast<int>(((base + off[m]) >> G.q) & mask)];
                    Acc<T> ar, ai;
                    cmac(ar, ai, w.re, w.im, xr[m], xi[m]);
                    xr[m] = ar.get();
                    xi[m] = ai.get();
                }
            } else if (G.p + G.k <= 6) {
                if (G.k == 1) regs_gate_lanes<T, E, 1>(G, srp, st, xr, xi);
                else regs_gate_lanes<T, E, 2>(G, srp, st, xr, xi);
            } else if (G.p >= 8) {
                if (G.k == 1) regs_gate_regs_dispatch<T, E, 1, 0>(G.p - 8, G, srp, st, xr, xi);
                else regs_gate_regs_dispatch<T, E, 2, 0>(G.p - 8, G, srp, st, xr, xi);
            } else {  // touches a wave bit: exchange through LDS
                __syncthreads();  // earlier readers of the LDS copy are done
#pragma unroll
                for (int m = 0; m < E; ++m) {
                    sr[threadIdx.x + 256 * m] = xr[m];
                    si[threadIdx.x + 256 * m] = xi[m];
                }
                __syncthreads();
                const int GM = 1 << G.k;
#pragma unroll
                for (int m = 0; m < E; ++m) {
                    const int e = threadIdx.x + 256 * m;
                    const int r = (e >> G.p) & (GM - 1);
                    Acc<T> ar, ai;
                    const int t1 = srp[G.rp + r + 1];
                    for (int tt = srp[G.rp + r]; tt < t1; ++tt) {
                        const GateTerm<T> w = st[tt];
                        const int a = e ^ ((w.c ^ r) << G.p);
                        cmac(ar, ai, w.re, w.im, sr[a], si[a]);
                    }
                    xr[m] = ar.get();
                    xi[m] = ai.get();
                }
            }
        }
#pragma unroll
        for (int m = 0; m < E; ++m) {
            re[base + off[m]] = xr[m];
            im[base + off[m]] = xi[m];
        }
        __syncthreads();  // the next tile's LDS exchanges start after every lane is done with this one's
    }
}

}  // namespace dq

// A planned gate sequence: passes, gate ops, row pointers and terms (copies of the gates' values), built once and run
// any number of times.
struct dq_gate_plan {
    int n_qubits = 0, precision = DQ_F64, L = 11, kernel = 0;
    std::vector<dq::PassHdr> passes;
    std::vector<std::unique_ptr<dq_devmat>> big;  // copies of the non-diagonal gates on three or four qubits
    dq::DevBuf d_passes, d_gops, d_rowptr, d_terms;
};

namespace dq {

static int g_gates_tile_qubits = 11;  // dq_set_tuning "gates_tile_qubits": 10, 11 or 12
static int g_gates_low_qubits = 5;    // dq_set_tuning "gates_low_qubits": 4, 5 or 6 (qubits always in a tile)
static int g_gates_kernel = 0;        // dq_set_tuning "gates_kernel": 0 = LDS tiles, 1 = register layout

bool gates_set_tuning(const std::string &key, int value) {
    if (key == "gates_kernel") {
        if (value != 0 && value != 1) throw ArgError("gates_kernel must be 0 or 1");
        g_gates_kernel = value;
        return true;
    }
    if (key == "gates_low_qubits") {
        if (value < 4 || value > 6) throw ArgError("gates_low_qubits must be 4, 5 or 6");
        g_gates_low_qubits = value;
        return true;
    }
    if (key != "gates_tile_qubits") return false;
    if (value < 10 || value > 12) throw ArgError("gates_tile_qubits must be 10, 11 or 12");
    g_gates_tile_qubits = value;
    return true;
}

static int gate_qubits(const dq_devmat &G) {
    if (G.rows != G.cols) throw ArgError("gates: every gate must be square");
    int k = 0;
    while ((1u << k) < G.rows) ++k;
    if ((1u << k) != G.rows || k < 1 || k > GATES_MAX_K)
        throw ArgError("gates: a gate must be 2^k x 2^k with 1 <= k <= 4");
    return k;
}

// One gate on the host: qubits, and its terms row by row in slot order.
struct HostGate {
    int k = 0, q = 0;
    bool diag = false;
    std::vector<int> rowptr;                // g + 1
    std::vector<int> col;                   // per term
    std::vector<double> vre, vim;           // per term (exact copies of the gate's FP64 or FP32 values)
};

static HostGate host_gate(const dq_devmat &G, int k, int q) {
    HostGate H;
    H.k = k;
    H.q = q;
    H.diag = G.n() == 1 && G.offs[0] == 0;
    const int g = 1 << k;
    const size_t eb = elem_bytes(G.precision);
    std::vector<char> re(G.total * eb), im(G.total * eb);
    if (G.total) {
        DQ_HIP(hipMemcpy(re.data(), G.re.p, re.size(), hipMemcpyDeviceToHost));
        DQ_HIP(hipMemcpy(im.data(), G.im.p, im.size(), hipMemcpyDeviceToHost));
    }
    auto val = [&](const std::vector<char> &v, int64_t i) {
        return G.precision == DQ_F64 ? reinterpret_cast<const double *>(v.data())[i]
                                     : static_cast<double>(reinterpret_cast<const float *>(v.data())[i]);
    };
    H.rowptr.push_back(0);
    for (int r = 0; r < g; ++r) {
        for (int s = 0; s < G.n(); ++s) {
            const DiagMeta &m = G.meta[s];
            if (r >= m.rlo && r < m.rhi) {
                H.col.push_back(r + m.off);
                H.vre.push_back(val(re, m.rowbase + r));
                H.vim.push_back(val(im, m.rowbase + r));
            }
        }
        H.rowptr.push_back(static_cast<int>(H.col.size()));
    }
    return H;
}

// Greedy cut of the sequence into passes: a non-diagonal gate joins the current pass while the qubits of the pass's
// non-diagonal gates, together with qubits 0..low-1, number at most L, and while the pass's tables fit in LDS.
template <typename T>
static void plan_passes(dq_gate_plan &P, const std::vector<HostGate> &gates,
                        const std::vector<const dq_devmat *> &devgates) {
    const int n = P.n_qubits, L = P.L;
    std::vector<GateOp> ops;
    std::vector<int> rowptr;
    std::vector<GateTerm<T>> terms;
    auto close = [&](int g0, int g1, uint64_t S) {
        if (g1 <= g0) return;
        uint64_t Tm = S;
        for (int b = 0; b < n && __builtin_popcountll(Tm) < L; ++b) Tm |= uint64_t(1) << b;
        PassHdr D{};
        D.single = -1;
        int i = 0, c = 0;
        for (int b = 0; b < n; ++b) {
            if ((Tm >> b) & 1) D.tpos[i++] = b;
            else D.cpos[c++] = b;
        }
        D.nc = c;
        D.gop = static_cast<int64_t>(ops.size());
        D.rowptr = static_cast<int64_t>(rowptr.size());
        D.term = static_cast<int64_t>(terms.size());
        int nrp = 0, nt = 0;
        for (int j = g0; j < g1; ++j) {
            const HostGate &H = gates[j];
            GateOp op{};
            op.k = H.k;
            op.q = H.q;
            op.rp = nrp;
            const uint64_t Q = ((uint64_t(1) << H.k) - 1) << H.q;
            if (H.diag) {
                op.kind = (Q & Tm) ? 1 : 2;
            } else {
                int pos = 0;
                while (D.tpos[pos] != H.q) ++pos;
                op.p = pos;  // the gate's qubits are consecutive, so are their tile bits
            }
            for (int r : H.rowptr) rowptr.push_back(nt + r);
            for (size_t t = 0; t < H.col.size(); ++t)
                terms.push_back(GateTerm<T>{static_cast<T>(H.vre[t]), static_cast<T>(H.vim[t]), H.col[t], 0});
            nrp += static_cast<int>(H.rowptr.size());
            nt += static_cast<int>(H.col.size());
            ops.push_back(op);
        }
        D.n_gates = g1 - g0;
        D.n_rowptr = nrp;
        D.n_terms = nt;
        P.passes.push_back(D);
    };
    const uint64_t low = (uint64_t(1) << g_gates_low_qubits) - 1;
    uint64_t S = 0;
    int g0 = 0, nrp = 0, nt = 0;
    for (int j = 0; j < static_cast<int>(gates.size()); ++j) {
        const HostGate &H = gates[j];
        if (!H.diag && H.k > 2) {  // a pass of its own: the single-gate kernel, on a copy the plan keeps
            if (devgates[j]->n() > 64) throw ArgError("gates: a gate has more than 64 diagonals");
            close(g0, j, S);
            PassHdr D{};
            D.single = static_cast<int>(P.big.size());
            D.single_q = H.q;
            P.big.push_back(copy_devmat(*devgates[j], nullptr));
            P.passes.push_back(D);
            g0 = j + 1;
            S = 0;
            nrp = nt = 0;
            continue;
        }
        const uint64_t Q = H.diag ? 0 : ((uint64_t(1) << H.k) - 1) << H.q;
        const int hr = static_cast<int>(H.rowptr.size()), ht = static_cast<int>(H.col.size());
        if (ht > PASS_MAX_TERMS) throw ArgError("gates: a gate has more than 512 nonzero terms");
        if (j > g0 && (__builtin_popcountll(S | Q | low) > L || j - g0 >= PASS_MAX_GATES ||
                       nrp + hr > PASS_MAX_ROWPTR || nt + ht > PASS_MAX_TERMS)) {
            close(g0, j, S);
            g0 = j;
            S = 0;
            nrp = nt = 0;
        }
        S |= Q;
        nrp += hr;
        nt += ht;
    }
    close(g0, static_cast<int>(gates.size()), S);
    DQ_HIP(hipDeviceSynchronize());  // the copies of the large gates are complete
    P.d_passes = DevBuf(sizeof(PassHdr) * P.passes.size());
    P.d_gops = DevBuf(sizeof(GateOp) * std::max<size_t>(1, ops.size()));
    P.d_rowptr = DevBuf(sizeof(int) * std::max<size_t>(1, rowptr.size()));
    P.d_terms = DevBuf(sizeof(GateTerm<T>) * std::max<size_t>(1, terms.size()));
    DQ_HIP(hipMemcpy(P.d_passes.p, P.passes.data(), P.d_passes.bytes, hipMemcpyHostToDevice));
    if (!ops.empty()) DQ_HIP(hipMemcpy(P.d_gops.p, ops.data(), sizeof(GateOp) * ops.size(), hipMemcpyHostToDevice));
    if (!rowptr.empty())
        DQ_HIP(hipMemcpy(P.d_rowptr.p, rowptr.data(), sizeof(int) * rowptr.size(), hipMemcpyHostToDevice));
    if (!terms.empty())
        DQ_HIP(hipMemcpy(P.d_terms.p, terms.data(), sizeof(GateTerm<T>) * terms.size(), hipMemcpyHostToDevice));
}

template <typename T>
static void run_passes(const dq_gate_plan &P, void *re, void *im, hipStream_t s) {
    const int64_t ntiles = int64_t(1) << (P.n_qubits - P.L);
    const unsigned grid = blocks_for(ntiles, 1, 2048);
    const PassHdr *dp = P.d_passes.as<PassHdr>();
    const GateOp *dg = P.d_gops.as<GateOp>();
    const int *dr = P.d_rowptr.as<int>();
    const GateTerm<T> *dt = P.d_terms.as<GateTerm<T>>();
    for (size_t i = 0; i < P.passes.size(); ++i) {
        if (P.passes[i].single >= 0) {
            const dq_devmat &G = *P.big[P.passes[i].single];
            int k = 0;
            while ((1u << k) < G.rows) ++k;
            const int q = P.passes[i].single_q;
            kron_apply(G, P.n_qubits - q - k, q, re, im, re, im, false, s);
            continue;
        }
#define DQ_GATES_PASS(LL)                                                                                         \
    if (P.kernel == 1)                                                                                            \
        hipLaunchKernelGGL((gates_regs_kernel<T, LL>), dim3(grid), dim3(256), 0, s, dp + i, dg, dr, dt,            \
                           static_cast<T *>(re), static_cast<T *>(im), ntiles);                                    \
    else                                                                                                          \
        hipLaunchKernelGGL((gates_pass_kernel<T, LL>), dim3(grid), dim3(256), 0, s, dp + i, dg, dr, dt,            \
                           static_cast<T *>(re), static_cast<T *>(im), ntiles)
        if (P.L == 10) {
            DQ_GATES_PASS(10);
        } else if (P.L == 12) {
            DQ_GATES_PASS(12);
        } else {
            DQ_GATES_PASS(11);
        }
#undef DQ_GATES_PASS
        DQ_LAUNCH_CHECK();
    }
}

}  // namespace dq

using namespace dq;

extern "C" int dq_gate_plan_create(int n_qubits, int n_gates, const dq_devmat *const *gates,
                                   const uint32_t *first_qubit, dq_gate_plan **out) {
    return guard([&] {
        if (!out) throw ArgError("out is NULL");
        if (n_gates < 1 || !gates || !first_qubit) throw ArgError("gates: need at least one gate");
        const int L = g_gates_tile_qubits;
        if (n_qubits < L || n_qubits > 40)
            throw ArgError("gates: n_qubits must be in [" + std::to_string(L) + ", 40] (smaller states: "
                           "dq_kron_apply_device per gate)");
        auto P = std::make_unique<dq_gate_plan>();
        P->n_qubits = n_qubits;
        P->L = L;
        P->kernel = g_gates_kernel;
        std::vector<HostGate> hg;
        for (int j = 0; j < n_gates; ++j) {
            const dq_devmat *G = gates[j];
            if (!G) throw ArgError("gates: NULL gate");
            const int k = gate_qubits(*G);
            if (j == 0) P->precision = G->precision;
            else if (G->precision != P->precision) throw ArgError("gates: all gates need one precision");
            if (first_qubit[j] + k > static_cast<uint32_t>(n_qubits)) throw ArgError("gates: gate beyond n_qubits");
            hg.push_back(host_gate(*G, k, static_cast<int>(first_qubit[j])));
        }
        std::vector<const dq_devmat *> dg(gates, gates + n_gates);
        if (P->precision == DQ_F64) plan_passes<double>(*P, hg, dg);
        else plan_passes<float>(*P, hg, dg);
        *out = P.release();
    });
}

extern "C" int dq_gate_plan_passes(const dq_gate_plan *plan, int *n_passes) {
    return guard([&] {
        if (!plan || !n_passes) throw ArgError("NULL argument");
        *n_passes = static_cast<int>(plan->passes.size());
    });
}

extern "C" int dq_gate_plan_run(const dq_gate_plan *plan, void *re, void *im, int repetitions, uintptr_t stream) {
    return guard([&] {
        if (!plan) throw ArgError("plan is NULL");
        if (!re || !im) throw ArgError("gates: NULL state");
        if (repetitions < 0) throw ArgError("gates: repetitions must be >= 0");
        hipStream_t s = as_stream(stream);
        for (int r = 0; r < repetitions; ++r) {
            if (plan->precision == DQ_F64) run_passes<double>(*plan, re, im, s);
            else run_passes<float>(*plan, re, im, s);
        }
    });
}

extern "C" void dq_gate_plan_destroy(dq_gate_plan *plan) { delete plan; }
