// diaq_gpu: the Python module of the MI355X DiaQ GPU path, a drop-in for the reference's diaq_gpu module
// (include/pythonWrappers.cpp:31-35, GPU block :589-654, shared entries :41-587).
//
// Host C++ only: the host DIA container, host vectors and conversions live here; every GPU operation goes through the
// C ABI of libdiaq_hip.so (include/diaq_hip.h). C ABI failures are rethrown as RuntimeError with the library's
// message, as the reference's std::runtime_error are (pythonWrappers.cpp; src/cuda_check.hpp:11-18).
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <array>
#include <cmath>
#include <complex>
#include <cstring>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "diaq_hip.h"

namespace py = pybind11;
using cplx = std::complex<double>;

static void check(int rc) {
    if (rc != DQ_OK) throw std::runtime_error(dq_last_error());
}

// ---------------------------------------------------------------------------------------------------------------------
// Host DIA container (src/diaQMatrix.hpp:43-63): split real / imaginary storage with every diagonal start rounded up
// to 8 elements (the reference's AVX-512 build: initStorage, :290-357), so diagOffsets / totalLength read the same.
// ---------------------------------------------------------------------------------------------------------------------
static constexpr unsigned VEC_WIDTH = 8;

static unsigned diag_len(int d, const std::array<unsigned, 2> &s) {
    long long n = d < 0 ? (long long)s[0] + d : (d > 0 ? (long long)s[1] - d : std::min(s[0], s[1]));
    if (n > s[0]) n = s[0];
    if (n > s[1]) n = s[1];
    return n < 0 ? 0u : (unsigned)n;
}

static std::vector<int> symmetric_order(unsigned n) {  // src/diaQMatrix.hpp:25-39
    std::vector<int> o;
    for (unsigned i = 0; i < n; ++i) o.push_back(i == 0 ? 0 : (i % 2 == 1 ? -(int)((i + 1) / 2) : (int)(i / 2)));
    return o;
}

struct HostDia {
    std::array<unsigned, 2> matrix_shape{0, 0};
    std::vector<unsigned> shape;
    std::vector<int> dIndices;
    std::vector<unsigned> diagOffsets;
    unsigned totalLength = 0;
    std::vector<double> re, im;

    HostDia() = default;
    HostDia(std::array<unsigned, 2> ms, std::vector<int> d, std::vector<unsigned> ts)
        : matrix_shape(ms), shape(std::move(ts)), dIndices(std::move(d)) {
        init_storage();
    }
    virtual ~HostDia() = default;

    void init_storage() {
        diagOffsets.resize(dIndices.size());
        unsigned total = 0;
        for (size_t k = 0; k < dIndices.size(); ++k) {
            if (total % VEC_WIDTH) total += VEC_WIDTH - total % VEC_WIDTH;
            diagOffsets[k] = total;
            total += diag_len(dIndices[k], matrix_shape);
        }
        totalLength = total;
        re.assign(total, 0.0);
        im.assign(total, 0.0);
    }
    unsigned len(size_t k) const { return diag_len(dIndices[k], matrix_shape); }
    int slot(int d) const {
        for (size_t k = 0; k < dIndices.size(); ++k)
            if (dIndices[k] == d) return (int)k;
        return -1;
    }
    std::shared_ptr<HostDia> copy() const { return std::make_shared<HostDia>(*this); }
};

struct HostIdentity : HostDia {  // src/specialMatrices.hpp:9-24
    explicit HostIdentity(unsigned s) : HostDia({s, s}, {0}, {s, s}) {
        std::fill(re.begin(), re.begin() + s, 1.0);
    }
};

// permutationMatrix(s, row_order): P[i, row_order[i]] = 1, stored as the ascending nonzero diagonals of that dense
// matrix, as the reference builds it (src/specialMatrices.hpp:26-51); entries with row_order[i] >= s are absent.
struct HostPermutation : HostDia {
    HostPermutation(unsigned s, const std::vector<unsigned> &row_order) {
        if (row_order.size() != s) throw std::invalid_argument("permutationMatrix: row_order needs s entries");
        std::vector<int> ds;
        for (unsigned i = 0; i < s; ++i)
            if (row_order[i] < s) ds.push_back(static_cast<int>(row_order[i]) - static_cast<int>(i));
        std::sort(ds.begin(), ds.end());
        ds.erase(std::unique(ds.begin(), ds.end()), ds.end());
        matrix_shape = {s, s};
        shape = {s, s};
        dIndices = ds;
        init_storage();
        for (unsigned i = 0; i < s; ++i) {
            if (row_order[i] >= s) continue;
            const int d = static_cast<int>(row_order[i]) - static_cast<int>(i);
            const size_t k = std::lower_bound(ds.begin(), ds.end(), d) - ds.begin();
            const unsigned pos = d >= 0 ? i : row_order[i];  // position p holds A[p, p+d] (d >= 0) or A[p-d, p]
            re[diagOffsets[k] + pos] = 1.0;
        }
    }
};

// denseMatrix: the reference's dense container (src/denseMatrix.hpp:8-170), kept on the host; convertToDiaQ applies
// the dense -> ascending nonzero diagonals rule (src/dense_to_diagonals.hpp:20-58).
struct HostDense {
    std::array<unsigned, 2> matrix_shape{0, 0};
    std::vector<unsigned> shape;
    std::vector<double> re, im;
    HostDense(unsigned s, std::vector<double> r, std::vector<double> i)
        : matrix_shape{s, s}, shape{s, s}, re(std::move(r)), im(std::move(i)) {
        const size_t n = static_cast<size_t>(s) * s;
        if (re.size() != n || im.size() != n)
            throw std::invalid_argument("denseMatrix: the value arrays need s * s entries");
    }
};

struct HostVector {  // src/vector.hpp:8-299 (the parts the GPU path uses)
    unsigned numRows = 0;
    std::vector<double> re, im;
    explicit HostVector(unsigned n) : numRows(n), re(n, 0.0), im(n, 0.0) {}
};

// Shared shape / tensor-shape checks of the bindings.
static std::array<unsigned, 2> pair_of(const std::vector<unsigned> &s) {
    if (s.size() != 2) throw std::invalid_argument("a matrix shape needs exactly 2 entries");
    return {s[0], s[1]};
}

// ---------------------------------------------------------------------------------------------------------------------
// Device matrices (DeviceMatrix, src/deviceMatrix_gpu.hpp:37-68): an owning handle around the C ABI object.
// ---------------------------------------------------------------------------------------------------------------------
struct DeviceMatrix {
    std::shared_ptr<dq_devmat> h;
    std::vector<unsigned> tensor_shape;
    explicit DeviceMatrix(dq_devmat *p, std::vector<unsigned> ts = {})
        : h(p, dq_devmat_destroy), tensor_shape(std::move(ts)) {}
    std::array<unsigned, 2> matrix_shape() const {
        uint32_t r, c;
        check(dq_devmat_info(h.get(), &r, &c, nullptr, nullptr, nullptr));
        return {r, c};
    }
    int n_diags() const {
        int32_t n;
        check(dq_devmat_info(h.get(), nullptr, nullptr, &n, nullptr, nullptr));
        return n;
    }
    int precision() const {
        int p;
        check(dq_devmat_info(h.get(), nullptr, nullptr, nullptr, &p, nullptr));
        return p;
    }
    std::vector<int> diagonals() const {
        std::vector<int32_t> o(n_diags());
        check(dq_devmat_offsets(h.get(), o.data()));
        return std::vector<int>(o.begin(), o.end());
    }
};

static int parse_precision(const std::string &p) {
    if (p == "fp64" || p == "float64" || p == "complex128") return DQ_F64;
    if (p == "fp32" || p == "float32" || p == "complex64") return DQ_F32;
    throw std::invalid_argument("precision must be 'fp64' or 'fp32'");
}

static std::shared_ptr<DeviceMatrix> upload(const HostDia &A, int precision = DQ_F64) {
    std::vector<int32_t> offs(A.dIndices.begin(), A.dIndices.end());
    std::vector<int64_t> starts(A.diagOffsets.begin(), A.diagOffsets.end());
    dq_devmat *p = nullptr;
    check(dq_devmat_upload(A.matrix_shape[0], A.matrix_shape[1], (int32_t)offs.size(), offs.data(), starts.data(),
                           A.re.data(), A.im.data(), precision, &p));
    return std::make_shared<DeviceMatrix>(p, A.shape);
}

static std::shared_ptr<HostDia> download(const DeviceMatrix &D) {
    auto ms = D.matrix_shape();
    std::vector<unsigned> ts = D.tensor_shape;
    if (ts.empty()) ts = {ms[0], ms[1]};
    auto H = std::make_shared<HostDia>(ms, D.diagonals(), ts);
    std::vector<int64_t> starts(H->diagOffsets.begin(), H->diagOffsets.end());
    check(dq_devmat_download(D.h.get(), starts.data(), H->re.data(), H->im.data()));
    return H;
}

static std::shared_ptr<DeviceMatrix> wrap(dq_devmat *p, std::vector<unsigned> ts = {}) {
    return std::make_shared<DeviceMatrix>(p, std::move(ts));
}

// ---------------------------------------------------------------------------------------------------------------------
// Conversions (src/conversions.cpp:70-212, src/dense_to_diagonals.hpp:20-58)
// ---------------------------------------------------------------------------------------------------------------------
static std::shared_ptr<HostDia> diagonals_from_dense(std::array<unsigned, 2> ms, const std::vector<unsigned> &ts,
                                                     const cplx *v) {
    const long long rows = ms[0], cols = ms[1];
    std::vector<int> diags;
    for (long long d = -(rows - 1); d < cols; ++d) {
        const long long r0 = d >= 0 ? 0 : -d, c0 = d >= 0 ? d : 0;
        const long long n = std::min(rows - r0, cols - c0);
        for (long long j = 0; j < n; ++j) {
            const cplx z = v[(r0 + j) * cols + c0 + j];
            if (z.real() != 0.0 || z.imag() != 0.0) { diags.push_back((int)d); break; }
        }
    }
    auto M = std::make_shared<HostDia>(ms, diags, ts);
    for (size_t k = 0; k < diags.size(); ++k) {
        const long long d = diags[k], r0 = d >= 0 ? 0 : -d, c0 = d >= 0 ? d : 0;
        const long long n = std::min(rows - r0, cols - c0);
        for (long long j = 0; j < n; ++j) {
            const cplx z = v[(r0 + j) * cols + c0 + j];
            M->re[M->diagOffsets[k] + j] = z.real();
            M->im[M->diagOffsets[k] + j] = z.imag();
        }
    }
    return M;
}

static std::shared_ptr<HostDia> from_numpy_shape(const cplx *v, const std::vector<unsigned> &shape) {
    if (shape.size() > 2) {  // higher-order tensor: the closest factor pair as the matrix shape (conversions.cpp:193-222)
        unsigned long long n = 1;
        for (unsigned s : shape) n *= s;
        unsigned x = 1, y = (unsigned)n;
        for (unsigned i = (unsigned)std::sqrt((double)n); i > 0; --i)
            if (n % i == 0) { x = i; y = (unsigned)(n / i); break; }
        return diagonals_from_dense({x, y}, shape, v);
    }
    if (shape.size() != 2) throw std::runtime_error("from_numpy: needs a matrix (2 or more dimensions)");
    return diagonals_from_dense({shape[0], shape[1]}, shape, v);
}

static py::array_t<cplx> to_numpy(const HostDia &A) {
    py::array_t<cplx> out({(py::ssize_t)A.matrix_shape[0], (py::ssize_t)A.matrix_shape[1]});
    cplx *o = out.mutable_data();
    std::fill(o, o + (size_t)A.matrix_shape[0] * A.matrix_shape[1], cplx(0, 0));
    for (size_t k = 0; k < A.dIndices.size(); ++k) {
        const int d = A.dIndices[k];
        for (unsigned j = 0; j < A.len(k); ++j) {
            const size_t r = d >= 0 ? j : j - d, c = d >= 0 ? j + d : j;
            o[r * A.matrix_shape[1] + c] = cplx(A.re[A.diagOffsets[k] + j], A.im[A.diagOffsets[k] + j]);
        }
    }
    return out;
}

static std::shared_ptr<HostVector> vector_from(py::array_t<cplx, py::array::c_style | py::array::forcecast> a) {
    auto v = std::make_shared<HostVector>((unsigned)a.size());
    const cplx *p = a.data();
    for (unsigned i = 0; i < v->numRows; ++i) { v->re[i] = p[i].real(); v->im[i] = p[i].imag(); }
    return v;
}

static py::array_t<cplx> vector_to(const HostVector &v) {
    py::array_t<cplx> out(v.numRows);
    cplx *o = out.mutable_data();
    for (unsigned i = 0; i < v.numRows; ++i) o[i] = cplx(v.re[i], v.im[i]);
    return out;
}

// ---------------------------------------------------------------------------------------------------------------------
// Operations
// ---------------------------------------------------------------------------------------------------------------------
static std::shared_ptr<HostVector> spmv_dev(const DeviceMatrix &A, const HostVector &x) {
    auto ms = A.matrix_shape();
    if (x.numRows != ms[1]) throw std::runtime_error("spmv_cuda: dimension mismatch (x length vs A.cols)");
    auto y = std::make_shared<HostVector>(ms[0]);
    check(dq_spmv_host(A.h.get(), x.re.data(), x.im.data(), y->re.data(), y->im.data()));
    return y;
}

static void spmv_into(const HostDia &A, const HostVector &x, HostVector &y) {
    if (x.numRows != A.matrix_shape[1]) throw std::runtime_error("spmv_cuda: dimension mismatch (x length vs A.cols)");
    if (y.numRows != A.matrix_shape[0]) throw std::runtime_error("spmv_cuda: dimension mismatch (y length vs A.rows)");
    auto D = upload(A);  // a host matrix is uploaded for this call, as in the reference (src/spMV_gpu.cu:187-192)
    check(dq_spmv_host(D->h.get(), x.re.data(), x.im.data(), y.re.data(), y.im.data()));
}

static void spmv_ntimes_into(const HostDia &A, const HostVector &x, HostVector &y, int n) {
    if (n < 0) throw std::runtime_error("spmv_ntimes_cuda: num_times must be >= 0");
    if (x.numRows != A.matrix_shape[1])
        throw std::runtime_error("spmv_ntimes_cuda: dimension mismatch (x length vs A.cols)");
    if (y.numRows != A.matrix_shape[0])
        throw std::runtime_error("spmv_ntimes_cuda: dimension mismatch (y length vs A.rows)");
    auto D = upload(A);
    check(dq_spmv_ntimes_host(D->h.get(), x.re.data(), x.im.data(), y.re.data(), y.im.data(), n));
}

// A planned gate sequence (dq_gate_plan): keeps its gates alive for as long as the plan lives.
struct GatePlan {
    std::shared_ptr<dq_gate_plan> h;
    std::vector<std::shared_ptr<DeviceMatrix>> gates;
    int n_qubits = 0;
    GatePlan(int n, const std::vector<std::pair<std::shared_ptr<DeviceMatrix>, unsigned>> &seq) : n_qubits(n) {
        std::vector<const dq_devmat *> g;
        std::vector<uint32_t> q;
        for (const auto &e : seq) {
            if (!e.first) throw std::invalid_argument("GatePlan: None gate");
            gates.push_back(e.first);
            g.push_back(e.first->h.get());
            q.push_back(e.second);
        }
        dq_gate_plan *p = nullptr;
        check(dq_gate_plan_create(n, static_cast<int>(g.size()), g.data(), q.data(), &p));
        h.reset(p, dq_gate_plan_destroy);
    }
    int passes() const {
        int n = 0;
        check(dq_gate_plan_passes(h.get(), &n));
        return n;
    }
};

// y += (I_{2^before} (x) G (x) I_{2^after}) x on host vectors (src/spKron.cpp:286-335, 433-571): the gate is uploaded
// for the call, like the reference's other host-matrix GPU entry points.
static std::shared_ptr<HostVector> kron_apply_into(unsigned before, const HostDia &G, unsigned after,
                                                   const HostVector &x, std::shared_ptr<HostVector> y) {
    if (G.matrix_shape[0] != G.matrix_shape[1]) throw std::runtime_error("kron times vector: the gate must be square");
    const unsigned long long n = (unsigned long long)G.matrix_shape[0] << before << after;
    if (x.numRows != n || y->numRows != n)
        throw std::runtime_error("kron times vector: dimension mismatch (x and y need 2^before * gate rows * 2^after)");
    auto D = upload(G);
    check(dq_kron_apply_host(D->h.get(), before, after, x.re.data(), x.im.data(), y->re.data(), y->im.data()));
    return y;
}

static std::shared_ptr<DeviceMatrix> product_dev(const DeviceMatrix &A, const DeviceMatrix &B) {
    dq_devmat *p = nullptr;
    check(dq_product_device(A.h.get(), B.h.get(), &p));
    return wrap(p);
}

static std::shared_ptr<HostDia> product_host(const HostDia &A, const HostDia &B) {
    if (A.matrix_shape[1] != B.matrix_shape[0])
        throw std::runtime_error("diaQMatrix product_cuda: dimension mismatch.");
    auto dA = upload(A), dB = upload(B);
    dq_devmat *p = nullptr, *q = nullptr;
    check(dq_product_device(dA->h.get(), dB->h.get(), &p));
    DeviceMatrix P(p);
    check(dq_prune(P.h.get(), &q));  // product() prunes all-zero diagonals (src/spGEMM.cpp:516-534)
    return download(DeviceMatrix(q));
}

// tensordot on device matrices (dq_tensordot); the reference's argument errors are std::invalid_argument, which
// pybind11 raises as ValueError (src/tensordotKernels.cpp:24-61)
static std::shared_ptr<DeviceMatrix> tensordot_dev(const DeviceMatrix &A, const std::vector<unsigned> &ta,
                                                   const DeviceMatrix &B, const std::vector<unsigned> &tb,
                                                   const std::vector<std::vector<unsigned>> &axes) {
    if (axes.size() != 2 || axes[0].size() != axes[1].size())
        throw std::invalid_argument("tensordot: axes must be two lists of the same length");
    std::vector<uint32_t> sa(ta.begin(), ta.end()), sb(tb.begin(), tb.end());
    std::vector<uint32_t> xa(axes[0].begin(), axes[0].end()), xb(axes[1].begin(), axes[1].end());
    std::vector<uint32_t> out_shape(std::max<size_t>(2, sa.size() + sb.size() + 4));
    int out_rank = 0;
    dq_devmat *p = nullptr;
    const int rc = dq_tensordot(A.h.get(), (int)sa.size(), sa.data(), B.h.get(), (int)sb.size(), sb.data(),
                                (int)xa.size(), xa.data(), xb.data(), &p, out_shape.data(), &out_rank);
    if (rc == DQ_ERR_ARG) throw std::invalid_argument(dq_last_error());
    check(rc);
    return wrap(p, std::vector<unsigned>(out_shape.begin(), out_shape.begin() + out_rank));
}

static std::shared_ptr<HostDia> tensordot_host(const HostDia &A, const HostDia &B,
                                               const std::vector<std::vector<unsigned>> &axes) {
    auto dA = upload(A), dB = upload(B);
    return download(*tensordot_dev(*dA, A.shape, *dB, B.shape, axes));
}

static std::shared_ptr<HostDia> pow_host(const HostDia &A, long long n) {
    if (n < 0) throw std::runtime_error("pow_cuda: negative exponent not supported");
    auto D = upload(A);
    dq_devmat *p = nullptr;
    check(dq_pow(D->h.get(), n, &p));
    return download(DeviceMatrix(p));
}

static std::shared_ptr<HostDia> expm_host(const HostDia &A, int max_it, double cutoff, double tol) {
    auto D = upload(A);
    dq_devmat *p = nullptr;
    check(dq_expm_taylor_scaled(D->h.get(), max_it, cutoff, tol, &p));
    return download(DeviceMatrix(p));
}

static std::shared_ptr<HostDia> expm_negative_host(const HostDia &A, double t, int max_it, double cutoff,
                                                   double tol) {
    auto D = upload(A);
    dq_devmat *p = nullptr;
    check(dq_expm_negative_taylor_scaled(D->h.get(), t, max_it, cutoff, tol, &p, nullptr, nullptr));
    return download(DeviceMatrix(p));
}

static cplx trace_dev(const DeviceMatrix &D) {
    double r = 0, i = 0;
    check(dq_trace(D.h.get(), &r, &i));
    return cplx(r, i);
}

static cplx trace_host(const HostDia &A) {
    const int k = A.slot(0);
    if (k < 0) return cplx(0, 0);
    HostDia main({A.matrix_shape[0], A.matrix_shape[1]}, {0}, A.shape);  // only the main diagonal goes up
    std::copy(A.re.begin() + A.diagOffsets[k], A.re.begin() + A.diagOffsets[k] + A.len(k), main.re.begin());
    std::copy(A.im.begin() + A.diagOffsets[k], A.im.begin() + A.diagOffsets[k] + A.len(k), main.im.begin());
    return trace_dev(*upload(main));
}

// ---------------------------------------------------------------------------------------------------------------------
PYBIND11_MODULE(diaq_gpu, m) {
    m.doc() = "DiaQ GPU path on MI355X (gfx950): diagonal-format sparse linear algebra in HIP";
    m.def("version", []() { return std::string(dq_version()); }, "the library version");

    py::class_<HostDia, std::shared_ptr<HostDia>>(m, "diaQMatrix")
        .def(py::init([](unsigned s) { return std::make_shared<HostDia>(std::array<unsigned, 2>{s, s},
                                                                         std::vector<int>{}, std::vector<unsigned>{s, s}); }))
        .def(py::init([](unsigned s, unsigned nd) {
            return std::make_shared<HostDia>(std::array<unsigned, 2>{s, s}, symmetric_order(nd),
                                             std::vector<unsigned>{s, s});
        }))
        .def(py::init([](std::array<unsigned, 2> ms) {
            return std::make_shared<HostDia>(ms, std::vector<int>{}, std::vector<unsigned>{ms[0], ms[1]});
        }))
        .def(py::init([](std::array<unsigned, 2> ms, const std::vector<int> &d) {
            return std::make_shared<HostDia>(ms, d, std::vector<unsigned>{ms[0], ms[1]});
        }))
        .def(py::init([](std::array<unsigned, 2> ms, const std::vector<int> &d, std::vector<unsigned> ts) {
            return std::make_shared<HostDia>(ms, d, ts);
        }))
        .def_readonly("matrix_shape", &HostDia::matrix_shape)
        .def_readonly("shape", &HostDia::shape)
        .def_readonly("dIndices", &HostDia::dIndices)
        .def_readonly("diagOffsets", &HostDia::diagOffsets)
        .def_readonly("totalLength", &HostDia::totalLength)
        .def("num_diags", [](const HostDia &A) { return (unsigned)A.dIndices.size(); })
        .def("get_diagonals", [](const HostDia &A) { return A.dIndices; })
        .def("copy", &HostDia::copy)
        .def("set_at_diag", [](HostDia &A, int d, unsigned pos, double r, double i) {
            const int k = A.slot(d);
            if (k < 0) throw std::runtime_error("set_diag: diagonal index not present in dIndices.");
            A.re[A.diagOffsets[k] + pos] = r;
            A.im[A.diagOffsets[k] + pos] = i;
        })
        .def("getItem", [](const HostDia &A, std::array<unsigned, 2> rc) {
            const int k = A.slot((int)rc[1] - (int)rc[0]);
            if (k < 0) return cplx(0, 0);
            const unsigned p = (int)rc[1] - (int)rc[0] >= 0 ? rc[0] : rc[1];
            return cplx(A.re[A.diagOffsets[k] + p], A.im[A.diagOffsets[k] + p]);
        })
        .def("getMatrixPosition", [](const HostDia &A, int d, unsigned p) {
            std::array<unsigned, 2> rc = d >= 0 ? std::array<unsigned, 2>{p, (unsigned)d + p}
                                                : std::array<unsigned, 2>{(unsigned)(-d) + p, p};
            if (rc[0] >= A.matrix_shape[0] || rc[1] >= A.matrix_shape[1])
                throw std::length_error("out of matrix bounds");
            return rc;
        }, py::arg("diagonal_idx"), py::arg("diagonal_pos"))
        .def("multiplyScalar", [](HostDia &A, double r, double i) {
            for (unsigned t = 0; t < A.totalLength; ++t) {
                const double a = A.re[t], b = A.im[t];
                A.re[t] = a * r - b * i;
                A.im[t] = a * i + b * r;
            }
        }, py::arg("real_scale_factor"), py::arg("imag_scale_factor"))
        .def("pruneMatrix", [](HostDia &A) {
            std::vector<int> keep;
            for (size_t k = 0; k < A.dIndices.size(); ++k) {
                bool nz = false;
                for (unsigned j = 0; j < A.len(k) && !nz; ++j)
                    nz = A.re[A.diagOffsets[k] + j] != 0.0 || A.im[A.diagOffsets[k] + j] != 0.0;
                if (nz) keep.push_back((int)k);
            }
            std::vector<int> d;
            for (int k : keep) d.push_back(A.dIndices[k]);
            HostDia P(A.matrix_shape, d, A.shape);
            for (size_t j = 0; j < keep.size(); ++j)
                for (unsigned t = 0; t < P.len(j); ++t) {
                    P.re[P.diagOffsets[j] + t] = A.re[A.diagOffsets[keep[j]] + t];
                    P.im[P.diagOffsets[j] + t] = A.im[A.diagOffsets[keep[j]] + t];
                }
            A = P;
        })
        .def("get_diagonal", [](const HostDia &A, int d) {
            const int k = A.slot(d);
            if (k < 0) throw std::runtime_error("diagonal index not present in dIndices.");
            py::array_t<cplx> out(A.len(k));
            cplx *o = out.mutable_data();
            for (unsigned j = 0; j < A.len(k); ++j) o[j] = cplx(A.re[A.diagOffsets[k] + j], A.im[A.diagOffsets[k] + j]);
            return out;
        }, "values of diagonal d (position order) as a complex128 array")
        .def_static("from_diagonals", [](std::array<unsigned, 2> ms, const std::vector<int> &offs,
                                         const std::vector<py::array_t<cplx, py::array::c_style | py::array::forcecast>> &vals) {
            if (offs.size() != vals.size()) throw std::runtime_error("from_diagonals: one value array per offset");
            auto M = std::make_shared<HostDia>(ms, offs, std::vector<unsigned>{ms[0], ms[1]});
            for (size_t k = 0; k < offs.size(); ++k) {
                if ((size_t)vals[k].size() != M->len(k))
                    throw std::runtime_error("from_diagonals: diagonal " + std::to_string(offs[k]) + " needs " +
                                             std::to_string(M->len(k)) + " values");
                const cplx *p = vals[k].data();
                for (unsigned j = 0; j < M->len(k); ++j) {
                    M->re[M->diagOffsets[k] + j] = p[j].real();
                    M->im[M->diagOffsets[k] + j] = p[j].imag();
                }
            }
            return M;
        }, "bulk constructor from per-diagonal complex arrays (no reference counterpart)")
        .def("print_meta", [](const HostDia &A) {
            std::cout << "matrix_shape (" << A.matrix_shape[0] << "," << A.matrix_shape[1] << ") diagonals "
                      << A.dIndices.size() << " totalLength " << A.totalLength << std::endl;
        })
        .def("print", [](const HostDia &A) {  // src/diaQMatrix.hpp:749-790
            std::string t;
            for (size_t i = 0; i < A.shape.size(); ++i) t += (i ? "," : "") + std::to_string(A.shape[i]);
            py::print("matrix_shape (" + std::to_string(A.matrix_shape[0]) + "," + std::to_string(A.matrix_shape[1]) +
                      ")");
            py::print("tensor_shape (" + t + ")");
            py::print("#Diagonals = " + std::to_string(A.dIndices.size()));
            for (size_t k = 0; k < A.dIndices.size(); ++k) {
                std::string line = "(dIndex=" + std::to_string(A.dIndices[k]) + ") (dLength=" +
                                   std::to_string(A.len(k)) + ") [";
                for (unsigned j = 0; j < A.len(k); ++j)
                    line += (j ? ",(" : "(") + std::to_string(A.re[A.diagOffsets[k] + j]) + "," +
                            std::to_string(A.im[A.diagOffsets[k] + j]) + ")";
                py::print(line + "]");
            }
        })
        .def("sparsity", [](const HostDia &A) {  // src/diaQMatrix.hpp:1361-1375: 1 - stored nonzeros / size
            size_t nnz = 0;
            for (size_t t = 0; t < A.re.size(); ++t) nnz += (A.re[t] != 0.0 || A.im[t] != 0.0);
            const double size = static_cast<double>(A.matrix_shape[0]) * A.matrix_shape[1];
            return 1.0 - static_cast<double>(nnz) / size;
        })
        .def("convertToDenseMatrix", [](const HostDia &A) {  // src/conversions.cpp:11-48
            const size_t n = static_cast<size_t>(A.matrix_shape[0]) * A.matrix_shape[1];
            auto D = std::make_shared<HostDense>(0u, std::vector<double>{}, std::vector<double>{});
            D->matrix_shape = A.matrix_shape;
            D->shape = A.shape;
            D->re.assign(n, 0.0);
            D->im.assign(n, 0.0);
            for (size_t k = 0; k < A.dIndices.size(); ++k) {
                const int d = A.dIndices[k];
                for (unsigned j = 0; j < A.len(k); ++j) {
                    const size_t r = d >= 0 ? j : j - d, c = d >= 0 ? j + d : j;
                    D->re[r * A.matrix_shape[1] + c] = A.re[A.diagOffsets[k] + j];
                    D->im[r * A.matrix_shape[1] + c] = A.im[A.diagOffsets[k] + j];
                }
            }
            return D;
        })
        .def("addMatrixScaled", [](HostDia &A, std::shared_ptr<HostDia> B, double rs, double is) {
            // A = A + (rs + i is) B over the ascending union of their diagonals (src/diaQMatrix.hpp:1061-1142)
            if (A.matrix_shape != B->matrix_shape) throw std::runtime_error("addMatrixScaled: matrix shapes do not match");
            std::vector<int> u(A.dIndices);
            u.insert(u.end(), B->dIndices.begin(), B->dIndices.end());
            std::sort(u.begin(), u.end());
            u.erase(std::unique(u.begin(), u.end()), u.end());
            HostDia C(A.matrix_shape, u, A.shape);
            for (size_t k = 0; k < u.size(); ++k) {
                const int sa = A.slot(u[k]), sb = B->slot(u[k]);
                for (unsigned j = 0; j < C.len(k); ++j) {
                    const double ar = sa >= 0 ? A.re[A.diagOffsets[sa] + j] : 0.0;
                    const double ai = sa >= 0 ? A.im[A.diagOffsets[sa] + j] : 0.0;
                    const double br = sb >= 0 ? B->re[B->diagOffsets[sb] + j] : 0.0;
                    const double bi = sb >= 0 ? B->im[B->diagOffsets[sb] + j] : 0.0;
                    C.re[C.diagOffsets[k] + j] = ar + (rs * br - is * bi);
                    C.im[C.diagOffsets[k] + j] = ai + (rs * bi + is * br);
                }
            }
            A = C;
        }, py::arg("B"), py::arg("real_scale_factor"), py::arg("imag_scale_factor"));

    py::class_<HostIdentity, HostDia, std::shared_ptr<HostIdentity>>(m, "identityMatrix").def(py::init<unsigned>());
    py::class_<HostPermutation, HostDia, std::shared_ptr<HostPermutation>>(m, "permutationMatrix")
        .def(py::init<unsigned, const std::vector<unsigned> &>());

    auto dense_values = [](py::array_t<double, py::array::c_style | py::array::forcecast> a) {
        return std::vector<double>(a.data(), a.data() + a.size());
    };
    auto dense_to_dia = [](const HostDense &D, std::array<unsigned, 2> ms) {
        const size_t n = D.re.size();
        if (static_cast<size_t>(ms[0]) * ms[1] != n)
            throw std::invalid_argument("convertToDiaQ: the new shape must hold the same number of entries");
        std::vector<cplx> v(n);
        for (size_t i = 0; i < n; ++i) v[i] = cplx(D.re[i], D.im[i]);
        return diagonals_from_dense(ms, D.shape, v.data());
    };
    py::class_<HostDense, std::shared_ptr<HostDense>>(m, "denseMatrix")
        .def(py::init([dense_values](unsigned s, py::array_t<double, py::array::c_style | py::array::forcecast> r) {
            return std::make_shared<HostDense>(s, dense_values(r), std::vector<double>(r.size(), 0.0));
        }))
        .def(py::init([dense_values](unsigned s, py::array_t<double, py::array::c_style | py::array::forcecast> r,
                                     py::array_t<double, py::array::c_style | py::array::forcecast> i) {
            return std::make_shared<HostDense>(s, dense_values(r), dense_values(i));
        }))
        .def_readwrite("matrix_shape", &HostDense::matrix_shape)
        .def_readwrite("shape", &HostDense::shape)
        .def("getItem", [](const HostDense &D, const std::vector<unsigned> &pos) {  // getFlatIndex, denseMatrix.hpp
            if (pos.size() != D.shape.size()) throw std::invalid_argument("getItem: one index per tensor axis");
            unsigned long long flat = 0, prod = 1;
            const unsigned long long size = static_cast<unsigned long long>(D.matrix_shape[0]) * D.matrix_shape[1];
            for (size_t i = 0; i < D.shape.size(); ++i) {
                prod *= D.shape[i];
                flat += (size / prod) * pos[i];
            }
            if (flat >= D.re.size()) throw std::out_of_range("getItem: index outside the matrix");
            return cplx(D.re[flat], D.im[flat]);
        })
        .def("set", [](HostDense &D, unsigned i, unsigned j, cplx v) {
            const size_t f = static_cast<size_t>(D.matrix_shape[1]) * i + j;
            D.re.at(f) = v.real();
            D.im.at(f) = v.imag();
        })
        .def("get", [](const HostDense &D, unsigned i, unsigned j) {
            const size_t f = static_cast<size_t>(D.matrix_shape[1]) * i + j;
            return cplx(D.re.at(f), D.im.at(f));
        })
        .def("print", [](const HostDense &D) {
            py::print("matrix_shape (" + std::to_string(D.matrix_shape[0]) + "," + std::to_string(D.matrix_shape[1]) +
                      ")");
            for (unsigned i = 0; i < D.matrix_shape[0]; ++i) {
                std::string line;
                for (unsigned j = 0; j < D.matrix_shape[1]; ++j) {
                    const size_t f = static_cast<size_t>(D.matrix_shape[1]) * i + j;
                    line += (j ? "," : "") + std::string("(") + std::to_string(D.re[f]) + "," +
                            std::to_string(D.im[f]) + ")";
                }
                py::print(line);
            }
        })
        .def("isPartOf", [](const HostDense &D, int dIndex, int a, int b, int x, int y) {
            const long long size = static_cast<long long>(D.matrix_shape[0]) * D.matrix_shape[1];
            const int p = dIndex >= size ? x + y - a - b : x - y + b - a;
            return p == 0;
        })
        .def("sparsity", [](const HostDense &D) {
            size_t z = 0;
            for (size_t i = 0; i < D.re.size(); ++i) z += (D.re[i] == 0.0 && D.im[i] == 0.0);
            return D.re.empty() ? 0.0 : 100.0 * static_cast<double>(z) / static_cast<double>(D.re.size());
        })
        .def("convertToDiaQ", [dense_to_dia](const HostDense &D, std::vector<unsigned> s) {
            return dense_to_dia(D, pair_of(s));
        }, py::arg())
        .def("convertToDiaQ", [dense_to_dia](const HostDense &D) { return dense_to_dia(D, D.matrix_shape); });

    py::class_<HostVector, std::shared_ptr<HostVector>>(m, "vector")
        .def(py::init<unsigned>())
        .def_readonly("numRows", &HostVector::numRows)
        .def("setValue", [](HostVector &v, unsigned i, double r, double im) { v.re.at(i) = r; v.im.at(i) = im; })
        .def("getArray", [](const HostVector &v) {
            std::vector<cplx> out(v.numRows);
            for (unsigned i = 0; i < v.numRows; ++i) out[i] = cplx(v.re[i], v.im[i]);
            return out;
        })
        .def("copy", [](const HostVector &v) { return std::make_shared<HostVector>(v); })
        .def("isEqualTo", [](const HostVector &v, std::shared_ptr<HostVector> w, double tol) {  // src/vector.hpp:119-131
            if (v.numRows != w->numRows) return false;
            for (unsigned i = 0; i < v.numRows; ++i)
                if (std::abs(v.re[i] - w->re[i]) > tol || std::abs(v.im[i] - w->im[i]) > tol) return false;
            return true;
        }, py::arg("v"), py::arg("tol") = 1e-6)
        .def("print", [](const HostVector &v) {
            std::string line;
            for (unsigned i = 0; i < v.numRows; ++i)
                line += (i ? ",(" : "(") + std::to_string(v.re[i]) + "," + std::to_string(v.im[i]) + ")";
            py::print(line);
        })
        .def("norm", [](const HostVector &v) {
            double s = 0;
            for (unsigned i = 0; i < v.numRows; ++i) s += v.re[i] * v.re[i] + v.im[i] * v.im[i];
            return std::sqrt(s);
        });

    // conversions
    m.def("from_numpy", [](py::array_t<cplx, py::array::c_style | py::array::forcecast> vals, std::vector<unsigned> shape) {
        if (vals.ndim() != 1) throw std::runtime_error("Input arrays must be 1-dimensional.");
        return from_numpy_shape(vals.data(), shape);
    });
    m.def("from_numpy", [](py::array_t<cplx, py::array::c_style | py::array::forcecast> vals) {
        std::vector<unsigned> shape;
        for (py::ssize_t i = 0; i < vals.ndim(); ++i) shape.push_back((unsigned)vals.shape(i));
        return from_numpy_shape(vals.data(), shape);
    });
    m.def("to_numpy", [](std::shared_ptr<HostDia> A) { return to_numpy(*A); });
    m.def("from_numpy_vector", &vector_from, py::arg());
    m.def("to_numpy_vector", [](std::shared_ptr<HostVector> v) { return vector_to(*v); }, py::arg());

    m.def("dia_viability", [](int64_t n_rows, int64_t n_cols, int64_t n_diags, int64_t nnz, int64_t n_right,
                              int itemsize, int index_bytes) {
        dq_dia_viability_t v;
        check(dq_dia_viability(n_rows, n_cols, n_diags, nnz, n_right, itemsize, index_bytes, &v));
        py::dict d;
        d["viable"] = (bool)v.viable;
        d["storage_ok"] = (bool)v.storage_ok;
        d["diagonals_ok"] = (bool)v.diagonals_ok;
        d["dia_bytes"] = v.dia_bytes;
        d["csr_bytes"] = v.csr_bytes;
        d["density"] = v.density;
        d["diagonal_budget"] = v.diagonal_budget;
        d["reason"] = std::string(v.reason);
        return d;
    }, py::arg("n_rows"), py::arg("n_cols"), py::arg("n_diags"), py::arg("nnz"), py::arg("n_right"),
       py::arg("itemsize") = 16, py::arg("index_bytes") = 8);

    // kernels on host matrices (each call uploads, as the reference's host API does)
    m.def("dia_viability_mi355x", [](int64_t n_rows, int64_t n_cols, int64_t n_diags, int64_t nnz, int64_t n_right,
                              int itemsize, int index_bytes) {
        dq_dia_viability_t v;
        check(dq_dia_viability_mi355x(n_rows, n_cols, n_diags, nnz, n_right, itemsize, index_bytes, &v));
        py::dict d;
        d["viable"] = (bool)v.viable;
        d["storage_ok"] = (bool)v.storage_ok;
        d["diagonals_ok"] = (bool)v.diagonals_ok;
        d["dia_bytes"] = v.dia_bytes;
        d["csr_bytes"] = v.csr_bytes;
        d["density"] = v.density;
        d["diagonal_budget"] = v.diagonal_budget;
        d["reason"] = std::string(v.reason);
        return d;
    }, py::arg("n_rows"), py::arg("n_cols"), py::arg("n_diags"), py::arg("nnz"), py::arg("n_right"),
       py::arg("itemsize") = 16, py::arg("index_bytes") = 8);

    // kernels on host matrices (each call uploads, as the reference's host API does)
    m.def("product", [](std::shared_ptr<HostDia> A, std::shared_ptr<HostDia> B) { return product_host(*A, *B); });
    m.def("trace", [](std::shared_ptr<HostDia> A) { return trace_host(*A); });
    // tensordot and the reference's numpy overloads for quimb (include/pythonWrappers.cpp:270-330); the contraction
    // runs on the GPU
    using Axes = std::vector<std::vector<unsigned>>;
    m.def("tensordot", [](std::shared_ptr<HostDia> A, std::shared_ptr<HostDia> B, Axes axes) {
        return tensordot_host(*A, *B, axes);
    }, py::arg(), py::arg(), py::arg("axes") = Axes{});
    m.def("tensordot", [](py::array_t<cplx, py::array::c_style | py::array::forcecast> a,
                          py::array_t<cplx, py::array::c_style | py::array::forcecast> b, Axes axes) {
        std::vector<unsigned> sa(a.shape(), a.shape() + a.ndim()), sb(b.shape(), b.shape() + b.ndim());
        return tensordot_host(*from_numpy_shape(a.data(), sa), *from_numpy_shape(b.data(), sb), axes);
    }, py::arg("arr1"), py::arg("arr2"), py::arg("axes") = Axes{});
    m.def("tensordot", [](std::shared_ptr<HostDia> A, py::array_t<cplx, py::array::c_style | py::array::forcecast> b,
                          Axes axes) {
        std::vector<unsigned> sb(b.shape(), b.shape() + b.ndim());
        return tensordot_host(*A, *from_numpy_shape(b.data(), sb), axes);
    }, py::arg("mat"), py::arg("arr"), py::arg("axes") = Axes{});
    m.def("tensordot", [](py::array_t<cplx, py::array::c_style | py::array::forcecast> a, std::shared_ptr<HostDia> B,
                          Axes axes) {
        std::vector<unsigned> sa(a.shape(), a.shape() + a.ndim());
        return tensordot_host(*from_numpy_shape(a.data(), sa), *B, axes);
    }, py::arg("arr"), py::arg("mat"), py::arg("axes") = Axes{});
    m.def("spMV", [](std::shared_ptr<HostDia> A, std::shared_ptr<HostVector> x) {
        auto y = std::make_shared<HostVector>(A->matrix_shape[0]);
        spmv_into(*A, *x, *y);
        return y;
    }, py::arg(), py::arg());
    m.def("spMV", [](std::shared_ptr<HostDia> A, std::shared_ptr<HostVector> x, std::shared_ptr<HostVector> y) {
        spmv_into(*A, *x, *y);
        return y;
    }, py::arg(), py::arg(), py::arg());
    m.def("spMV_ntimes", [](std::shared_ptr<HostDia> A, std::shared_ptr<HostVector> x, int n) {
        auto y = std::make_shared<HostVector>(A->matrix_shape[0]);
        spmv_ntimes_into(*A, *x, *y, n);
        return y;
    }, py::arg(), py::arg(), py::arg());
    m.def("spMV_ntimes", [](std::shared_ptr<HostDia> A, std::shared_ptr<HostVector> x, std::shared_ptr<HostVector> y,
                            int n) {
        spmv_ntimes_into(*A, *x, *y, n);
        return y;
    }, py::arg(), py::arg(), py::arg(), py::arg());
    m.def("identity_kron_matrix_kron_identity_times_vector",
          [](unsigned before, std::shared_ptr<HostDia> gate, unsigned after, std::shared_ptr<HostVector> x,
             std::shared_ptr<HostVector> y) { return kron_apply_into(before, *gate, after, *x, y); },
          py::arg("before"), py::arg("gate"), py::arg("after"), py::arg("x"), py::arg("y"));
    m.def("identity_kron_matrix_times_vector",
          [](unsigned n, std::shared_ptr<HostDia> A, std::shared_ptr<HostVector> x, std::shared_ptr<HostVector> y) {
              return kron_apply_into(n, *A, 0, *x, y);
          }, py::arg("n"), py::arg("A"), py::arg("x"), py::arg("y"));
    m.def("matrix_kron_identity_times_vector",
          [](std::shared_ptr<HostDia> A, unsigned n, std::shared_ptr<HostVector> x, std::shared_ptr<HostVector> y) {
              return kron_apply_into(0, *A, n, *x, y);
          }, py::arg("A"), py::arg("n"), py::arg("x"), py::arg("y"));
    m.def("spmv_ntimes_device", [](std::shared_ptr<DeviceMatrix> A, std::uintptr_t xr, std::uintptr_t xi,
                                   std::uintptr_t yr, std::uintptr_t yi, int num_times, std::uintptr_t work_r,
                                   std::uintptr_t work_i, std::uintptr_t stream) {
        check(dq_spmv_ntimes_device(A->h.get(), reinterpret_cast<const void *>(xr), reinterpret_cast<const void *>(xi),
                                    reinterpret_cast<void *>(yr), reinterpret_cast<void *>(yi), num_times,
                                    reinterpret_cast<void *>(work_r), reinterpret_cast<void *>(work_i), stream));
    }, py::arg("A"), py::arg("xr"), py::arg("xi"), py::arg("yr"), py::arg("yi"), py::arg("num_times"),
       py::arg("work_r") = 0, py::arg("work_i") = 0, py::arg("stream") = 0,
       "y = A^num_times x on device vectors (x == y allowed with work vectors); spMV_ntimes on resident data");
    m.def("kron_apply_device", [](std::shared_ptr<DeviceMatrix> G, unsigned before, unsigned after, std::uintptr_t xr,
                                  std::uintptr_t xi, std::uintptr_t yr, std::uintptr_t yi, bool accumulate,
                                  std::uintptr_t stream) {
        check(dq_kron_apply_device(G->h.get(), before, after, reinterpret_cast<const void *>(xr),
                                   reinterpret_cast<const void *>(xi), reinterpret_cast<void *>(yr),
                                   reinterpret_cast<void *>(yi), accumulate ? 1 : 0, stream));
    }, py::arg("G"), py::arg("before"), py::arg("after"), py::arg("xr"), py::arg("xi"), py::arg("yr"), py::arg("yi"),
       py::arg("accumulate") = true, py::arg("stream") = 0,
       "y (+)= (I_{2^before} (x) G (x) I_{2^after}) x on device vectors (split re/im, G's precision); x and y may be "
       "the same vectors for gates up to 16 x 16 (no reference counterpart for device vectors)");
    m.def("pow", [](std::shared_ptr<HostDia> A, long long n) { return pow_host(*A, n); }, py::arg(), py::arg());
    m.def("expm", [](std::shared_ptr<HostDia> A, int max_iterations, double scale_cutoff, double tolerance) {
        return expm_host(*A, max_iterations, scale_cutoff, tolerance);
    }, py::arg("A"), py::arg("max_iterations") = 100, py::arg("scale_cutoff") = 1.0, py::arg("tolerance") = 1e-12);
    m.def("expm_negative", [](std::shared_ptr<HostDia> A, double t, int max_iterations, double scale_cutoff,
                              double tolerance) {
        return expm_negative_host(*A, t, max_iterations, scale_cutoff, tolerance);
    }, py::arg("A"), py::arg("t"), py::arg("max_iterations") = 100, py::arg("scale_cutoff") = 1.0,
       py::arg("tolerance") = 1e-12);

    py::class_<GatePlan, std::shared_ptr<GatePlan>>(m, "GatePlan",
        "A gate sequence on an n-qubit state, planned into fused passes: GatePlan(n_qubits, [(gate, first_qubit), ...]) "
        "with 2^k x 2^k DeviceMatrix gates (k <= 4) acting on qubits first_qubit .. first_qubit + k - 1 "
        "(no reference counterpart)")
        .def(py::init<int, const std::vector<std::pair<std::shared_ptr<DeviceMatrix>, unsigned>> &>(),
             py::arg("n_qubits"), py::arg("gates"))
        .def_property_readonly("n_passes", &GatePlan::passes)
        .def_readonly("n_qubits", &GatePlan::n_qubits)
        .def("run", [](const GatePlan &P, std::uintptr_t re, std::uintptr_t im, int repetitions, std::uintptr_t stream) {
            check(dq_gate_plan_run(P.h.get(), reinterpret_cast<void *>(re), reinterpret_cast<void *>(im), repetitions,
                                   stream));
        }, py::arg("re"), py::arg("im"), py::arg("repetitions") = 1, py::arg("stream") = 0,
           "state <- (G_last ... G_first)^repetitions state in place on split re/im device vectors");

    // device-resident matrices
    py::class_<DeviceMatrix, std::shared_ptr<DeviceMatrix>>(m, "DeviceMatrix",
        "A DIA matrix resident in HBM. The matrix is uploaded once; spMV on it moves only the vectors, products "
        "of two of them stay on the device, and to_host() brings a copy back as a diaQMatrix.")
        .def(py::init([](std::shared_ptr<HostDia> A, const std::string &precision) {
            return upload(*A, parse_precision(precision));
        }), py::arg("A"), py::arg("precision") = "fp64")
        .def("to_host", [](const DeviceMatrix &D) { return download(D); })
        .def("copy", [](const DeviceMatrix &D) {
            dq_devmat *p = nullptr;
            check(dq_devmat_copy(D.h.get(), &p));
            return wrap(p, D.tensor_shape);
        })
        .def("astype", [](const DeviceMatrix &D, const std::string &precision) {
            dq_devmat *p = nullptr;
            check(dq_devmat_convert(D.h.get(), parse_precision(precision), &p));
            return wrap(p, D.tensor_shape);
        }, py::arg("precision"), "a copy in 'fp64' or 'fp32' (no reference counterpart)")
        .def("get_diagonal", [](const DeviceMatrix &D, int offset) {
            auto ms = D.matrix_shape();
            const unsigned n = diag_len(offset, ms);
            std::vector<double> r(n), i(n);
            check(dq_devmat_get_diagonal(D.h.get(), offset, r.data(), i.data()));
            py::array_t<cplx> out(n);
            cplx *o = out.mutable_data();
            for (unsigned j = 0; j < n; ++j) o[j] = cplx(r[j], i[j]);
            return out;
        }, py::arg("offset"), "values of one diagonal (position order), downloaded (no reference counterpart)")
        .def_property_readonly("shape", &DeviceMatrix::matrix_shape)
        .def_property_readonly("num_diags", &DeviceMatrix::n_diags)
        .def_property_readonly("diagonals", &DeviceMatrix::diagonals)
        .def_property_readonly("precision", [](const DeviceMatrix &D) {
            return std::string(D.precision() == DQ_F32 ? "fp32" : "fp64");
        })
        .def_property_readonly("stored", [](const DeviceMatrix &D) {
            int64_t s;
            check(dq_devmat_info(D.h.get(), nullptr, nullptr, nullptr, nullptr, &s));
            return s;
        })
        .def_property_readonly("handle", [](const DeviceMatrix &D) { return reinterpret_cast<uintptr_t>(D.h.get()); },
                               "the dq_devmat pointer for the C ABI");
    m.def("to_device", [](std::shared_ptr<HostDia> A, const std::string &precision) {
        return upload(*A, parse_precision(precision));
    }, py::arg("A"), py::arg("precision") = "fp64", "Copy a host diaQMatrix into HBM and return the DeviceMatrix holding it.");
    m.def("spMV", [](std::shared_ptr<DeviceMatrix> A, std::shared_ptr<HostVector> x) { return spmv_dev(*A, *x); },
          py::arg(), py::arg());
    m.def("pauli_sum_device", [](int n_qubits, std::vector<uint64_t> x_masks, std::vector<uint64_t> z_masks,
                                 std::vector<std::complex<double>> coefs, const std::string &precision,
                                 std::uintptr_t stream) {
        if (x_masks.size() != z_masks.size() || x_masks.size() != coefs.size())
            throw std::invalid_argument("pauli_sum_device: x_masks, z_masks and coefs must have the same length");
        std::vector<double> cr(coefs.size()), ci(coefs.size());
        for (size_t j = 0; j < coefs.size(); ++j) { cr[j] = coefs[j].real(); ci[j] = coefs[j].imag(); }
        dq_devmat *p = nullptr;
        check(dq_devmat_from_pauli_sum(n_qubits, static_cast<int>(x_masks.size()), x_masks.data(), z_masks.data(),
                                       cr.data(), ci.data(), parse_precision(precision), stream, &p));
        return wrap(p);
    }, py::arg("n_qubits"), py::arg("x_masks"), py::arg("z_masks"), py::arg("coefs"), py::arg("precision") = "fp64",
       py::arg("stream") = 0,
       "H = sum_j coefs[j] P_j built on the device (a DeviceMatrix); term j flips the qubits of x_masks[j] (X or Y) "
       "and has Z or Y on those of z_masks[j]; qubit k is bit k of the basis index (no reference counterpart)");
    m.def("from_dense_device", [](std::uintptr_t dense, uint32_t rows, uint32_t cols, int64_t ld,
                                  const std::string &precision, std::uintptr_t stream) {
        dq_devmat *p = nullptr;
        check(dq_devmat_from_dense_device(rows, cols, reinterpret_cast<const double *>(dense), ld < 0 ? cols : ld,
                                          parse_precision(precision), stream, &p));
        return wrap(p);
    }, py::arg("dense"), py::arg("rows"), py::arg("cols"), py::arg("ld") = -1, py::arg("precision") = "fp64",
       py::arg("stream") = 0,
       "from_numpy on the device: the nonzero diagonals (ascending) of a row-major complex128 matrix at device "
       "address dense (row stride ld, default cols), as a DeviceMatrix");
    m.def("to_dense_device", [](std::shared_ptr<DeviceMatrix> A, std::uintptr_t dense, int64_t ld,
                                std::uintptr_t stream) {
        check(dq_devmat_to_dense_device(A->h.get(), reinterpret_cast<double *>(dense),
                                        ld < 0 ? static_cast<int64_t>(A->matrix_shape()[1]) : ld, stream));
    }, py::arg("A"), py::arg("dense"), py::arg("ld") = -1, py::arg("stream") = 0,
       "to_numpy on the device: write A as a row-major complex128 matrix at device address dense (row stride ld)");
    m.def("product", [](std::shared_ptr<DeviceMatrix> A, std::shared_ptr<DeviceMatrix> B) {
        return product_dev(*A, *B);
    }, py::arg(), py::arg());
    m.def("trace", [](std::shared_ptr<DeviceMatrix> A) { return trace_dev(*A); });
    m.def("tensordot", [](std::shared_ptr<DeviceMatrix> A, std::shared_ptr<DeviceMatrix> B, Axes axes) {
        return tensordot_dev(*A, A->tensor_shape, *B, B->tensor_shape, axes);
    }, py::arg(), py::arg(), py::arg("axes") = Axes{},
       "np.tensordot of two device-resident tensors; the result stays on the device with its tensor shape");
    // pow / expm / expm_negative on a DeviceMatrix: the result stays on the device (the host-matrix forms above return
    // a host diaQMatrix, as the reference's do, which for a near-dense result means downloading it)
    m.def("pow", [](std::shared_ptr<DeviceMatrix> A, long long n) {
        dq_devmat *p = nullptr;
        check(dq_pow(A->h.get(), n, &p));
        return wrap(p);
    }, py::arg(), py::arg());
    m.def("expm", [](std::shared_ptr<DeviceMatrix> A, int max_iterations, double scale_cutoff, double tolerance) {
        dq_devmat *p = nullptr;
        check(dq_expm_taylor_scaled(A->h.get(), max_iterations, scale_cutoff, tolerance, &p));
        return wrap(p);
    }, py::arg("A"), py::arg("max_iterations") = 100, py::arg("scale_cutoff") = 1.0, py::arg("tolerance") = 1e-12);
    m.def("expm_negative", [](std::shared_ptr<DeviceMatrix> A, double t, int max_iterations, double scale_cutoff,
                              double tolerance) {
        dq_devmat *p = nullptr;
        check(dq_expm_negative_taylor_scaled(A->h.get(), t, max_iterations, scale_cutoff, tolerance, &p, nullptr,
                                             nullptr));
        return wrap(p);
    }, py::arg("A"), py::arg("t"), py::arg("max_iterations") = 100, py::arg("scale_cutoff") = 1.0,
       py::arg("tolerance") = 1e-12);
    m.def("prune", [](std::shared_ptr<DeviceMatrix> A) {
        dq_devmat *p = nullptr;
        check(dq_prune(A->h.get(), &p));
        return wrap(p);
    }, "drop all-zero diagonals on the device (no reference counterpart for a DeviceMatrix)");
    m.def("estimate_1norm", [](std::shared_ptr<DeviceMatrix> A) {
        double v;
        check(dq_estimate_1norm(A->h.get(), &v));
        return v;
    });
    m.def("norm", [](std::shared_ptr<DeviceMatrix> A) {
        double v;
        check(dq_frobenius_norm(A->h.get(), &v));
        return v;
    });
    m.def("norm_diff", [](std::shared_ptr<DeviceMatrix> A, std::shared_ptr<DeviceMatrix> B) {
        double v;
        check(dq_frobenius_diff(A->h.get(), B->h.get(), &v));
        return v;
    }, "||A - B||_F on the device for matrices with the same diagonals (no reference counterpart)");
    m.def("spmv_device", [](std::shared_ptr<DeviceMatrix> A, uintptr_t xr, uintptr_t xi, uintptr_t yr, uintptr_t yi,
                            uintptr_t stream) {
        check(dq_spmv_device(A->h.get(), (const void *)xr, (const void *)xi, (void *)yr, (void *)yi, stream));
    }, py::arg("A"), py::arg("x_real"), py::arg("x_imag"), py::arg("y_real"), py::arg("y_imag"), py::arg("stream") = 0,
       "Device-pointer SpMV: x and y are the split real and imaginary device arrays (float64; float32 for an fp32 "
       "DeviceMatrix) given as integer addresses, e.g. torch data_ptr(). y is written, not accumulated, and the "
       "call returns before the kernel ends (it runs on `stream`).");
    m.def("spmm_dia_device", [](uintptr_t offsets, uintptr_t vals, int n_diags, int64_t n_rows, int64_t n_cols,
                                uintptr_t B, int64_t n_right, uintptr_t C, bool is_complex, uintptr_t stream) {
        check(dq_spmm_dia_device(offsets, vals, n_diags, n_rows, n_cols, B, n_right, C, is_complex ? 1 : 0, stream));
    }, py::arg("offsets"), py::arg("vals"), py::arg("n_diags"), py::arg("n_rows"), py::arg("n_cols"), py::arg("B"),
       py::arg("n_right"), py::arg("C"), py::arg("is_complex"), py::arg("stream") = 0,
       "Dense product C = A B on device pointers. A is stored row-aligned: vals[k * n_rows + r] holds A[r, r + "
       "offsets[k]] (offsets int64). B has n_cols rows and C n_rows rows, both with n_right columns in row-major "
       "order, real or complex double. C is overwritten.");
    m.def("hamiltonian_simulation_cuda", [](std::shared_ptr<HostDia> H, std::shared_ptr<HostVector> psi0, double dt,
                                            unsigned long r, int max_iterations, double scale_cutoff,
                                            double tolerance) {
        (void)scale_cutoff;  // accepted and ignored, as in the reference (src/spHamSim_gpu_all_fused.cu:293-294)
        (void)tolerance;
        if (H->matrix_shape[0] != H->matrix_shape[1])
            throw std::runtime_error("hamiltonian_simulation_cuda: H must be square");
        if (psi0->numRows != H->matrix_shape[1])
            throw std::runtime_error("hamiltonian_simulation_cuda: psi0 dimension mismatch");
        auto out = std::make_shared<HostVector>(psi0->numRows);
        if (psi0->numRows == 0) return out;
        auto D = upload(*H);
        check(dq_hamsim_host(D->h.get(), psi0->re.data(), psi0->im.data(), dt, r, max_iterations, out->re.data(),
                             out->im.data()));
        return out;
    }, py::arg("H"), py::arg("psi0"), py::arg("dt"), py::arg("r"), py::arg("max_iterations") = 100,
       py::arg("scale_cutoff") = 1.0, py::arg("tolerance") = 1e-12);
    m.def("hamsim_device", [](std::shared_ptr<DeviceMatrix> H, uintptr_t pr, uintptr_t pi, double t, uint64_t r, int m,
                              uintptr_t stream) {
        check(dq_hamsim_device(H->h.get(), (void *)pr, (void *)pi, t, r, m, stream));
    }, py::arg("H"), py::arg("psi_real"), py::arg("psi_imag"), py::arg("t"), py::arg("r"), py::arg("m"),
       py::arg("stream") = 0, "the fused evolution in place on device vectors (no reference counterpart)");
    m.def("hamsim_device_fp64_state", [](std::shared_ptr<DeviceMatrix> H, uintptr_t pr, uintptr_t pi, double t,
                                         uint64_t r, int m, uintptr_t stream) {
        check(dq_hamsim_device_f64state(H->h.get(), (double *)pr, (double *)pi, t, r, m, stream));
    }, py::arg("H"), py::arg("psi_real"), py::arg("psi_imag"), py::arg("t"), py::arg("r"), py::arg("m"),
       py::arg("stream") = 0,
       "the fused evolution in place on FP64 device vectors for either precision of H; an fp32 H runs its Taylor "
       "terms in FP32, sums each step in FP64 and renormalises in FP64 (no reference counterpart)");

    m.def("install_info", []() {
        py::dict info;
        info["VEC_WIDTH"] = (int)VEC_WIDTH;
        info["GPU_ARCH"] = "gfx950";
        info["HIP_ENABLED"] = true;
        info["CUDA_ENABLED"] = false;
        info["valType"] = "d";
        info["library"] = std::string(dq_version());
        return info;
    }, "build configuration as a dict");
}
