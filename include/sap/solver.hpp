// Header-only C++ facade with the reference's interface
// (spikegpu/SaPLibrary sap/solver.h:52-225, sap/banded_matrix.h, sap/spmv.h,
// and the cusp::csr_matrix input of Solver::setup) over the C ABI of
// include/sap_c.h.  Array types are any device array with
// `value_type`, `size()` and `data()` returning a device pointer (for example
// sap::DeviceArray below); the reference used cusp::array1d<..., device_memory>.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "../sap_c.h"

namespace sap {

enum KrylovSolverType { BiCGStab_C, GMRES_C, CG_C, CR_C, BiCGStab1, BiCGStab2, BiCGStab, MINRES };
enum FactorizationMethod { LU_UL, LU_only };
enum PreconditionerType { Spike, Block, None };

// sap::system_error (sap/exception.h:13-43)
class system_error : public std::runtime_error {
public:
    enum Reason { Zero_pivoting = -1, Negative_DB_weight = -2, Illegal_update = -3, Illegal_solve = -4,
                  Matrix_singular = -5 };
    system_error(int reason, const std::string& what) : std::runtime_error(what), m_reason(reason) {}
    int reason() const { return m_reason; }

private:
    int m_reason;
};

struct Options : sap_options {
    Options() { sap_default_options(this); }
};
typedef sap_stats Stats;

template <class T>
class DeviceArray {
public:
    typedef T value_type;
    DeviceArray() = default;
    explicit DeviceArray(size_t n) { resize(n); }
    ~DeviceArray() { if (p_) (void)hipFree(p_); }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    void resize(size_t n)
    {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = n;
        if (n && hipMalloc(&p_, n * sizeof(T)) != hipSuccess) throw std::bad_alloc();
    }
    T* data() const { return p_; }
    size_t size() const { return n_; }

private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

// Row-major FP64 band: A(i, j) at band[i*(2k+1) + k + (j-i)] (sap/banded_matrix.h:74-92).
template <class Array>
struct BandedMatrix {
    int num_rows = 0;
    int m_k = 0;
    const Array* band = nullptr;
    BandedMatrix(int n, int k, const Array& b) : num_rows(n), m_k(k), band(&b) {}
};

// SPMV functor over a BandedMatrix (sap/spmv.h:27-74).
template <class Array>
class MVBanded {
public:
    explicit MVBanded(const BandedMatrix<Array>& A) : m_A(A), num_rows(A.num_rows) {}
    void operator()(const Array& v, Array& Av) const { spmv_raw(v.data(), Av.data(), num_rows); }
    void spmv_raw(const double* v, double* Av, int n) const
    {
        int rc = sap_band_spmv(n, m_A.m_k, m_A.band->data(), v, Av, nullptr);
        if (rc < 0) throw system_error(rc, "sap_band_spmv failed");
    }
    const BandedMatrix<Array>& m_A;
    int num_rows;
};

// FP64 CSR matrix on the device with zero-based int indices, the fields of
// cusp::csr_matrix<int, double, cusp::device_memory>.
template <class IndexArray, class ValueArray>
struct CsrMatrix {
    int num_rows = 0, num_cols = 0, num_entries = 0;
    const IndexArray* row_offsets = nullptr;
    const IndexArray* column_indices = nullptr;
    const ValueArray* values = nullptr;
    CsrMatrix(int n, int nnz, const IndexArray& rp, const IndexArray& ci, const ValueArray& val)
        : num_rows(n), num_cols(n), num_entries(nnz), row_offsets(&rp), column_indices(&ci), values(&val) {}
};

// SPMV functor over a CsrMatrix (SpmvCusp, sap/spmv.h:83-129).
template <class Matrix>
class SpmvCusp {
public:
    explicit SpmvCusp(const Matrix& A) : m_A(A), num_rows(A.num_rows) {}
    template <class Array>
    void operator()(const Array& v, Array& Av) const { spmv_raw(v.data(), Av.data(), num_rows); }
    void spmv_raw(const double* v, double* Av, int n) const
    {
        int rc = sap_csr_spmv(n, m_A.num_entries, m_A.row_offsets->data(), m_A.column_indices->data(),
                              m_A.values->data(), v, Av, nullptr);
        if (rc < 0) throw system_error(rc, "sap_csr_spmv failed");
    }
    const Matrix& m_A;
    int num_rows;
};

template <class Array, class PrecValueType>
class Solver {
public:
    Solver(int numPartitions, const Options& opts)
    {
        Options o = opts;
        o.precValueType = sizeof(PrecValueType) == sizeof(double) ? SAP_PREC_DOUBLE : SAP_PREC_FLOAT;
        m_h = sap_create(numPartitions, &o);
        if (!m_h) throw system_error(SAP_ERR_UNSUPPORTED, "sap_create rejected the options");
    }
    ~Solver() { sap_destroy(m_h); }
    Solver(const Solver&) = delete;
    Solver& operator=(const Solver&) = delete;

    template <class A>
    bool setup(const BandedMatrix<A>& M)
    {
        check(sap_setup_banded(m_h, M.num_rows, M.m_k, M.band->data()));
        return true;
    }

    // CSR input: DB / RCM / drop-off as the options ask.
    template <class I, class V>
    bool setup(const CsrMatrix<I, V>& M)
    {
        check(sap_setup_csr(m_h, M.num_rows, M.num_entries, M.row_offsets->data(), M.column_indices->data(),
                            M.values->data()));
        return true;
    }

    // New values, same pattern (sap/solver.h:421-462); needs trackReordering.
    template <class Array1>
    bool update(const Array1& entries)
    {
        int rc = sap_update(m_h, entries.data(), (int64_t)entries.size());
        check(rc);
        return rc == 1;
    }

    // Any SPMV functor with operator()(const Array&, Array&); MVBanded over
    // the setup matrix takes the built-in path.
    template <class SpmvOperator>
    bool solve(SpmvOperator& spmv, const Array& b, Array& x)
    {
        m_ctx = Ctx{&spmv, b.size()};
        int rc = sap_solve(m_h, &Solver::trampoline<SpmvOperator>, &m_ctx, b.data(), x.data(), 1);
        check(rc);
        return rc == 1;
    }

    Stats getStats() const
    {
        Stats s;
        sap_get_stats(m_h, &s);
        return s;
    }
    int getMonitorCode() const { return sap_monitor_code(m_h); }
    std::string getMonitorMessage() const { return sap_monitor_message(m_h); }

private:
    struct Ctx { void* op; size_t n; };
    // Non-owning Array views over raw device pointers for the user functor.
    struct View {
        typedef typename Array::value_type value_type;
        value_type* p; size_t n;
        value_type* data() const { return p; }
        size_t size() const { return n; }
    };
    template <class Op>
    static void trampoline(void* ctx, const double* v, double* Av, int n)
    {
        Ctx* c = static_cast<Ctx*>(ctx);
        invoke(*static_cast<Op*>(c->op), v, Av, n, 0);
    }
    // functors written against Array views
    template <class Op>
    static auto invoke(Op& op, const double* v, double* Av, int n, int)
        -> decltype(op(std::declval<const View&>(), std::declval<View&>()), void())
    {
        View vin{const_cast<double*>(v), (size_t)n}, vout{Av, (size_t)n};
        op(vin, vout);
    }
    // MVBanded, SpmvCusp and other functors bound to Array itself: their raw entry
    template <class Op>
    static void invoke(Op& op, const double* v, double* Av, int n, long)
    {
        op.spmv_raw(v, Av, n);
    }
    void check(int rc) const
    {
        if (rc < 0) throw system_error(rc, sap_last_error(m_h));
    }
    sap_solver* m_h = nullptr;
    Ctx m_ctx{};
};

}  // namespace sap
