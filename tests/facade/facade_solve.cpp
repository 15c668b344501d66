// Runs the reference-shaped C++ interface (include/sap/solver.hpp) end to
// end on the GPU: banded setup + solve with MVBanded, a user SPMV functor,
// update() with new values, a CSR setup with reordering + SpmvCusp, and
// Illegal_solve before setup.  Prints one line per check and exits 0 only
// when every check passes.  Built by tests/facade/Makefile, run by
// tests/test_gpu_api.py.
#include <cmath>
#include <cstdio>
#include <vector>

#include "sap/solver.hpp"

typedef sap::DeviceArray<double> Vec;

static std::vector<double> to_host(const Vec& v)
{
    std::vector<double> h(v.size());
    if (hipMemcpy(h.data(), v.data(), v.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        throw std::runtime_error("copy failed");
    return h;
}

static void to_dev(Vec& v, const std::vector<double>& h)
{
    if (hipMemcpy(v.data(), h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        throw std::runtime_error("copy failed");
}

static double norm(const std::vector<double>& a)
{
    double s = 0;
    for (double v : a) s += v * v;
    return std::sqrt(s);
}

// |b - A x| / |b| with the library's banded SpMV
template <class Op>
static double rel_residual(Op& op, const Vec& b, const Vec& x)
{
    Vec ax(b.size());
    op(x, ax);
    std::vector<double> hb = to_host(b), hax = to_host(ax);
    for (size_t i = 0; i < hb.size(); ++i) hax[i] = hb[i] - hax[i];
    return norm(hax) / norm(hb);
}

// a user functor written against Array views (the reference's SPMV concept)
struct UserBand {
    int n, k;
    const double* band;
    int calls = 0;
    template <class A>
    void operator()(const A& v, A& Av)
    {
        ++calls;
        if (sap_band_spmv(n, k, band, v.data(), Av.data(), nullptr) < 0) throw std::runtime_error("spmv");
    }
};

static int failures = 0;
static void check(bool ok, const char* what, double val)
{
    std::printf("%s %s %.3e\n", ok ? "PASS" : "FAIL", what, val);
    if (!ok) ++failures;
}

int main()
{
    const int n = 20000, k = 30, P = 8;
    const size_t w = 2 * k + 1;
    Vec band(n * w), xt(n), b(n), x(n);
    if (sap_gen_banded(n, k, 0.9, 5, band.data(), nullptr) || sap_gen_vector(n, 6, xt.data(), nullptr)) return 2;
    sap::BandedMatrix<Vec> A(n, k, band);
    sap::MVBanded<Vec> spmv(A);
    spmv(xt, b);

    sap::Options opts;
    opts.factMethod = sap::LU_UL;
    opts.precondType = sap::Spike;
    opts.relTol = 1e-10;
    opts.refine = 6;
    opts.trackReordering = 1;

    // Illegal_solve before setup
    {
        sap::Solver<Vec, float> s0(P, opts);
        int reason = 0;
        try {
            s0.solve(spmv, b, x);
        } catch (const sap::system_error& e) {
            reason = e.reason();
        }
        check(reason == sap::system_error::Illegal_solve, "illegal_solve_before_setup", reason);
    }

    sap::Solver<Vec, float> solver(P, opts);
    solver.setup(A);
    to_dev(x, std::vector<double>(n, 0.0));
    bool ok = solver.solve(spmv, b, x);
    check(ok && solver.getMonitorCode() == 1, "banded_solve_converged", solver.getStats().relResidualNorm);
    check(rel_residual(spmv, b, x) <= 1e-10, "banded_true_residual", rel_residual(spmv, b, x));
    check(solver.getStats().numPartitions == P, "partitions", solver.getStats().numPartitions);

    // a user functor
    UserBand user{n, k, band.data()};
    to_dev(x, std::vector<double>(n, 0.0));
    ok = solver.solve(user, b, x);
    check(ok && user.calls == solver.getStats().numSpmv && user.calls > 0, "user_functor_solve", user.calls);
    check(rel_residual(spmv, b, x) <= 1e-10, "user_functor_true_residual", rel_residual(spmv, b, x));

    // update(): every value scaled by 1.02 except the diagonal (same pattern)
    std::vector<double> hb = to_host(band);
    for (int i = 0; i < n; ++i)
        for (size_t j = 0; j < w; ++j)
            if ((int)j != k) hb[i * w + j] *= 1.02;
    Vec band2(n * w), b2(n);
    to_dev(band2, hb);
    sap::BandedMatrix<Vec> A2(n, k, band2);
    sap::MVBanded<Vec> spmv2(A2);
    spmv2(xt, b2);
    check(solver.update(band2), "update_accepted", 1);
    to_dev(x, std::vector<double>(n, 0.0));
    ok = solver.solve(spmv2, b2, x);
    check(ok && rel_residual(spmv2, b2, x) <= 1e-10, "updated_solve", rel_residual(spmv2, b2, x));
    Vec shorter(10);
    check(!solver.update(shorter), "update_count_mismatch_returns_false", 0);

    // CSR setup with reordering (DB, RCM, scaling) and the SpmvCusp functor
    {
        std::vector<int> rp(1, 0), ci;
        std::vector<double> val;
        std::vector<double> hband = to_host(band);
        for (int i = 0; i < n; ++i) {
            for (int j = std::max(0, i - k); j <= std::min(n - 1, i + k); ++j) {
                ci.push_back(j);
                val.push_back(hband[i * w + k + (j - i)]);
            }
            rp.push_back((int)ci.size());
        }
        sap::DeviceArray<int> drp(rp.size()), dci(ci.size());
        Vec dval(val.size());
        if (hipMemcpy(drp.data(), rp.data(), rp.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dci.data(), ci.data(), ci.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
            return 2;
        to_dev(dval, val);
        typedef sap::CsrMatrix<sap::DeviceArray<int>, Vec> Csr;
        Csr C(n, (int)val.size(), drp, dci, dval);
        sap::SpmvCusp<Csr> cs(C);
        sap::Options o2;
        o2.relTol = 1e-10;
        o2.maxBandwidth = 1000;   // RCM on
        sap::Solver<Vec, double> s2(P, o2);
        s2.setup(C);
        to_dev(x, std::vector<double>(n, 0.0));
        ok = s2.solve(cs, b, x);
        check(ok && s2.getMonitorCode() == 1, "csr_reordered_solve", s2.getStats().relResidualNorm);
        check(rel_residual(cs, b, x) <= 1e-10, "csr_true_residual", rel_residual(cs, b, x));
        check(s2.getStats().bandwidth > 0 && s2.getStats().bandwidthReorder > 0, "csr_stats_bandwidths",
              s2.getStats().bandwidth);
    }
    std::printf("%s: %d failure(s)\n", failures ? "FAILED" : "ALL PASSED", failures);
    return failures ? 1 : 0;
}
