// Host solver: the Krylov methods and monitors (sap/bicgstab2.h,
// sap/minres.h, sap/monitor.h), the solver state machine (sap/solver.h) and
// the extern "C" entry points declared in include/sap_c.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <string>

#include "batch.hpp"
#include "precond.hpp"

namespace sap_amd {

// --------------------------------------------------------------- Profiler
Profiler::~Profiler()
{
    for (auto& p : pending_) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
    for (auto e : pool_) (void)hipEventDestroy(e);
}
hipEvent_t Profiler::get_event()
{
    if (!pool_.empty()) { hipEvent_t e = pool_.back(); pool_.pop_back(); return e; }
    hipEvent_t e;
    SAP_HIP_CHECK(hipEventCreate(&e));
    return e;
}
void Profiler::begin(const char* name)
{
    int idx = -1;
    for (size_t i = 0; i < entries.size(); ++i)
        if (entries[i].name == name) { idx = (int)i; break; }
    if (idx < 0) { entries.push_back({name, 0, 0.0}); idx = (int)entries.size() - 1; }
    cur_ = idx;
    cur_a_ = get_event();
    SAP_HIP_CHECK(hipEventRecord(cur_a_, stream));
}
void Profiler::end()
{
    if (cur_ < 0) return;
    hipEvent_t b = get_event();
    SAP_HIP_CHECK(hipEventRecord(b, stream));
    pending_.push_back({cur_, cur_a_, b});
    entries[cur_].launches++;
    cur_ = -1;
    if (pending_.size() > 4096) flush();
}
void Profiler::flush()
{
    for (auto& p : pending_) {
        float ms = 0.f;
        SAP_HIP_CHECK(hipEventSynchronize(p.b));
        SAP_HIP_CHECK(hipEventElapsedTime(&ms, p.a, p.b));
        entries[p.idx].ms += ms;
        pool_.push_back(p.a);
        pool_.push_back(p.b);
    }
    pending_.clear();
}
void Profiler::reset()
{
    flush();
    entries.clear();
}

namespace {

double wall_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// BiCGStabLMonitor (sap/monitor.h:153-335)
struct BiCGStabLMonitor {
    int    maxIterations = 100;
    float  iterations = 0.f;
    int    maxmSteps = 8;
    int    stag = 0, moreSteps = 0, maxStagSteps = 3;
    double relTol = 1e-6, absTol = 0, rhsNorm = 0, rNorm = std::numeric_limits<double>::max();
    int    code = 0;
    std::string message;

    void   reset(int maxIt, int maxm, double rt, double at)
    {
        maxIterations = maxIt; maxmSteps = maxm; relTol = rt; absTol = at;
        iterations = 0.f; stag = 0; moreSteps = 0; code = 0; message.clear();
        rNorm = std::numeric_limits<double>::max();
    }
    void   init(double rhs_norm) { rhsNorm = rhs_norm; iterations = 0; code = 0; message = ""; }
    double tolerance() const { return absTol + relTol * rhsNorm; }
    void   stop(int c, const char* m) { code = c; message = m; }
    void   increment(float x) { iterations += x; }
    void   incrementStag() { stag++; }
    void   resetStag() { stag = 0; }
    bool   finished()
    {
        if (code != 0) return true;
        if (iterations > maxIterations) stop(-1, "Maximum number of iterations was reached");
        return code != 0;
    }
    bool   finished(double r)
    {
        rNorm = r;
        if (code != 0) return true;
        if (std::isnan(rNorm)) stop(-2, "Residual norm is NaN");
        else if (rNorm <= tolerance()) stop(1, "Converged");
        else if (iterations > maxIterations) stop(-1, "Maximum number of iterations was reached");
        else {
            if (stag >= maxStagSteps && moreSteps == 0) stag = 0;
            moreSteps++;
            if (moreSteps >= maxmSteps) stop(-3, "Stagnation encountered");
        }
        return code != 0;
    }
    bool   needCheckConvergence(double r)
    {
        rNorm = r;
        if (code != 0) return false;
        if (std::isnan(rNorm)) stop(-2, "Residual norm is NaN");
        return code == 0 && (rNorm <= tolerance() || stag >= maxStagSteps || moreSteps);
    }
    bool   converged() const { return code > 0; }
};

// Monitor (sap/monitor.h:25-147)
struct Monitor {
    int    maxIterations = 100;
    float  iterations = 0.f;
    double relTol = 1e-6, absTol = 0, rhsNorm = 0, rNorm = std::numeric_limits<double>::max();
    int    code = 0;
    std::string message;
    void   reset(int maxIt, double rt, double at)
    {
        maxIterations = maxIt; relTol = rt; absTol = at; iterations = 0.f; code = 0; message.clear();
        rNorm = std::numeric_limits<double>::max();
    }
    void   init(double rhs_norm) { rhsNorm = rhs_norm; iterations = 0; code = 0; message = ""; }
    double tolerance() const { return absTol + relTol * rhsNorm; }
    void   stop(int c, const char* m) { code = c; message = m; }
    size_t iteration_count() const { return (size_t)(iterations + 0.5f); }
    bool   finished(double r)
    {
        rNorm = r;
        if (code != 0) return true;
        if (std::isnan(rNorm)) stop(-2, "Residual norm is NaN");
        else if (rNorm <= tolerance()) stop(1, "Converged");
        else if (iterations > maxIterations) stop(-1, "Maximum number of iterations was reached");
        return code != 0;
    }
    bool   converged() const { return code > 0; }
};

}  // namespace

// ------------------------------------------------------------------ Solver
class Solver {
public:
    sap_options opt;
    int P_req;
    hipStream_t stream = nullptr;
    Profiler prof;
    std::unique_ptr<Precond<float>> pf;
    std::unique_ptr<Precond<double>> pd;
    bool setup_done = false;
    int64_t n = 0;
    int k = 0;
    const double* d_band = nullptr;     // non-owning (the MVBanded reference)
    // non-owning CSR of a CSR setup (the SpmvCusp functor keeps a reference
    // to the matrix, sap/spmv.h:83-129)
    const int* d_rp = nullptr;
    const int* d_ci = nullptr;
    const double* d_val = nullptr;
    int64_t nnz = 0;
    sap_spmv_fn user_spmv = nullptr;
    void* user_ctx = nullptr;
    BiCGStabLMonitor bmon;
    Monitor mmon;
    sap_stats stats;
    std::string last_error;
    int code = 0;
    std::string message;

    DevBuf<double> work_, partials_;
    DevBuf<double> dgam_{1};   // a BiCGStab(l) gamma, kept on the device
    // When set (the allocation-free solve), the vector pool and the
    // reduction partials are taken from this caller-supplied device memory
    // instead of work_ / partials_.
    double* ext_work_ = nullptr;
    size_t ext_work_size_ = 0;
    double* ext_partials_ = nullptr;
    double* hscalar_ = nullptr;
    double* dscalar_map_ = nullptr;   // hscalar_ in the device's address space
    std::vector<double*> vecs_;
    size_t vec_next_ = 0;
    // b - A x of the last BiCGStab(L) solve when it converged (on a passed
    // true-residual check; x its returned solution), else null; a pool
    // vector valid until the next vec().
    const double* last_res_ = nullptr;
    double last_res_dot_ = -1.0;   // (last_res_, last_res_) as dot() gave it

    Solver(int P, const sap_options& o) : opt(o), P_req(P)
    {
        std::memset(&stats, 0, sizeof(stats));
        stats.rhsNorm = stats.residualNorm = stats.relResidualNorm = std::numeric_limits<double>::max();
        // reduction results land here directly (mapped, coherent pinned memory)
        SAP_HIP_CHECK(hipHostMalloc((void**)&hscalar_, sizeof(double) * 8,
                                    hipHostMallocMapped | hipHostMallocCoherent));
        SAP_HIP_CHECK(hipHostGetDevicePointer((void**)&dscalar_map_, hscalar_, 0));
        prof.enabled = opt.profile != 0;
        make_precond();
    }
    ~Solver()
    {
        if (hscalar_) (void)hipHostFree(hscalar_);
        batch_.reset();
        if (own_stream_) (void)hipStreamDestroy(own_stream_);
    }

    void make_precond()
    {
        PrecondConfig c;
        c.numPartitions = P_req;
        c.factMethod = opt.factMethod;
        c.precondType = opt.precondType;
        c.safeFactorization = opt.safeFactorization;
        c.autotune = opt.autotune;
        c.bcr = opt.useBCR;
        if (opt.precValueType == SAP_PREC_DOUBLE) {
            pd = std::make_unique<Precond<double>>();
            pd->cfg = c; pd->stream = stream; pd->prof = &prof;
            pf.reset();
        } else {
            pf = std::make_unique<Precond<float>>();
            pf->cfg = c; pf->stream = stream; pf->prof = &prof;
            pd.reset();
        }
    }
    void set_stream(hipStream_t s)
    {
        settle();   // the deferred setup work is ordered on the old stream
        stream = s;
        prof.stream = s;
        if (pf) pf->stream = s;
        if (pd) pd->stream = s;
    }
    int num_partitions() const { return pf ? pf->P : (pd ? pd->P : 0); }
    // waits for the setup work left on the preconditioner's second stream
    // and copies the statistics it produces
    void settle()
    {
        auto fin = [&](auto& p) {
            p.finish_pending();
            stats.time_assembly = p.time_assembly;
            stats.time_fullLU = p.time_fullLU;
            stats.actual_nnz = (int)std::min<int64_t>(p.actual_nnz, INT_MAX);
        };
        if (pf) fin(*pf);
        else if (pd) fin(*pd);
    }

    // ---------------------------------------------------------- vectors
    void reserve_vectors(int count)
    {
        const size_t need = (size_t)count * n;
        double* base;
        if (ext_work_) {
            if (ext_work_size_ < need) throw Error(SAP_ERR_BADARG, "solve work space too small");
            base = ext_work_;
        } else {
            if (work_.size() < need) work_.resize(need);
            partials_.resize(dot_partials_size());
            base = work_.get();
        }
        vecs_.clear();
        for (int i = 0; i < count; ++i) vecs_.push_back(base + (size_t)i * n);
        vec_next_ = 0;
    }
    double* vec() { return vecs_.at(vec_next_++); }
    double* partials() const { return ext_partials_ ? ext_partials_ : partials_.get(); }

    double dot(const double* a, const double* b)
    {
        launch_dot(n, a, b, partials(), dscalar_map_, stream);
        SAP_HIP_CHECK(hipStreamSynchronize(stream));
        return hscalar_[0];
    }
    double nrm2(const double* a) { return std::sqrt(dot(a, a)); }
    // norms of up to three vectors with one pass and one synchronization
    // (each value as nrm2 gives it)
    void nrm2s(int m, const double* const* v, double* out)
    {
        launch_dots(n, m, v, v, partials(), dscalar_map_, stream);
        SAP_HIP_CHECK(hipStreamSynchronize(stream));
        for (int q = 0; q < m; ++q) out[q] = std::sqrt(hscalar_[q]);
    }
    void axpby(const double* x, const double* y, double* z, double alpha, double beta)
    {
        launch_axpby(n, alpha, x, beta, y, z, stream);
    }
    void axpy(const double* x, double* y, double alpha) { launch_axpby(n, alpha, x, 1.0, y, y, stream); }
    void copy(const double* x, double* y) { launch_copy(n, x, y, stream); }

    void spmv(const double* x, double* y)
    {
        stats.numSpmv++;
        ProfScope ps(&prof, "spmv");
        if (user_spmv) user_spmv(user_ctx, x, y, (int)n);
        else if (d_rp) launch_csr_spmv(n, nnz, d_rp, d_ci, d_val, x, y, stream);
        else launch_band_spmv(n, k, d_band, x, y, stream);
    }
    void prec(const double* v, double* z)
    {
        stats.numApplies++;
        if (pf) pf->apply(v, z);
        else pd->apply(v, z);
    }

    // ------------------------------------------- BiCGStab(L), sap/bicgstab2.h:31-354
    template <int L>
    void bicgstabl(double* x, const double* b, BiCGStabLMonitor& monitor, bool x_is_zero, bool correction,
                   double b_dot)
    {
        // for a zero guess the initial residual is b, whose norm the monitor holds
        const double eps = 1e-20;
        double rho0 = 1, alpha = 0, omega = 1, rho1;
        double* r0 = vec();
        double* xx = vec();
        double* Pv = vec();
        double* rr[L + 1];
        double* uu[L + 1];
        for (int i = 0; i <= L; ++i) { rr[i] = vec(); uu[i] = vec(); }
        double* x_min = vec();
        double* Pxx = vec();
        double* APxx = vec();
        double* rsave = vec();
        double tao[L + 1][L + 1] = {};
        double gamma[L + 2] = {}, gamma_prime[L + 2] = {}, gamma_pp[L + 2] = {}, sigma[L + 2] = {};

        // r0 = rr(0) <- b - A x, uu(0) <- 0, xx <- x.  A*0 = 0 exactly, so a
        // zero guess needs no SpMV (and x is not read).  rr(1..L) and
        // uu(1..L) are written before they are read, so they are not
        // cleared; x_min (zero until a first minimum is recorded) is only
        // read by the non-converged ending, which clears it then.
        if (x_is_zero) {
            double* cp[2] = {r0, rr[0]};
            double* zs[2] = {uu[0], xx};
            launch_spread(n, b, cp, 2, zs, 2, stream);
        } else {
            spmv(x, r0);
            axpby(b, r0, r0, 1.0, -1.0);
            double* cp[2] = {rr[0], nullptr};
            double* zs[1] = {uu[0]};
            launch_spread(n, r0, cp, 1, zs, 1, stream);
            copy(x, xx);
        }
        bool have_xmin = false;
        double r_norm_min = x_is_zero ? monitor.rhsNorm : nrm2(r0);
        double r_norm = r_norm_min;
        double r_norm_act = r_norm;
        bool pxx_valid = false;    // Pxx == M^{-1} xx for the current xx

        last_res_ = nullptr;
        auto check_true = [&]() -> bool {
            prec(xx, Pxx);
            spmv(Pxx, APxx);
            axpby(b, APxx, rr[0], 1.0, -1.0);
            pxx_valid = true;
            r_norm_act = nrm2(rr[0]);
            const double rdot = hscalar_[0];
            const bool done = monitor.finished(r_norm_act);
            if (done && monitor.converged()) {
                last_res_ = rr[0];
                last_res_dot_ = rdot;
            }
            return done;
        };

        // A refinement correction solve (x = 0) whose first step already
        // meets the tolerance returns x = alpha M^{-1} uu(0), using the
        // application made for uu(1): its residual b - alpha A M^{-1} uu(0)
        // is the recursive rr(0) (up to FP64 rounding of the products), so
        // the check needs no further application or SpMV.  (The reference
        // returns M^{-1}(alpha uu(0)), which differs only by the rounding of
        // the preconditioner; the refinement forms b - A x afresh before it
        // declares convergence.)
        bool first_step = correction && x_is_zero;
        // With a zero guess rr(0) = r0 = b on the first step, so (rr(0), r0)
        // is (b, b), which the caller may already hold as dot() gives it
        // (b_dot >= 0): the same kernel on the same values, so the same bits.
        bool rho_known = x_is_zero && b_dot >= 0.0;
        while (true) {
            rho0 = -omega * rho0;
            monitor.increment(0.25f);
            for (int j = 0; j < L; ++j) {
                if (rho_known) {
                    rho1 = b_dot;
                    rho_known = false;
                } else {
                    rho1 = dot(rr[j], r0);
                }
                if (rho0 == 0) { monitor.stop(-10, "rho0 is zero"); break; }
                const double beta = alpha * rho1 / rho0;
                rho0 = rho1;
                for (int i = 0; i <= j; ++i) axpby(rr[i], uu[i], uu[i], 1.0, -beta);
                prec(uu[j], Pv);
                spmv(Pv, uu[j + 1]);
                // gamma = (uu(j+1), r0) stays on the device: the updates
                // rr(i) -= alpha uu(i+1) take alpha = rho0 / gamma from there,
                // and the host reads gamma with the norms below (one
                // synchronization instead of two; the same division, so the
                // same alpha)
                {
                    const double* ga[1] = {uu[j + 1]};
                    const double* gb[1] = {r0};
                    launch_dots(n, 1, ga, gb, partials(), dgam_.get(), stream);
                }
                for (int i = 0; i <= j; ++i) launch_axpy_ratio(n, rho0, dgam_.get(), -1.0, uu[i + 1], rr[i], stream);
                launch_copy(1, dgam_.get(), dscalar_map_ + 4, stream);
                double nv[3];
                {
                    const double* vs[3] = {rr[0], uu[0], xx};
                    nrm2s(3, vs, nv);   // |rr(0)|, and |uu(0)|, |xx| for the stagnation test below
                }
                const double rr0_dot = hscalar_[0];
                const double gam = hscalar_[4];
                if (gam == 0) { monitor.stop(-11, "gamma is zero"); break; }
                alpha = rho0 / gam;
                r_norm_act = r_norm = nv[0];
                // rr(j+1) <- A P^{-1} rr(j) of the recursive rr(j).  The
                // reference forms it before the true-residual check; here it
                // is formed after a check that does not end the solve (the
                // check overwrites rr(0), so a recursive rr(0) is saved
                // first), which gives the same iterates and saves one
                // preconditioner application and one SpMV on the converging
                // step.
                if (std::fabs(alpha) * nv[1] < eps * nv[2]) monitor.incrementStag();
                else monitor.resetStag();
                axpy(uu[0], xx, alpha);
                pxx_valid = false;
                const bool first = first_step;
                first_step = false;
                if (monitor.needCheckConvergence(r_norm)) {
                    if (first) {
                        axpby(Pv, Pv, Pxx, alpha, 0.0);   // Pv = M^{-1} uu(0) until rr(1) is formed
                        pxx_valid = true;
                        r_norm_act = r_norm;
                        const bool done = monitor.finished(r_norm_act);
                        if (done && monitor.converged()) {
                            last_res_ = rr[0];
                            last_res_dot_ = rr0_dot;
                        }
                        if (done) break;
                        // Not reached: on a first step the monitor has no
                        // stagnation count >= 3 and no extra steps, so it asks
                        // for a check only when r_norm <= tol, and finished()
                        // then converges (a NaN norm stops in
                        // needCheckConvergence).  The reference would hold
                        // b - A M^{-1} xx in rr(0) here; the recursive rr(0)
                        // is kept.
                        prec(rr[0], Pv);
                    } else {
                        const double* src = rr[j];
                        if (j == 0) { copy(rr[0], rsave); src = rsave; }
                        if (check_true()) break;
                        prec(src, Pv);
                    }
                } else {
                    prec(rr[j], Pv);
                }
                spmv(Pv, rr[j + 1]);
                if (r_norm_act < r_norm_min) { r_norm_min = r_norm_act; copy(xx, x_min); have_xmin = true; }
                if (monitor.finished()) break;
            }
            if (monitor.finished()) break;

            for (int j = 1; j <= L; ++j) {
                for (int i = 1; i < j; ++i) {
                    tao[i][j] = dot(rr[j], rr[i]) / sigma[i];
                    axpy(rr[i], rr[j], -tao[i][j]);
                }
                sigma[j] = dot(rr[j], rr[j]);
                if (sigma[j] == 0) { monitor.stop(-12, "a sigma value is zero"); break; }
                gamma_prime[j] = dot(rr[j], rr[0]) / sigma[j];
            }
            if (monitor.finished()) break;

            gamma[L] = gamma_prime[L];
            omega = gamma[L];
            for (int j = L - 1; j > 0; --j) {
                gamma[j] = gamma_prime[j];
                for (int i = j + 1; i <= L; ++i) gamma[j] -= tao[j][i] * gamma[i];
            }
            for (int j = 1; j < L; ++j) {
                gamma_pp[j] = gamma[j + 1];
                for (int i = j + 1; i < L; ++i) gamma_pp[j] += tao[j][i] * gamma[i + 1];
            }
            {
                double nv[2];
                const double* vs[2] = {rr[0], xx};
                nrm2s(2, vs, nv);
                if (std::fabs(gamma[1]) * nv[0] < eps * nv[1]) monitor.incrementStag();
                else monitor.resetStag();
            }
            axpy(rr[0], xx, gamma[1]);
            axpy(rr[L], rr[0], -gamma_prime[L]);
            axpy(uu[L], uu[0], -gamma[L]);
            pxx_valid = false;
            r_norm_act = r_norm = nrm2(rr[0]);
            monitor.increment(0.25f);
            if (monitor.needCheckConvergence(r_norm)) {
                if (check_true()) break;
            }
            if (r_norm_act < r_norm_min) { r_norm_min = r_norm_act; copy(xx, x_min); have_xmin = true; }
            if (monitor.finished()) break;
            monitor.increment(0.25f);

            bool brk = false;
            for (int j = 1; j < L; ++j) {
                axpy(uu[j], uu[0], -gamma[j]);
                {
                    double nv[2];
                    const double* vs[2] = {rr[j], xx};
                    nrm2s(2, vs, nv);
                    if (std::fabs(gamma_pp[j]) * nv[0] < eps * nv[1]) monitor.incrementStag();
                    else monitor.resetStag();
                }
                axpy(rr[j], xx, gamma_pp[j]);
                axpy(rr[j], rr[0], -gamma_prime[j]);
                pxx_valid = false;
                r_norm_act = r_norm = nrm2(rr[0]);
                if (monitor.needCheckConvergence(r_norm)) {
                    if (check_true()) { brk = true; break; }
                }
                if (r_norm_act < r_norm_min) { r_norm_min = r_norm_act; copy(xx, x_min); have_xmin = true; }
                if (monitor.finished()) { brk = true; break; }
            }
            if (brk || monitor.finished()) break;
            monitor.increment(0.25f);
        }

        if (monitor.converged()) {
            // x <- M^{-1} xx; the preconditioner is deterministic, so a
            // still-valid M^{-1} xx from the last true-residual check is reused.
            if (pxx_valid) copy(Pxx, x);
            else prec(xx, x);
        } else {
            double* Pxmin = vec();
            double* APxmin = vec();
            double* rc = vec();
            double* rcm = vec();
            prec(xx, Pxx);
            spmv(Pxx, APxx);
            axpby(b, APxx, rc, 1.0, -1.0);
            if (!have_xmin) launch_fill(n, 0.0, x_min, stream);
            prec(x_min, Pxmin);
            spmv(Pxmin, APxmin);
            axpby(b, APxmin, rcm, 1.0, -1.0);
            const double rcn = nrm2(rc), rcmn = nrm2(rcm);
            if (rcn < rcmn) { copy(Pxx, x); monitor.rNorm = rcn; }
            else            { copy(Pxmin, x); monitor.rNorm = rcmn; }
        }
    }

    // ------------------------------------------------- MINRES, sap/minres.h:38-181
    void minres(double* x, const double* b, Monitor& monitor)
    {
        const double eps = std::numeric_limits<double>::epsilon();
        double* y = vec();
        double* r1 = vec();
        double* v = vec();
        double* w = vec();
        double* w1 = vec();
        double* w2 = vec();
        double* r2 = vec();
        double* tmp = vec();
        spmv(x, r1);
        axpby(b, r1, r1, 1.0, -1.0);
        prec(r1, y);
        double beta1 = dot(r1, y);
        if (beta1 < 0) { monitor.stop(-10, "The preconditioner is not positive definite"); beta1 *= -1; }
        else if (beta1 > 0) beta1 = std::sqrt(beta1);
        double Anorm = 0, ynorm = 0, phibar = beta1;
        double oldb = 0, beta = beta1, dbar = 0, epsln = 0, oldeps = 0;
        double phi = 0, rhs1 = beta1, rhs2 = 0, tnorm2 = 0;
        double cs = -1, sn = 0, gmax = 0, gmin = std::numeric_limits<double>::max();
        double alpha = 0, gamma = 0, delta = 0, gbar = 0, z = 0;
        launch_fill(n, 0.0, w, stream);
        launch_fill(n, 0.0, w1, stream);
        launch_fill(n, 0.0, w2, stream);
        copy(r1, r2);
        while (!monitor.finished(phibar)) {
            const double s = 1 / beta;
            axpby(y, y, v, s, 0.0);
            spmv(v, y);
            if (monitor.iteration_count() > 0) axpby(y, r1, y, 1.0, -beta / oldb);
            alpha = dot(v, y);
            axpby(y, r2, y, 1.0, -alpha / beta);
            copy(r2, r1);
            copy(y, r2);
            prec(r2, y);
            oldb = beta;
            beta = dot(r2, y);
            if (beta < 0) { monitor.stop(-10, "The preconditioner is not positive definite"); break; }
            beta = std::sqrt(beta);
            tnorm2 += alpha * alpha + oldb * oldb + beta * beta;
            if (monitor.iteration_count() == 0) {
                if (beta / beta1 <= 10 * eps) monitor.stop(10, "beta2 = 0. If M = I, b and x are eigenvectors");
            }
            oldeps = epsln;
            delta = cs * dbar + sn * alpha;
            gbar = sn * dbar - cs * alpha;
            epsln = sn * beta;
            dbar = -cs * beta;
            const double root = std::sqrt(gbar * gbar + dbar * dbar);
            gamma = std::sqrt(gbar * gbar + beta * beta);
            gamma = std::max(gamma, eps);
            cs = gbar / gamma;
            sn = beta / gamma;
            phi = cs * phibar;
            phibar = sn * phibar;
            const double denom = 1 / gamma;
            copy(w2, w1);
            copy(w, w2);
            // w = denom*v - oldeps*denom*w1 - delta*denom*w2  (cusp::blas::axpbypcz)
            axpby(v, w1, tmp, denom, -oldeps * denom);
            axpby(tmp, w2, w, 1.0, -delta * denom);
            axpby(x, w, x, 1.0, phi);
            gmax = std::max(gmax, gamma);
            gmin = std::min(gmin, gamma);
            z = rhs1 / gamma;
            rhs1 = rhs2 - delta * z;
            rhs2 = -epsln * z;
            Anorm = std::sqrt(tnorm2);
            ynorm = nrm2(x);
            if (monitor.code != 10) {
                const double test2 = root / Anorm;
                const double Acond = gmax / gmin;
                if (Acond >= 0.1 / eps) monitor.stop(11, "x has converged to an eigenvector");
                if (Anorm * eps * ynorm >= beta1) monitor.stop(12, "Reasonable accuracy achieved, given eps");
                if (test2 <= monitor.absTol) monitor.stop(13, "A least-squares solution was found, given tol");
            }
            monitor.iterations += 1.f;
        }
    }

    bool use_minres() const { return opt.solverType == SAP_MINRES; }

    // One Krylov solve of A x = b (the reference's solve path).
    // b_norm: |b| when the caller has it (as nrm2 gives it), else negative;
    // x_is_zero: start from x = 0 without reading x; b_dot: (b, b) as dot()
    // gives it when the caller has it, else negative
    int krylov(double* x, const double* b, double rtol, double atol, bool x_is_zero, bool correction, float* its,
               double* rnorm, double* rhs_norm, double b_norm = -1.0, double b_dot = -1.0)
    {
        last_res_ = nullptr;
        if (use_minres()) {
            if (x_is_zero) launch_fill(n, 0.0, x, stream);
            mmon.reset(opt.maxNumIterations, rtol, atol);
            mmon.init(nrm2(b));
            minres(x, b, mmon);
            *its += mmon.iterations;
            *rnorm = mmon.rNorm;
            *rhs_norm = mmon.rhsNorm;
            message = mmon.message;
            return mmon.code;
        }
        bmon.reset(opt.maxNumIterations, 8, rtol, atol);   // maxmSteps = 8 (sap/solver.h:314-319)
        bmon.init(b_norm >= 0 ? b_norm : nrm2(b));
        if (opt.solverType == SAP_BiCGStab1) bicgstabl<1>(x, b, bmon, x_is_zero, correction, b_dot);
        else bicgstabl<2>(x, b, bmon, x_is_zero, correction, b_dot);
        *its += bmon.iterations;
        *rnorm = bmon.rNorm;
        *rhs_norm = bmon.rhsNorm;
        message = bmon.message;
        return bmon.code;
    }

    // Solver::setup statistics (sap/solver.h:354-399)
    template <class PC>
    void harvest(const PC& p, double t0)
    {
        stats.timeSetup = wall_ms() - t0;
        stats.time_DB = p.time_DB;
        stats.time_reorder = p.time_reorder;
        stats.time_dropOff = p.time_dropOff;
        stats.time_cpu_assemble = p.time_cpu_assemble;
        stats.time_transfer = p.time_transfer;
        stats.time_toBanded = p.time_toBanded;
        stats.time_offDiags = p.time_offDiags;
        stats.time_bandLU = p.time_bandLU;
        stats.time_bandUL = p.time_bandUL;
        stats.time_assembly = p.time_assembly;
        stats.time_fullLU = p.time_fullLU;
        stats.numPartitions = p.P;
        stats.actual_nnz = (int)std::min<int64_t>(p.actual_nnz, INT_MAX);
        stats.bandwidth = p.k;
        stats.bandwidthReorder = p.k_reorder;
        stats.bandwidthDB = p.k_db;
        stats.actualDropOff = p.dropoff_actual;
        stats.d_p1 = p.d_p1;
        stats.d_p1_ori = p.d_p1_ori;
        stats.diag_dom = p.diag_dom;
        stats.diag_dom_ori = p.diag_dom_ori;
        // nuKf = (2 K n - sum ks_row - sum ks_col) / (2 K n) and flops_LU =
        // sum ks_row ks_col / (time_bandLU 1e6); the extents are only
        // assembled by the reordering path (empty otherwise, so nuKf = 1 and
        // flops_LU = 0 there, as in the reference)
        stats.nuKf = p.k == 0 ? 0.0
                              : (2.0 * p.k * (double)n - (p.profile.sum_row + p.profile.sum_col)) /
                                    (2.0 * p.k * (double)n);
        stats.flops_LU = stats.time_bandLU == 0 ? 0.0 : p.profile.sum_prod / (stats.time_bandLU * 1e6);
    }

    int setup_banded(int64_t n_, int k_, const double* band)
    {
        const double t0 = wall_ms();
        setup_done = false;
        n = n_;
        k = k_;
        d_band = band;
        d_rp = d_ci = nullptr;
        d_val = nullptr;
        nnz = 0;
        int rc = pf ? pf->setup(n, k, band) : pd->setup(n, k, band);
        SAP_HIP_CHECK(hipStreamSynchronize(stream));
        if (rc != 0) return rc;
        if (pf) harvest(*pf, t0); else harvest(*pd, t0);
        setup_done = true;
        return 0;
    }

    int setup_csr(int64_t n_, int64_t nnz_, const int* rp, const int* ci, const double* val)
    {
        const double t0 = wall_ms();
        setup_done = false;
        n = n_;
        d_band = nullptr;
        d_rp = rp;
        d_ci = ci;
        d_val = val;
        nnz = nnz_;
        ReorderConfig rc;
        rc.reorder = opt.performReorder != 0;
        rc.doDB = opt.performDB != 0;
        rc.dbFirstStageOnly = opt.dbFirstStageOnly != 0;
        rc.scale = opt.applyScaling != 0;
        rc.maxBandwidth = opt.maxBandwidth;
        rc.dropOffFraction = opt.dropOffFraction;
        rc.track = opt.trackReordering != 0;
        rc.variable = opt.variableBandwidth != 0;
        rc.testDB = opt.testDB != 0;
        int r = pf ? pf->setup_csr(n, nnz, rp, ci, val, rc) : pd->setup_csr(n, nnz, rp, ci, val, rc);
        SAP_HIP_CHECK(hipStreamSynchronize(stream));
        if (r != 0) return r;
        k = pf ? pf->k : pd->k;
        if (pf) harvest(*pf, t0); else harvest(*pd, t0);
        // the DB test builds no preconditioner: a later solve is illegal
        setup_done = !rc.testDB;
        return 0;
    }

    // Solver::update (sap/solver.h:421-462): 1 done, 0 when the number of
    // entries does not match the setup matrix, negative on error.
    int update(const double* entries, int64_t count)
    {
        if (!setup_done) throw Error(SAP_ERR_ILLEGAL_UPDATE, "Illegal call to update() before setup().");
        if (!opt.trackReordering)
            throw Error(SAP_ERR_ILLEGAL_UPDATE, "Illegal call to update() with reordering tracking disabled.");
        const double t0 = wall_ms();
        int rc = 0;
        const bool done = pf ? pf->update(entries, count, &rc) : pd->update(entries, count, &rc);
        SAP_HIP_CHECK(hipStreamSynchronize(stream));
        if (!done) return 0;
        if (rc != 0) {
            setup_done = false;
            return rc;
        }
        stats.timeUpdate = wall_ms() - t0;
        auto fill = [&](const auto& p) {
            stats.time_reorder = 0;
            stats.time_cpu_assemble = p.time_cpu_assemble;
            stats.time_transfer = p.time_transfer;
            stats.time_toBanded = p.time_toBanded;
            stats.time_offDiags = p.time_offDiags;
            stats.time_bandLU = p.time_bandLU;
            stats.time_bandUL = p.time_bandUL;
            stats.time_assembly = p.time_assembly;
            stats.time_fullLU = p.time_fullLU;
        };
        if (pf) fill(*pf); else fill(*pd);
        // the built-in SpMV follows the updated values (non-owning, as at setup)
        if (d_rp) d_val = entries;
        else d_band = entries;
        return 1;
    }

    // ------------------------------------------ batched solve (nrhs > 1)
    // The operators of one system for nr columns (column c at c * n).
    template <class PC>
    struct OneSystemOps : BatchOps {
        Solver& s;
        PC& pc;
        int nr;
        OneSystemOps(Solver& s_, PC& p, int r) : s(s_), pc(p), nr(r) {}
        void apply(const double* v, double* z, const int* gate) override
        {
            pc.apply_multi(v, s.n, z, s.n, nr, gate);
        }
        void spmv(const double* x, double* y, const int* gate) override
        {
            ProfScope ps(&s.prof, "spmv");
            if (s.d_rp) launch_csr_spmm(s.n, s.nnz, s.d_rp, s.d_ci, s.d_val, x, s.n, y, s.n, nr, gate, s.stream);
            else launch_band_spmm(s.n, s.k, s.d_band, x, s.n, y, s.n, nr, gate, s.stream);
        }
    };
    std::unique_ptr<BatchKrylov> batch_;
    int64_t batch_n_ = 0;

    bool can_batch(sap_spmv_fn fn, int nrhs) const
    {
        return nrhs > 1 && opt.batch && !fn && opt.solverType == SAP_BiCGStab2;
    }

    static const char* bicgstab_message(int c, bool refined)
    {
        switch (c) {
        case 1: return "Converged";
        case -1: return refined ? "Maximum number of refinement steps was reached"
                                : "Maximum number of iterations was reached";
        case -2: return "Residual norm is NaN";
        case -3: return "Stagnation encountered";
        case -10: return "rho0 is zero";
        case -11: return "gamma is zero";
        case -12: return "a sigma value is zero";
        default: return "";
        }
    }

    // A HIP graph cannot be captured on the null stream: a batched solve
    // with the graph on a solver that uses the null stream runs on a stream
    // of its own, after the work already queued on the null stream (the
    // solve ends with a synchronization, so later work is ordered after it).
    hipStream_t own_stream_ = nullptr;
    struct StreamSwap {
        Solver& s;
        hipStream_t saved;
        bool swapped = false;
        StreamSwap(Solver& s_, bool need) : s(s_), saved(s_.stream)
        {
            if (!need || saved != nullptr) return;
            if (!s.own_stream_) SAP_HIP_CHECK(hipStreamCreateWithFlags(&s.own_stream_, hipStreamNonBlocking));
            SAP_HIP_CHECK(hipStreamSynchronize(nullptr));
            s.set_stream(s.own_stream_);
            swapped = true;
        }
        ~StreamSwap()
        {
            if (swapped) s.set_stream(saved);
        }
    };

    // All nr columns of b / x (leading dimension n) solved together.
    int solve_batch(const double* d_b, double* d_x, int nr)
    {
        if (!setup_done) return SAP_ERR_ILLEGAL_SOLVE;
        StreamSwap swap(*this, opt.graph != 0 && !opt.profile);
        user_spmv = nullptr;
        stats.numApplies = stats.numSpmv = stats.numRefine = 0;
        if (!batch_ || batch_->columns() != nr || batch_n_ != n || batch_->stream() != stream) {
            batch_.reset();
            std::vector<int64_t> off(nr);
            std::vector<int> len(nr, (int)n);
            for (int c = 0; c < nr; ++c) off[c] = (int64_t)c * n;
            batch_ = std::make_unique<BatchKrylov>(stream, off, len);
            batch_n_ = n;
        }
        // the preconditioner's work buffers exist before any graph capture
        if (pf) pf->reserve_multi(nr);
        else pd->reserve_multi(nr);
        BatchSettings bs;
        bs.relTol = opt.relTol;
        bs.absTol = opt.absTol;
        bs.refineTol = opt.refineTol;
        bs.maxIt = opt.maxNumIterations;
        bs.refine = opt.refine;
        bs.graph = opt.graph != 0 && !opt.profile;
        const double t0 = wall_ms();
        std::vector<BatchResult> res;
        if (pf) {
            OneSystemOps<Precond<float>> ops(*this, *pf, nr);
            batch_->solve(ops, d_b, d_x, bs, res);
        } else {
            OneSystemOps<Precond<double>> ops(*this, *pd, nr);
            batch_->solve(ops, d_b, d_x, bs, res);
        }
        stats.timeSolve = wall_ms() - t0;
        float its = 0.f;
        double rel = 0.0;
        int first_bad = -1;
        for (int c = 0; c < nr; ++c) {
            its = std::max(its, res[c].iterations);
            const double rr = res[c].rNorm / res[c].rhsNorm;
            if (c == 0 || rr > rel || std::isnan(rr)) {
                rel = rr;
                stats.residualNorm = res[c].rNorm;
                stats.rhsNorm = res[c].rhsNorm;
            }
            if (res[c].code <= 0 && first_bad < 0) first_bad = c;
        }
        const BatchResult& pick = res[first_bad >= 0 ? first_bad : nr - 1];
        code = pick.code;
        message = bicgstab_message(code, opt.refine > 0);
        bmon.code = code;
        bmon.message = message;
        stats.relResidualNorm = rel;
        stats.numIterations = its;
        return first_bad < 0 ? 1 : 0;
    }

    int solve(sap_spmv_fn fn, void* ctx, const double* d_b, double* d_x)
    {
        if (!setup_done) return SAP_ERR_ILLEGAL_SOLVE;
        user_spmv = fn;
        user_ctx = ctx;
        stats.numApplies = stats.numSpmv = stats.numRefine = 0;
        // b, x, three refinement vectors, and the Krylov method's vectors
        // (BiCGStab(2): 19 including the non-converged tail; MINRES: 8)
        reserve_vectors(32);
        double* bv = vec();
        double* xv = vec();
        copy(d_b, bv);            // SolverVector b_vector = b (sap/solver.h:488-489)
        copy(d_x, xv);
        const double t0 = wall_ms();
        float its = 0.f;
        double rn = 0, bn = 0;
        if (opt.refine > 0) {
            // Outer FP64 refinement: x += d, A d = b - A x solved by the
            // preconditioned Krylov method to refineTol.  With an FP32
            // preconditioner the right-preconditioned BiCGStab(l) cannot go
            // much below 1e-7 in relative residual on its own.
            double* r = vec();
            double* Ax = vec();
            double* dv = vec();
            const size_t mark = vec_next_;
            double bx[2];
            {
                const double* vs[2] = {bv, xv};
                nrm2s(2, vs, bx);
            }
            const double b_dot = hscalar_[0];   // (b, b) as dot() gives it
            bn = bx[0];
            const double tol = opt.absTol + opt.relTol * bn;
            // r = b - A x.  A*0 = 0 exactly, so a zero initial guess needs no
            // SpMV; after an inner solve that converged, r is that solve's
            // own b - A d (the residual of x + d up to FP64 rounding of the
            // update), and convergence is only declared on b - A x formed
            // afresh.  After an inner solve that did not converge, b - A x
            // is formed afresh.
            bool have_r = bx[1] == 0.0;
            if (have_r) copy(bv, r);
            double r_known = bn;   // |r| while have_r (the norm nrm2 gives for it)
            double r_dot = b_dot;  // and (r, r) as dot() gives it
            bool recurred = false;
            code = -1;
            for (int outer = 0;; ++outer) {
                if (!have_r) {
                    spmv(xv, Ax);
                    axpby(bv, Ax, r, 1.0, -1.0);
                    recurred = false;
                    rn = nrm2(r);
                    r_dot = hscalar_[0];
                } else {
                    rn = r_known;
                }
                if (!std::isnan(rn) && rn <= tol && recurred) {
                    spmv(xv, Ax);
                    axpby(bv, Ax, r, 1.0, -1.0);
                    recurred = false;
                    rn = nrm2(r);
                    r_dot = hscalar_[0];
                }
                have_r = false;
                if (std::isnan(rn)) { code = -2; break; }
                if (rn <= tol) { code = 1; break; }
                if (outer >= opt.refine) { code = -1; break; }
                const double rt = std::max(opt.refineTol, 0.5 * tol / rn);
                vec_next_ = mark;
                double inner_r, inner_b;
                const int c = krylov(dv, r, rt, 0.0, true, true, &its, &inner_r, &inner_b, rn, r_dot);
                stats.numRefine++;
                axpy(dv, xv, 1.0);
                if (c == -2) { code = -2; break; }
                if (last_res_) {
                    copy(last_res_, r);
                    have_r = recurred = true;
                    r_known = inner_r;   // the monitor's norm of that residual
                    r_dot = last_res_dot_;
                }
            }
            message = code == 1 ? "Converged"
                      : code == -2 ? "Residual norm is NaN" : "Maximum number of refinement steps was reached";
            bmon.code = mmon.code = code;
            bmon.message = mmon.message = message;
        } else {
            code = krylov(xv, bv, opt.relTol, opt.absTol, false, false, &its, &rn, &bn);
        }
        copy(xv, d_x);            // thrust::copy(x_vector -> x) (sap/solver.h:534)
        SAP_HIP_CHECK(hipStreamSynchronize(stream));
        stats.timeSolve = wall_ms() - t0;
        stats.rhsNorm = bn;
        stats.residualNorm = rn;
        stats.relResidualNorm = rn / bn;
        stats.numIterations = its;
        return code > 0 ? 1 : 0;
    }
};

// ---------------------------------------------------------- many systems
// Independent banded systems, each with its own preconditioner (its own
// Solver, setup as sap_setup_banded does), whose right-hand sides are solved
// together by one batched BiCGStab(2): the columns of all systems are the
// columns of one batch, every Krylov vector kernel covers all of them in
// one launch, and each iteration is one graph replay.
class SystemBatch {
public:
    sap_options opt;
    hipStream_t stream = nullptr;
    std::vector<std::unique_ptr<Solver>> sys;
    std::vector<BatchResult> res;        // per column of the last solve
    std::vector<int> col0;               // first column of each system
    double timeSetup = 0, timeSolve = 0;

    SystemBatch(int nsys, const int* P, const sap_options& o) : opt(o)
    {
        for (int s = 0; s < nsys; ++s) sys.push_back(std::make_unique<Solver>(P[s], o));
    }
    void set_stream(hipStream_t st)
    {
        null_user_ = false;
        switch_stream(st);
    }
    void switch_stream(hipStream_t st)
    {
        if (st == stream) return;
        stream = st;
        for (auto& p : sys) p->set_stream(st);
        kry_.reset();
    }
    int setup_banded(const int* n, const int* k, const double* const* bands)
    {
        const double t0 = wall_ms();
        for (size_t s = 0; s < sys.size(); ++s) {
            const int rc = sys[s]->setup_banded(n[s], k[s], bands[s]);
            if (rc != 0) return rc;
        }
        timeSetup = wall_ms() - t0;
        return 0;
    }

    struct Ops : BatchOps {
        SystemBatch& b;
        explicit Ops(SystemBatch& b_) : b(b_) {}
        void apply(const double* v, double* z, const int* gate) override
        {
            if (b.fused_jobs_ > 0) {
                // every system's sweeps in one launch
                ProfScope ps(&b.sys[0]->prof, "sweep_full");
                if (b.opt.precValueType == SAP_PREC_DOUBLE)
                    launch_sweep_full_jobs<double>(b.fused_.get(), b.fused_jobs_, b.fused_lds_, v, z, gate, b.stream);
                else
                    launch_sweep_full_jobs<float>(b.fused_.get(), b.fused_jobs_, b.fused_lds_, v, z, gate, b.stream);
                return;
            }
            for (size_t s = 0; s < b.sys.size(); ++s) {
                Solver& q = *b.sys[s];
                const int64_t o = b.base_[s];
                const int* gs = gate ? gate + b.col0[s] : nullptr;
                if (q.pf) q.pf->apply_multi(v + o, q.n, z + o, q.n, b.nr_[s], gs);
                else q.pd->apply_multi(v + o, q.n, z + o, q.n, b.nr_[s], gs);
            }
        }
        void spmv(const double* x, double* y, const int* gate) override
        {
            for (size_t s = 0; s < b.sys.size(); ++s) {
                Solver& q = *b.sys[s];
                const int64_t o = b.base_[s];
                launch_band_spmm(q.n, q.k, q.d_band, x + o, q.n, y + o, q.n, b.nr_[s],
                                 gate ? gate + b.col0[s] : nullptr, b.stream);
            }
        }
    };

    // B[s], X[s]: column-major n_s x nrhs[s] right-hand sides and initial
    // guesses / solutions of system s.  Returns 1 when every column
    // converged.
    int solve(const double* const* B, double* const* X, const int* nrhs)
    {
        for (auto& p : sys)
            if (!p->setup_done) return SAP_ERR_ILLEGAL_SOLVE;
        // a graph cannot be captured on the null stream (see Solver): with
        // the caller on the null stream the solves run on a stream of their
        // own, after the work queued on the null stream so far
        if (opt.graph && !opt.profile && (stream == nullptr || null_user_)) {
            if (!own_stream_) SAP_HIP_CHECK(hipStreamCreateWithFlags(&own_stream_, hipStreamNonBlocking));
            SAP_HIP_CHECK(hipStreamSynchronize(nullptr));
            switch_stream(own_stream_);
            null_user_ = true;
        }
        const int S = (int)sys.size();
        bool same = kry_ && (int)nr_.size() == S;
        for (int s = 0; same && s < S; ++s) same = nr_[s] == nrhs[s] && n_[s] == sys[s]->n;
        if (!same) {
            kry_.reset();
            nr_.assign(nrhs, nrhs + S);
            n_.resize(S);
            base_.resize(S);
            col0.resize(S);
            std::vector<int64_t> off;
            std::vector<int> len;
            int64_t at = 0;
            for (int s = 0; s < S; ++s) {
                n_[s] = sys[s]->n;
                base_[s] = at;
                col0[s] = (int)off.size();
                for (int r = 0; r < nr_[s]; ++r) {
                    off.push_back(at);
                    len.push_back((int)n_[s]);
                    at += n_[s];
                }
            }
            kry_ = std::make_unique<BatchKrylov>(stream, off, len);
            Bb_.resize(at);
            Xb_.resize(at);
        }
        for (int s = 0; s < S; ++s) {
            Solver& q = *sys[s];
            if (q.pf) q.pf->reserve_multi(nr_[s]);
            else q.pd->reserve_multi(nr_[s]);
        }
        // one sweep launch for all systems when every system's apply is a
        // single full sweep pass (the job table depends on the setup's
        // buffers, so it is built for every solve)
        fused_jobs_ = 0;
        {
            std::vector<SweepJob> jobs;
            size_t lds = 0;
            bool all = opt.fuse != 0;
            for (int s = 0; all && s < S; ++s) {
                Solver& q = *sys[s];
                all = q.pf ? q.pf->append_sweep_jobs(jobs, lds, base_[s], base_[s], nr_[s], col0[s])
                           : q.pd->append_sweep_jobs(jobs, lds, base_[s], base_[s], nr_[s], col0[s]);
            }
            if (all && !jobs.empty()) {
                fused_.resize(jobs.size());
                SAP_HIP_CHECK(hipMemcpyAsync(fused_.get(), jobs.data(), sizeof(SweepJob) * jobs.size(),
                                             hipMemcpyHostToDevice, stream));
                SAP_HIP_CHECK(hipStreamSynchronize(stream));
                fused_jobs_ = (int)jobs.size();
                fused_lds_ = lds;
            }
        }
        const double t0 = wall_ms();
        // batch_ is (re)made on the stream the solve runs on
        for (int s = 0; s < S; ++s) {
            const size_t bytes = sizeof(double) * (size_t)n_[s] * nr_[s];
            SAP_HIP_CHECK(hipMemcpyAsync(Bb_.get() + base_[s], B[s], bytes, hipMemcpyDeviceToDevice, stream));
            SAP_HIP_CHECK(hipMemcpyAsync(Xb_.get() + base_[s], X[s], bytes, hipMemcpyDeviceToDevice, stream));
        }
        BatchSettings bs;
        bs.relTol = opt.relTol;
        bs.absTol = opt.absTol;
        bs.refineTol = opt.refineTol;
        bs.maxIt = opt.maxNumIterations;
        bs.refine = opt.refine;
        bs.graph = opt.graph != 0 && !opt.profile;
        Ops ops(*this);
        kry_->solve(ops, Bb_.get(), Xb_.get(), bs, res);
        for (int s = 0; s < S; ++s)
            SAP_HIP_CHECK(hipMemcpyAsync(X[s], Xb_.get() + base_[s], sizeof(double) * (size_t)n_[s] * nr_[s],
                                         hipMemcpyDeviceToDevice, stream));
        SAP_HIP_CHECK(hipStreamSynchronize(stream));
        timeSolve = wall_ms() - t0;
        for (const auto& r : res)
            if (r.code <= 0) return 0;
        return 1;
    }

    ~SystemBatch()
    {
        kry_.reset();
        sys.clear();
        if (own_stream_) (void)hipStreamDestroy(own_stream_);
    }

private:
    hipStream_t own_stream_ = nullptr;
    bool null_user_ = false;   // the caller's stream is the null stream; the solves run on own_stream_
    DevBuf<SweepJob> fused_;   // the fused sweep's job table (fused_jobs_ > 0)
    int fused_jobs_ = 0;
    size_t fused_lds_ = 0;
    std::unique_ptr<BatchKrylov> kry_;
    std::vector<int> nr_;
    std::vector<int64_t> n_, base_;
    DevBuf<double> Bb_, Xb_;
};

}  // namespace sap_amd

// ====================================================================== C ABI
using sap_amd::Error;
using sap_amd::Solver;

struct sap_solver {
    std::unique_ptr<Solver> s;
    std::string err;
};

struct sap_batch {
    std::unique_ptr<sap_amd::SystemBatch> b;
    std::string err;
};

// ------------------------------------------- allocation-free solve: scratch
// The Krylov vector pool (kSolveVectors vectors of up to SAP_NOALLOC_MAX_N
// doubles) and the reduction partials of sap_solve_noalloc live in this
// array, one copy per GPU for the life of the process (32 MiB + 32 KiB).
// Host code never passes the array's name to a kernel: it is a host-side
// stand-in; the device address comes from hipGetSymbolAddress, once per
// device, and is kept in g_scratch_ptr.
constexpr int kSolveVectors = 32;           // what Solver::solve reserves
constexpr size_t kScratchPartials = 4096;   // >= dot_partials_size()
constexpr size_t kScratchDoubles = (size_t)kSolveVectors * SAP_NOALLOC_MAX_N + kScratchPartials;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of the scratch array (looked up on the first
// call on each device).
double* scratch_on_current_device()
{
    int dev = 0;
    SAP_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error(SAP_ERR_UNSUPPORTED, "device ordinal beyond 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        SAP_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    return g_scratch_ptr[dev];
}

// Points the solver's vector pool and reduction partials at the scratch
// array for the lifetime of the object.
struct ScratchPool {
    Solver& s;
    ScratchPool(Solver& s_, double* base) : s(s_)
    {
        s.ext_work_ = base;
        s.ext_work_size_ = (size_t)kSolveVectors * SAP_NOALLOC_MAX_N;
        s.ext_partials_ = base + s.ext_work_size_;
    }
    ~ScratchPool()
    {
        s.ext_work_ = nullptr;
        s.ext_work_size_ = 0;
        s.ext_partials_ = nullptr;
    }
};

template <class F>
int guarded(sap_solver* h, F&& f)
{
    try {
        return f();
    } catch (const Error& e) {
        if (h) h->err = e.what();
        return e.code;
    } catch (const std::bad_alloc& e) {
        if (h) h->err = e.what();
        return SAP_ERR_OOM;
    } catch (const std::exception& e) {
        if (h) h->err = e.what();
        return SAP_ERR_HIP;
    }
}

template <class F>
int guarded_batch(sap_batch* h, F&& f)
{
    try {
        return f();
    } catch (const Error& e) {
        h->err = e.what();
        return e.code;
    } catch (const std::bad_alloc& e) {
        h->err = e.what();
        return SAP_ERR_OOM;
    } catch (const std::exception& e) {
        h->err = e.what();
        return SAP_ERR_HIP;
    }
}

int check_options(const sap_options* o, std::string& why)
{
    if (o->gpuCount != 1) { why = "gpuCount > 1 is not supported (one GPU)"; return SAP_ERR_UNSUPPORTED; }
    if (o->useBCR && o->isSPD) { why = "BCR does not support SPD matrices"; return SAP_ERR_UNSUPPORTED; }
    if (o->ilu_level >= 0) { why = "ILU preconditioning is not supported"; return SAP_ERR_UNSUPPORTED; }
    if (o->saveMem) { why = "saveMem (SPD half-band storage) is not supported"; return SAP_ERR_UNSUPPORTED; }
    if (o->solverType != SAP_BiCGStab1 && o->solverType != SAP_BiCGStab2 && o->solverType != SAP_MINRES) {
        why = "only BiCGStab1, BiCGStab2 and MINRES are provided";
        return SAP_ERR_UNSUPPORTED;
    }
    if (o->factMethod != SAP_LU_UL && o->factMethod != SAP_LU_only) { why = "bad factMethod"; return SAP_ERR_BADARG; }
    if (o->precondType < SAP_Spike || o->precondType > SAP_None) { why = "bad precondType"; return SAP_ERR_BADARG; }
    if (o->precValueType != SAP_PREC_FLOAT && o->precValueType != SAP_PREC_DOUBLE) {
        why = "bad precValueType";
        return SAP_ERR_BADARG;
    }
    return SAP_OK;
}

}  // namespace

extern "C" {

void sap_default_options(sap_options* o)
{
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    // sap::Options::Options() (sap/solver.h:232-256)
    o->solverType = SAP_BiCGStab2;
    o->maxNumIterations = 100;
    o->gpuCount = 1;
    o->relTol = 1e-6;
    o->absTol = 0;
    o->testDB = 0;
    o->isSPD = 0;
    o->saveMem = 0;
    o->performReorder = 1;
    o->performDB = 1;
    o->dbFirstStageOnly = 0;
    o->applyScaling = 1;
    o->maxBandwidth = INT_MAX;
    o->dropOffFraction = 0;
    o->factMethod = SAP_LU_only;
    o->precondType = SAP_Spike;
    o->safeFactorization = 0;
    o->variableBandwidth = 1;
    o->trackReordering = 0;
    o->useBCR = 0;
    o->ilu_level = -1;
    o->precValueType = SAP_PREC_FLOAT;
    o->refine = 0;
    o->refineTol = 1e-6;
    o->profile = 0;
    o->batch = 1;
    o->graph = 1;
    o->fuse = 1;
    o->autotune = 0;
}

int sap_stream_read_ms(const double* d_x, int64_t n, int reps, double* best_ms)
{
    if (!d_x || n < 2 || (n & 1) || reps < 1 || !best_ms || ((uintptr_t)d_x & 15)) return SAP_ERR_BADARG;
    return guarded(nullptr, [&] {
        *best_ms = sap_amd::stream_read_best_ms(d_x, n, reps, nullptr);
        return 0;
    });
}

int sap_autotune_entries(sap_tune_entry* out, int max_entries)
{
    const auto e = sap_amd::Autotuner::instance().entries();
    for (int i = 0; out && i < max_entries && i < (int)e.size(); ++i) {
        out[i].family = e[i].family;
        out[i].k = e[i].k;
        out[i].elem_bytes = e[i].elem;
        out[i].value = e[i].value;
        out[i].default_value = e[i].default_value;
        out[i].candidates = e[i].candidates;
        out[i].best_ms = e[i].best_ms;
        out[i].default_ms = e[i].default_ms;
    }
    return (int)e.size();
}

sap_solver* sap_create(int numPartitions, const sap_options* opts)
{
    sap_options o;
    if (opts) o = *opts;
    else sap_default_options(&o);
    std::string why;
    if (numPartitions < 1 || check_options(&o, why) != SAP_OK) return nullptr;
    auto* h = new (std::nothrow) sap_solver();
    if (!h) return nullptr;
    const int rc = guarded(h, [&] {
        h->s = std::make_unique<Solver>(numPartitions, o);
        return 0;
    });
    if (rc != 0) {
        delete h;
        return nullptr;
    }
    return h;
}

void sap_destroy(sap_solver* h) { delete h; }

int sap_set_stream(sap_solver* h, void* stream)
{
    if (!h) return SAP_ERR_BADARG;
    h->s->set_stream((hipStream_t)stream);
    return SAP_OK;
}

int sap_setup_banded(sap_solver* h, int n, int k, const double* d_band)
{
    if (!h) return SAP_ERR_BADARG;
    if (n <= 0 || k < 0 || !d_band) { h->err = "bad matrix arguments"; return SAP_ERR_BADARG; }
    return guarded(h, [&] { return h->s->setup_banded(n, k, d_band); });
}

int sap_setup_csr(sap_solver* h, int n, int nnz, const int* d_row_offsets, const int* d_col_indices,
                  const double* d_values)
{
    if (!h) return SAP_ERR_BADARG;
    if (n <= 0 || nnz < 0 || !d_row_offsets || (nnz > 0 && (!d_col_indices || !d_values))) {
        h->err = "bad matrix arguments";
        return SAP_ERR_BADARG;
    }
    return guarded(h, [&] { return h->s->setup_csr(n, nnz, d_row_offsets, d_col_indices, d_values); });
}

int sap_update(sap_solver* h, const double* d_entries, int64_t count)
{
    if (!h) return SAP_ERR_BADARG;
    if (!d_entries || count < 0) { h->err = "bad update arguments"; return SAP_ERR_BADARG; }
    return guarded(h, [&] { return h->s->update(d_entries, count); });
}

int sap_solve(sap_solver* h, sap_spmv_fn spmv, void* ctx, const double* d_b, double* d_x, int nrhs)
{
    if (!h) return SAP_ERR_BADARG;
    if (!d_b || !d_x || nrhs < 1) { h->err = "bad solve arguments"; return SAP_ERR_BADARG; }
    return guarded(h, [&] {
        if (h->s->can_batch(spmv, nrhs)) return h->s->solve_batch(d_b, d_x, nrhs);
        int all = 1;
        for (int c = 0; c < nrhs; ++c) {
            const int r = h->s->solve(spmv, ctx, d_b + (size_t)c * h->s->n, d_x + (size_t)c * h->s->n);
            if (r < 0) return r;
            all &= r;
        }
        return all;
    });
}

int sap_solve_noalloc(sap_solver* h, sap_spmv_fn spmv, void* ctx, const double* d_b, double* d_x, int nrhs)
{
    if (!h) return SAP_ERR_BADARG;
    if (!d_b || !d_x || nrhs < 1) { h->err = "bad solve arguments"; return SAP_ERR_BADARG; }
    Solver& s = *h->s;
    if (!s.setup_done) return s.solve(spmv, ctx, d_b, d_x);   // Illegal_solve, as sap_solve
    if (s.n > SAP_NOALLOC_MAX_N) {
        h->err = "sap_solve_noalloc: n = " + std::to_string(s.n) + " exceeds SAP_NOALLOC_MAX_N = " +
                 std::to_string(SAP_NOALLOC_MAX_N) + " (use sap_solve)";
        return SAP_ERR_BADARG;
    }
    if (s.opt.profile) {
        h->err = "sap_solve_noalloc: profiling creates HIP events; use sap_solve";
        return SAP_ERR_UNSUPPORTED;
    }
    if ((size_t)sap_amd::dot_partials_size() > kScratchPartials) {
        h->err = "sap_solve_noalloc: reduction partials exceed the scratch array";
        return SAP_ERR_HIP;
    }
    return guarded(h, [&] {
        ScratchPool pool(s, scratch_on_current_device());
        int all = 1;
        for (int c = 0; c < nrhs; ++c) {
            const int r = s.solve(spmv, ctx, d_b + (size_t)c * s.n, d_x + (size_t)c * s.n);
            if (r < 0) return r;
            all &= r;
        }
        return all;
    });
}

sap_batch* sap_batch_create(int nsys, const int* numPartitions, const sap_options* opts)
{
    if (nsys < 1 || !numPartitions) return nullptr;
    sap_options o;
    if (opts) o = *opts;
    else sap_default_options(&o);
    std::string why;
    if (check_options(&o, why) != SAP_OK || o.solverType != SAP_BiCGStab2) return nullptr;
    for (int s = 0; s < nsys; ++s)
        if (numPartitions[s] < 1) return nullptr;
    auto* h = new (std::nothrow) sap_batch();
    if (!h) return nullptr;
    try {
        h->b = std::make_unique<sap_amd::SystemBatch>(nsys, numPartitions, o);
    } catch (...) {
        delete h;
        return nullptr;
    }
    return h;
}

void sap_batch_destroy(sap_batch* h) { delete h; }

int sap_batch_set_stream(sap_batch* h, void* stream)
{
    if (!h) return SAP_ERR_BADARG;
    h->b->set_stream((hipStream_t)stream);
    return SAP_OK;
}

int sap_batch_setup_banded(sap_batch* h, const int* n, const int* k, const double* const* d_bands)
{
    if (!h || !n || !k || !d_bands) return SAP_ERR_BADARG;
    for (size_t s = 0; s < h->b->sys.size(); ++s)
        if (n[s] <= 0 || k[s] < 0 || !d_bands[s]) { h->err = "bad matrix arguments"; return SAP_ERR_BADARG; }
    return guarded_batch(h, [&] { return h->b->setup_banded(n, k, d_bands); });
}

int sap_batch_solve(sap_batch* h, const double* const* d_b, double* const* d_x, const int* nrhs)
{
    if (!h || !d_b || !d_x || !nrhs) return SAP_ERR_BADARG;
    for (size_t s = 0; s < h->b->sys.size(); ++s)
        if (nrhs[s] < 1 || !d_b[s] || !d_x[s]) { h->err = "bad solve arguments"; return SAP_ERR_BADARG; }
    return guarded_batch(h, [&] { return h->b->solve(d_b, d_x, nrhs); });
}

int sap_batch_result(const sap_batch* h, int column, int* code, float* iterations, double* rel_residual)
{
    if (!h || column < 0 || column >= (int)h->b->res.size()) return SAP_ERR_BADARG;
    const auto& r = h->b->res[column];
    if (code) *code = r.code;
    if (iterations) *iterations = r.iterations;
    if (rel_residual) *rel_residual = r.rNorm / r.rhsNorm;
    return SAP_OK;
}

int sap_batch_times(const sap_batch* h, double* setup_ms, double* solve_ms)
{
    if (!h) return SAP_ERR_BADARG;
    if (setup_ms) *setup_ms = h->b->timeSetup;
    if (solve_ms) *solve_ms = h->b->timeSolve;
    return SAP_OK;
}

const char* sap_batch_last_error(const sap_batch* h) { return h ? h->err.c_str() : ""; }

int sap_get_stats(const sap_solver* h, sap_stats* out)
{
    if (!h || !out) return SAP_ERR_BADARG;
    const int rc = guarded(const_cast<sap_solver*>(h), [&] {
        h->s->settle();
        return 0;
    });
    if (rc != 0) return rc;
    *out = h->s->stats;
    return SAP_OK;
}

int sap_monitor_code(const sap_solver* h) { return h ? h->s->code : 0; }
const char* sap_monitor_message(const sap_solver* h) { return h ? h->s->message.c_str() : ""; }
const char* sap_last_error(const sap_solver* h) { return h ? h->err.c_str() : ""; }

int sap_precond_apply(sap_solver* h, const double* d_v, double* d_z)
{
    if (!h) return SAP_ERR_BADARG;
    if (!h->s->setup_done) return SAP_ERR_ILLEGAL_SOLVE;
    return guarded(h, [&] {
        h->s->prec(d_v, d_z);
        SAP_HIP_CHECK(hipStreamSynchronize(h->s->stream));
        return 0;
    });
}

int sap_band_spmv(int n, int k, const double* d_band, const double* d_x, double* d_y, void* stream)
{
    if (n <= 0 || k < 0 || !d_band || !d_x || !d_y) return SAP_ERR_BADARG;
    return guarded(nullptr, [&] {
        sap_amd::launch_band_spmv(n, k, d_band, d_x, d_y, (hipStream_t)stream);
        return 0;
    });
}

int sap_csr_spmv(int n, int nnz, const int* d_row_offsets, const int* d_col_indices, const double* d_values,
                 const double* d_x, double* d_y, void* stream)
{
    if (n <= 0 || nnz < 0 || !d_row_offsets || !d_x || !d_y || (nnz > 0 && (!d_col_indices || !d_values)))
        return SAP_ERR_BADARG;
    return guarded(nullptr, [&] {
        sap_amd::launch_csr_spmv(n, nnz, d_row_offsets, d_col_indices, d_values, d_x, d_y, (hipStream_t)stream);
        return 0;
    });
}

int sap_get_profile(const sap_solver* h, sap_kernel_profile* out, int max_entries)
{
    if (!h) return SAP_ERR_BADARG;
    auto& p = const_cast<sap_amd::Profiler&>(h->s->prof);
    return guarded(const_cast<sap_solver*>(h), [&] {
        p.flush();
        const int m = std::min<int>(max_entries, (int)p.entries.size());
        for (int i = 0; i < m && out; ++i) {
            out[i].name = p.entries[i].name.c_str();
            out[i].launches = p.entries[i].launches;
            out[i].total_ms = p.entries[i].ms;
        }
        return (int)p.entries.size();
    });
}

void sap_reset_profile(sap_solver* h)
{
    if (h) guarded(h, [&] { h->s->prof.reset(); return 0; });
}

int sap_num_partitions(const sap_solver* h) { return h ? h->s->num_partitions() : 0; }

int sap_get_factors(const sap_solver* h, double* h_out, int64_t count)
{
    if (!h || !h_out) return SAP_ERR_BADARG;
    Solver& S = *h->s;
    if (!S.setup_done) return SAP_ERR_ILLEGAL_SOLVE;
    return guarded(const_cast<sap_solver*>(h), [&] {
        S.settle();
        auto grab = [&](auto& p) {
            using T = typename std::remove_reference<decltype(*p.factors().get())>::type;
            if (p.variable_mode()) {
                // the partitions' own bands, second-level order, one after
                // the other (the device copy pads each to an aligned start)
                const std::vector<int64_t>& off = p.band_offsets();
                const std::vector<int>& ks = p.band_widths();
                sap_amd::PartInfo pi{S.n, p.P};
                int64_t total = 0;
                for (int q = 0; q < p.P; ++q) {
                    int64_t s, len;
                    pi.bounds(q, s, len);
                    total += len * (2 * (int64_t)ks[q] + 1);
                }
                if (count < total) throw Error(SAP_ERR_BADARG, "output too small");
                std::vector<T> tmp((size_t)off.back());
                SAP_HIP_CHECK(hipMemcpy(tmp.data(), p.factors().get(), sizeof(T) * tmp.size(), hipMemcpyDeviceToHost));
                int64_t o = 0;
                for (int q = 0; q < p.P; ++q) {
                    int64_t s, len;
                    pi.bounds(q, s, len);
                    const int64_t m = len * (2 * (int64_t)ks[q] + 1);
                    for (int64_t i = 0; i < m; ++i) h_out[o + i] = (double)tmp[off[q] + i];
                    o += m;
                }
                return;
            }
            const int64_t w = 2 * (int64_t)S.k + 1;
            if (count < w * S.n) throw Error(SAP_ERR_BADARG, "output too small");
            std::vector<T> tmp((size_t)w * S.n);
            SAP_HIP_CHECK(hipMemcpy(tmp.data(), p.factors().get(), sizeof(T) * tmp.size(), hipMemcpyDeviceToHost));
            if (p.spike_active() && p.lu_ul()) {
                // the last partition's factors are its UL factors
                sap_amd::PartInfo pi{S.n, p.P};
                int64_t s, len;
                pi.bounds(p.P - 1, s, len);
                SAP_HIP_CHECK(hipMemcpy(tmp.data() + w * s, p.factors_ul().get() + w * s, sizeof(T) * w * len,
                                        hipMemcpyDeviceToHost));
            }
            for (size_t i = 0; i < tmp.size(); ++i) h_out[i] = (double)tmp[i];
        };
        if (S.pf) grab(*S.pf); else grab(*S.pd);
        return 0;
    });
}

int sap_get_spikes(const sap_solver* h, double* h_wv, double* h_schur, int64_t count)
{
    if (!h) return SAP_ERR_BADARG;
    Solver& S = *h->s;
    if (!S.setup_done) return SAP_ERR_ILLEGAL_SOLVE;
    return guarded(const_cast<sap_solver*>(h), [&] {
        S.settle();
        auto grab = [&](auto& p) {
            using T = typename std::remove_reference<decltype(*p.spikes().get())>::type;
            const size_t nwv = p.spikes().size(), ns = p.schur().size();
            if ((size_t)count < nwv) throw Error(SAP_ERR_BADARG, "output too small");
            std::vector<T> tmp(std::max(nwv, ns));
            if (h_wv && nwv) {
                SAP_HIP_CHECK(hipMemcpy(tmp.data(), p.spikes().get(), sizeof(T) * nwv, hipMemcpyDeviceToHost));
                for (size_t i = 0; i < nwv; ++i) h_wv[i] = (double)tmp[i];
            }
            if (h_schur && ns) {
                SAP_HIP_CHECK(hipMemcpy(tmp.data(), p.schur().get(), sizeof(T) * ns, hipMemcpyDeviceToHost));
                for (size_t i = 0; i < ns; ++i) h_schur[i] = (double)tmp[i];
            }
        };
        if (S.pf) grab(*S.pf); else grab(*S.pd);
        return 0;
    });
}

int sap_gen_banded(int64_t n, int k, double dominance, uint64_t seed, double* d_band, void* stream)
{
    if (n <= 0 || k < 0 || !d_band) return SAP_ERR_BADARG;
    return guarded(nullptr, [&] {
        sap_amd::launch_gen_banded(n, k, dominance, seed, d_band, (hipStream_t)stream);
        return 0;
    });
}

int sap_gen_vector(int64_t n, uint64_t seed, double* d_x, void* stream)
{
    if (n <= 0 || !d_x) return SAP_ERR_BADARG;
    return guarded(nullptr, [&] {
        sap_amd::launch_gen_vector(n, seed, d_x, (hipStream_t)stream);
        return 0;
    });
}

struct sap_reordering {
    int precision = SAP_PREC_FLOAT;
    sap_amd::Reordering<float> f;
    sap_amd::Reordering<double> d;
    bool has_var = false;
    sap_amd::VarBand<float> vf;
    sap_amd::VarBand<double> vd;
};

sap_reordering* sap_reorder_csr(int n, const int* h_row_offsets, const int* h_col_indices, const double* h_values,
                                const sap_options* opts, int* err)
{
    auto fail = [&](int c) {
        if (err) *err = c;
        return (sap_reordering*)nullptr;
    };
    if (n <= 0 || !h_row_offsets || (h_row_offsets[n] > 0 && (!h_col_indices || !h_values)))
        return fail(SAP_ERR_BADARG);
    sap_options o;
    if (opts) o = *opts;
    else sap_default_options(&o);
    if (o.precValueType != SAP_PREC_FLOAT && o.precValueType != SAP_PREC_DOUBLE) return fail(SAP_ERR_BADARG);
    sap_amd::GraphOptions go;
    go.doDB = o.performDB != 0;
    go.scale = o.applyScaling != 0;
    go.dbFirstStageOnly = o.dbFirstStageOnly != 0;
    go.doRCM = o.maxBandwidth > 64;
    go.maxBandwidth = o.maxBandwidth;
    go.dropOffFrac = o.dropOffFraction;
    go.testDB = o.testDB != 0;
    auto* r = new (std::nothrow) sap_reordering();
    if (!r) return fail(SAP_ERR_OOM);
    r->precision = o.precValueType;
    const int rc = guarded(nullptr, [&] {
        if (r->precision == SAP_PREC_FLOAT)
            r->f = sap_amd::reorder_csr<float>(n, h_row_offsets, h_col_indices, h_values, go);
        else
            r->d = sap_amd::reorder_csr<double>(n, h_row_offsets, h_col_indices, h_values, go);
        return 0;
    });
    if (rc != 0) {
        delete r;
        return fail(rc);
    }
    if (err) *err = SAP_OK;
    return r;
}

int sap_reordering_info(const sap_reordering* r, int* k, int* k_db, int* k_reorder, double* drop_off, int* nnz)
{
    if (!r) return SAP_ERR_BADARG;
    auto get = [&](const auto& R) {
        if (k) *k = R.k;
        if (k_db) *k_db = R.k_db;
        if (k_reorder) *k_reorder = R.k_reorder;
        if (drop_off) *drop_off = R.dropoff_actual;
        if (nnz) *nnz = (int)R.ci.size();
    };
    if (r->precision == SAP_PREC_FLOAT) get(r->f); else get(r->d);
    return SAP_OK;
}

int sap_reordering_dominance(const sap_reordering* r, double* diag_dom_ori, double* d_p1_ori, double* diag_dom,
                             double* d_p1)
{
    if (!r) return SAP_ERR_BADARG;
    auto get = [&](const auto& R) {
        if (diag_dom_ori) *diag_dom_ori = R.diag_dom_ori;
        if (d_p1_ori) *d_p1_ori = R.d_p1_ori;
        if (diag_dom) *diag_dom = R.diag_dom;
        if (d_p1) *d_p1 = R.d_p1;
    };
    if (r->precision == SAP_PREC_FLOAT) get(r->f); else get(r->d);
    return SAP_OK;
}

int sap_reordering_get(const sap_reordering* r, int* db_perm, int* row_perm, int* col_reord, double* row_scale,
                       double* col_scale, int* row_offsets, int* col_indices, double* values)
{
    if (!r) return SAP_ERR_BADARG;
    auto get = [&](const auto& R) {
        const size_t n = (size_t)R.n, m = R.ci.size();
        if (db_perm) std::copy(R.db_perm.begin(), R.db_perm.end(), db_perm);
        if (row_perm) std::copy(R.row_perm.begin(), R.row_perm.end(), row_perm);
        if (col_reord) std::copy(R.col_reord.begin(), R.col_reord.end(), col_reord);
        for (size_t i = 0; i < n; ++i) {
            if (row_scale) row_scale[i] = (double)R.row_scale[i];
            if (col_scale) col_scale[i] = (double)R.col_scale[i];
        }
        if (row_offsets) std::copy(R.rp.begin(), R.rp.end(), row_offsets);
        if (col_indices) std::copy(R.ci.begin(), R.ci.end(), col_indices);
        if (values)
            for (size_t e = 0; e < m; ++e) values[e] = (double)R.val[e];
    };
    if (r->precision == SAP_PREC_FLOAT) get(r->f); else get(r->d);
    return SAP_OK;
}

int sap_reordering_assemble_var(sap_reordering* r, int numPartitions)
{
    if (!r || numPartitions < 1) return SAP_ERR_BADARG;
    return guarded(nullptr, [&] {
        auto run = [&](const auto& R, auto& V) {
            // partition clamp of the reordering path (sap/precond.h:1777-1783)
            const int P = std::min(numPartitions, std::max(R.n / (R.k + 1), 1));
            r->has_var = P > 1 && R.k > 0;
            if (r->has_var) V = sap_amd::assemble_var(R.n, R.k, P, R.rp, R.ci, R.val);
            return r->has_var ? P : 1;
        };
        return r->precision == SAP_PREC_FLOAT ? run(r->f, r->vf) : run(r->d, r->vd);
    });
}

int sap_reordering_var_get(const sap_reordering* r, int* ks, int64_t* boffsets, int* first_rows,
                           int* second_reorder, int* second_perm, int* widths_left, int* widths_right,
                           int* perms_left, int* perms_right, double* wv, double* offdiags, int* ks_col,
                           int* ks_row, int* diag_row_offsets, int* diag_col_indices, double* diag_values)
{
    if (!r || !r->has_var) return SAP_ERR_BADARG;
    auto get = [&](const auto& V) {
        auto cp = [](const auto& src, auto* dst) {
            if (dst)
                for (size_t i = 0; i < src.size(); ++i) dst[i] = src[i];
        };
        cp(V.ks, ks);
        cp(V.boffsets, boffsets);
        cp(V.first_rows, first_rows);
        cp(V.second_reorder, second_reorder);
        cp(V.second_perm, second_perm);
        cp(V.widths_left, widths_left);
        cp(V.widths_right, widths_right);
        cp(V.perms_left, perms_left);
        cp(V.perms_right, perms_right);
        cp(V.WV, wv);
        cp(V.offDiags, offdiags);
        cp(V.ks_col, ks_col);
        cp(V.ks_row, ks_row);
        cp(V.drp, diag_row_offsets);
        cp(V.dci, diag_col_indices);
        cp(V.dval, diag_values);
    };
    if (r->precision == SAP_PREC_FLOAT) get(r->vf); else get(r->vd);
    return SAP_OK;
}

void sap_reordering_free(sap_reordering* r) { delete r; }

}  // extern "C"
