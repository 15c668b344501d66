// Batched right-preconditioned BiCGStab(2) (sap/bicgstab2.h:31-354 with the
// BiCGStabLMonitor of sap/monitor.h:153-335) over many columns at once:
// several right-hand sides of one system, or right-hand sides of several
// systems.  Every scalar of the method and the monitor's state live on the
// device, one entry per column; a column that stops is switched off by its
// gate flag and every later kernel skips it.  One outer iteration is a fixed
// sequence of launches (dot products, one-thread-per-column scalar steps,
// gated vector updates, gated preconditioner applications and products), so
// it is captured once into a HIP graph and replayed; the host synchronizes
// once per iteration, to read how many columns are still running.
//
// Each column performs the operations of the single-vector solve
// (Solver::bicgstabl in solver.hip, which follows the reference loop) with
// the same kernels' arithmetic and the scalar steps rounded as the host
// rounds them (no contraction), so a batched solve without refinement gives
// every column the bits of solving it alone.  With refinement (refine > 0)
// each correction solve starts from b - A x formed afresh.
#include "batch.hpp"

#include <cmath>
#include <cstddef>

namespace sap_amd {

namespace {

#define BF(name) ((int)(offsetof(BCol, name) / sizeof(double)))

constexpr int kMaxStag = 3;   // maxStagSteps of BiCGStabLMonitor (sap/monitor.h:255-257)

// BiCGStabLMonitor, per column (rounding as on the host: no contraction)
__device__ double mon_tol(const BCol& m)
{
#pragma clang fp contract(off)
    return m.absTol + m.relTol * m.rhsNorm;
}
__device__ bool mon_finished0(BCol& m)
{
    if (m.code != 0) return true;
    if (m.iterations > (float)m.maxIt) m.code = -1;
    return m.code != 0;
}
__device__ bool mon_finished1(BCol& m, double r)
{
    m.rNorm = r;
    if (m.code != 0) return true;
    if (isnan(r)) m.code = -2;
    else if (r <= mon_tol(m)) m.code = 1;
    else if (m.iterations > (float)m.maxIt) m.code = -1;
    else {
        if (m.stag >= kMaxStag && m.moreSteps == 0) m.stag = 0;
        m.moreSteps++;
        if (m.moreSteps >= m.maxm) m.code = -3;
    }
    return m.code != 0;
}
__device__ bool mon_need_check(BCol& m, double r)
{
    m.rNorm = r;
    if (m.code != 0) return false;
    if (isnan(r)) m.code = -2;
    return m.code == 0 && (r <= mon_tol(m) || m.stag >= kMaxStag || m.moreSteps != 0);
}

__global__ void batch_step_kernel(int step, int C, BCol* __restrict__ cols, BGates g, double* __restrict__ count_out)
{
#pragma clang fp contract(off)
    const double eps = 1e-20;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (step == BS_COUNT) {
        // number of columns still running (inner solve, or outer refinement)
        if (c == 0) {
            int a = 0;
            for (int q = 0; q < C; ++q) a += g.act[q];
            *count_out = (double)a;
        }
        return;
    }
    if (c >= C) return;
    BCol& m = cols[c];
    switch (step) {
    case BS_OUTER_INIT: {   // after (b, b)
        m.bnorm = sqrt(m.dot[0]);
        m.tol_outer = m.absTol0 + m.relTol0 * m.bnorm;
        m.code_outer = 0;
        m.outer = 0;
        m.its_total = 0.f;
        g.oact[c] = 1;
        g.act[c] = 0;
        g.run[c] = 0;
        break;
    }
    case BS_OUTER_CHECK: {   // after r = b - A x and (r, r)
        g.run[c] = 0;
        g.act[c] = 0;
        if (!g.oact[c]) break;
        const double rn = sqrt(m.dot[0]);
        m.rn_outer = rn;
        if (isnan(rn)) { m.code_outer = -2; g.oact[c] = 0; break; }
        if (rn <= m.tol_outer) { m.code_outer = 1; g.oact[c] = 0; break; }
        if (m.outer >= m.refine) { m.code_outer = -1; g.oact[c] = 0; break; }
        m.outer++;
        // the inner solve of A d = r to max(refineTol, tol / (2 |r|)), d = 0
        const double half = 0.5 * m.tol_outer / rn;
        m.relTol = m.refineTol > half ? m.refineTol : half;
        m.absTol = 0.0;
        m.rhsNorm = rn;
        m.r_norm_min = m.r_norm = m.r_norm_act = rn;
        g.act[c] = g.run[c] = 1;
        break;
    }
    case BS_INNER_INIT0: {   // no refinement: after (b, b) and (r0, r0)
        m.relTol = m.relTol0;
        m.absTol = m.absTol0;
        m.rhsNorm = sqrt(m.dot[0]);
        m.r_norm_min = m.r_norm = m.r_norm_act = sqrt(m.dot[1]);
        g.act[c] = g.run[c] = 1;
        g.oact[c] = 0;
        break;
    }
    case BS_INNER_START: {   // reset of the monitor and the method's scalars
        if (!g.act[c]) break;
        m.rho0 = 1.0;
        m.alpha = 0.0;
        m.omega = 1.0;
        m.iterations = 0.f;
        m.stag = m.moreSteps = 0;
        m.code = 0;
        m.rNorm = 0.0;
        g.chk[c] = g.mcp[c] = 0;
        break;
    }
    case BS_ITER_START:
        if (!g.act[c]) break;
        m.rho0 = -m.omega * m.rho0;
        m.iterations += 0.25f;
        break;
    case BS_BETA:
        if (!g.act[c]) break;
        m.rho1 = m.dot[0];
        if (m.rho0 == 0) { m.code = -10; g.act[c] = 0; break; }
        m.beta = m.alpha * m.rho1 / m.rho0;
        m.rho0 = m.rho1;
        break;
    case BS_ALPHA:
        if (!g.act[c]) break;
        m.gam = m.dot[0];
        if (m.gam == 0) { m.code = -11; g.act[c] = 0; break; }
        m.alpha = m.rho0 / m.gam;
        break;
    case BS_NORMS: {   // after |rr0|^2, |uu0|^2, |xx|^2
        if (!g.act[c]) break;
        m.r_norm_act = m.r_norm = sqrt(m.dot[0]);
        if (fabs(m.alpha) * sqrt(m.dot[1]) < eps * sqrt(m.dot[2])) m.stag++;
        else m.stag = 0;
        break;
    }
    case BS_NEED_CHECK:
        g.chk[c] = 0;
        if (!g.act[c]) break;
        if (mon_need_check(m, m.r_norm)) g.chk[c] = 1;
        else if (m.code != 0) g.act[c] = 0;
        break;
    case BS_CHECK_DONE:   // after the true residual's |rr0|^2
        if (!g.chk[c]) break;
        g.chk[c] = 0;
        m.r_norm_act = sqrt(m.dot[0]);
        if (mon_finished1(m, m.r_norm_act)) g.act[c] = 0;
        break;
    case BS_MIN:
        g.mcp[c] = 0;
        if (!g.act[c]) break;
        if (m.r_norm_act < m.r_norm_min) {
            m.r_norm_min = m.r_norm_act;
            g.mcp[c] = 1;
        }
        if (mon_finished0(m)) g.act[c] = 0;
        break;
    case BS_TAO:
        if (!g.act[c]) break;
        m.tao12 = m.dot[0] / m.sigma1;
        break;
    case BS_SIGMA1:
    case BS_SIGMA2: {   // after (rr_j, rr_j), (rr_j, rr0)
        if (!g.act[c]) break;
        const double sg = m.dot[0];
        if (step == BS_SIGMA1) m.sigma1 = sg; else m.sigma2 = sg;
        if (sg == 0) { m.code = -12; g.act[c] = 0; break; }
        if (step == BS_SIGMA1) m.gp1 = m.dot[1] / sg; else m.gp2 = m.dot[1] / sg;
        break;
    }
    case BS_GAMMA:
        if (!g.act[c]) break;
        m.g2 = m.gp2;
        m.omega = m.g2;
        m.g1 = m.gp1;
        m.g1 -= m.tao12 * m.g2;
        m.gpp1 = m.g2;
        break;
    case BS_STAG_G1:     // after |rr0|^2, |xx|^2
    case BS_STAG_GPP1: { // after |rr1|^2, |xx|^2
        if (!g.act[c]) break;
        const double f = step == BS_STAG_G1 ? m.g1 : m.gpp1;
        if (fabs(f) * sqrt(m.dot[0]) < eps * sqrt(m.dot[1])) m.stag++;
        else m.stag = 0;
        break;
    }
    case BS_MR_NORM:     // after |rr0|^2; the quarter step counted first
    case BS_MR2_NORM:
        g.chk[c] = 0;
        if (!g.act[c]) break;
        m.r_norm_act = m.r_norm = sqrt(m.dot[0]);
        if (step == BS_MR_NORM) m.iterations += 0.25f;
        if (mon_need_check(m, m.r_norm)) g.chk[c] = 1;
        else if (m.code != 0) g.act[c] = 0;
        break;
    case BS_INC:
        if (!g.act[c]) break;
        m.iterations += 0.25f;
        break;
    case BS_FINAL_SPLIT:   // columns that ran: converged -> x = M^{-1} xx, else the best of xx and x_min
        g.cnv[c] = g.ncv[c] = 0;
        if (!g.run[c]) break;
        if (m.code > 0) g.cnv[c] = 1;
        else g.ncv[c] = 1;
        break;
    case BS_FINAL_PICK: {   // after |b - A M^{-1} xx|^2, |b - A M^{-1} x_min|^2
        g.fa[c] = g.fb[c] = 0;
        if (!g.ncv[c]) break;
        const double rcn = sqrt(m.dot[0]), rcmn = sqrt(m.dot[1]);
        if (rcn < rcmn) { g.fa[c] = 1; m.rNorm = rcn; }
        else            { g.fb[c] = 1; m.rNorm = rcmn; }
        break;
    }
    case BS_OUTER_POST:   // after x += d
        if (!g.run[c]) break;
        m.its_total += m.iterations;
        if (m.code == -2) { m.code_outer = -2; g.oact[c] = 0; }
        break;
    case BS_OUTER_ACT:    // the running flags of the outer loop for BS_COUNT
        g.act[c] = g.oact[c];
        break;
    default:
        break;
    }
}

}  // namespace

BatchKrylov::BatchKrylov(hipStream_t st, const std::vector<int64_t>& off, const std::vector<int>& len)
    : stream_(st), C_((int)off.size())
{
    if (C_ < 1 || len.size() != off.size()) throw Error(SAP_ERR_BADARG, "batched solve: bad column layout");
    int64_t total = 0, maxlen = 0;
    for (int c = 0; c < C_; ++c) {
        total = std::max<int64_t>(total, off[c] + len[c]);
        maxlen = std::max<int64_t>(maxlen, len[c]);
    }
    total_ = total;
    off_.resize(C_);
    len_.resize(C_);
    SAP_HIP_CHECK(hipMemcpyAsync(off_.get(), off.data(), sizeof(int64_t) * C_, hipMemcpyHostToDevice, st));
    SAP_HIP_CHECK(hipMemcpyAsync(len_.get(), len.data(), sizeof(int) * C_, hipMemcpyHostToDevice, st));
    bc_.off = off_.get();
    bc_.len = len_.get();
    bc_.cols = C_;
    bc_.maxlen = maxlen;
    cols_.resize(C_);
    flags_.resize((size_t)kNumGates * C_);
    int* f = flags_.get();
    g_.act = f;
    g_.chk = f + C_;
    g_.mcp = f + 2 * C_;
    g_.oact = f + 3 * C_;
    g_.run = f + 4 * C_;
    g_.cnv = f + 5 * C_;
    g_.ncv = f + 6 * C_;
    g_.fa = f + 7 * C_;
    g_.fb = f + 8 * C_;
    partials_.resize(bdot_partials_size(C_));
    pool_.resize((size_t)kVectors * total_);
    SAP_HIP_CHECK(hipHostMalloc((void**)&count_host_, sizeof(double), hipHostMallocMapped | hipHostMallocCoherent));
    SAP_HIP_CHECK(hipHostGetDevicePointer((void**)&count_dev_, count_host_, 0));
    double* p = pool_.get();
    auto take = [&]() { double* q = p; p += total_; return q; };
    r0_ = take(); xx_ = take(); Pv_ = take();
    for (int i = 0; i < 3; ++i) { rr_[i] = take(); uu_[i] = take(); }
    xmin_ = take(); Pxx_ = take(); APxx_ = take(); Pxmin_ = take(); APxmin_ = take();
    rc_ = take(); rcm_ = take(); r_ = take(); Ax_ = take(); d_ = take(); xo_ = take();
    SAP_HIP_CHECK(hipStreamSynchronize(st));
}

BatchKrylov::~BatchKrylov()
{
    drop_graph();
    if (count_host_) (void)hipHostFree(count_host_);
}

void BatchKrylov::step(int s)
{
    const int tb = 64;
    hipLaunchKernelGGL(batch_step_kernel, dim3((C_ + tb - 1) / tb), dim3(tb), 0, stream_, s, C_, cols_.get(), g_,
                       count_dev_);
    SAP_LAUNCH_CHECK();
}

void BatchKrylov::dot1(const double* a, const double* b, const int* gate)
{
    const double* pa[1] = {a};
    const double* pb[1] = {b};
    launch_bdots(bc_, 1, pa, pb, gate, partials_.get(), reinterpret_cast<double*>(cols_.get()) + BF(dot),
                 (int)(sizeof(BCol) / sizeof(double)), stream_);
}

void BatchKrylov::dot2(const double* a0, const double* b0, const double* a1, const double* b1, const int* gate)
{
    const double* pa[2] = {a0, a1};
    const double* pb[2] = {b0, b1};
    launch_bdots(bc_, 2, pa, pb, gate, partials_.get(), reinterpret_cast<double*>(cols_.get()) + BF(dot),
                 (int)(sizeof(BCol) / sizeof(double)), stream_);
}

void BatchKrylov::dot3(const double* a0, const double* a1, const double* a2, const int* gate)
{
    const double* pa[3] = {a0, a1, a2};
    launch_bdots(bc_, 3, pa, pa, gate, partials_.get(), reinterpret_cast<double*>(cols_.get()) + BF(dot),
                 (int)(sizeof(BCol) / sizeof(double)), stream_);
}

void BatchKrylov::drop_graph()
{
    if (exec_) { (void)hipGraphExecDestroy(exec_); exec_ = nullptr; }
    if (graph_) { (void)hipGraphDestroy(graph_); graph_ = nullptr; }
    graph_b_ = nullptr;
}

void BatchKrylov::axpby(BCoef ka, const double* x, BCoef kb, const double* y, double* z, const int* gate)
{
    launch_baxpby(bc_, ka, x, kb, y, z, reinterpret_cast<const double*>(cols_.get()),
                  (int)(sizeof(BCol) / sizeof(double)), gate, stream_);
}

// the true-residual check of the monitor, for the columns flagged chk:
// rr0 <- b - A M^{-1} xx, then the monitor's finished(|rr0|)
void BatchKrylov::check_true(BatchOps& ops, const double* b)
{
    ops.apply(xx_, Pxx_, g_.chk);
    ops.spmv(Pxx_, APxx_, g_.chk);
    axpby(bconst(1.0), b, bconst(-1.0), APxx_, rr_[0], g_.chk);
    dot1(rr_[0], rr_[0], g_.chk);
    step(BS_CHECK_DONE);
    step(BS_MIN);
    launch_bcopy(bc_, xx_, xmin_, g_.mcp, stream_);
}

// One outer iteration of BiCGStab(2) (sap/bicgstab2.h:91-314) for every
// column whose act flag is set; b is the inner right-hand side.
void BatchKrylov::iteration(BatchOps& ops, const double* b)
{
    const int* act = g_.act;
    step(BS_ITER_START);
    for (int j = 0; j < 2; ++j) {
        dot1(rr_[j], r0_, act);
        step(BS_BETA);
        for (int i = 0; i <= j; ++i) axpby(bconst(1.0), rr_[i], bfield(BF(beta), -1.0), uu_[i], uu_[i], act);
        ops.apply(uu_[j], Pv_, act);
        ops.spmv(Pv_, uu_[j + 1], act);
        dot1(uu_[j + 1], r0_, act);
        step(BS_ALPHA);
        for (int i = 0; i <= j; ++i) axpby(bfield(BF(alpha), -1.0), uu_[i + 1], bconst(1.0), rr_[i], rr_[i], act);
        dot3(rr_[0], uu_[0], xx_, act);
        step(BS_NORMS);
        // rr(j+1) = A M^{-1} rr(j) of the recursive rr(j), before the check
        ops.apply(rr_[j], Pv_, act);
        ops.spmv(Pv_, rr_[j + 1], act);
        axpby(bfield(BF(alpha)), uu_[0], bconst(1.0), xx_, xx_, act);
        step(BS_NEED_CHECK);
        check_true(ops, b);   // includes BS_MIN and the x_min copy
    }
    // MR part
    dot2(rr_[1], rr_[1], rr_[1], rr_[0], act);
    step(BS_SIGMA1);
    dot1(rr_[2], rr_[1], act);
    step(BS_TAO);
    axpby(bfield(BF(tao12), -1.0), rr_[1], bconst(1.0), rr_[2], rr_[2], act);
    dot2(rr_[2], rr_[2], rr_[2], rr_[0], act);
    step(BS_SIGMA2);
    step(BS_GAMMA);
    dot2(rr_[0], rr_[0], xx_, xx_, act);
    step(BS_STAG_G1);
    axpby(bfield(BF(g1)), rr_[0], bconst(1.0), xx_, xx_, act);
    axpby(bfield(BF(gp2), -1.0), rr_[2], bconst(1.0), rr_[0], rr_[0], act);
    axpby(bfield(BF(g2), -1.0), uu_[2], bconst(1.0), uu_[0], uu_[0], act);
    dot1(rr_[0], rr_[0], act);
    step(BS_MR_NORM);
    check_true(ops, b);
    step(BS_INC);
    axpby(bfield(BF(g1), -1.0), uu_[1], bconst(1.0), uu_[0], uu_[0], act);
    dot2(rr_[1], rr_[1], xx_, xx_, act);
    step(BS_STAG_GPP1);
    axpby(bfield(BF(gpp1)), rr_[1], bconst(1.0), xx_, xx_, act);
    axpby(bfield(BF(gp1), -1.0), rr_[1], bconst(1.0), rr_[0], rr_[0], act);
    dot1(rr_[0], rr_[0], act);
    step(BS_MR2_NORM);
    check_true(ops, b);
    step(BS_INC);
    step(BS_COUNT);
}

int BatchKrylov::run_inner(BatchOps& ops, const double* b, bool use_graph)
{
    // inner solve until no column is active; one synchronization per iteration
    int launches = 0;
    const int cap = 4 * (max_it_ + 2) + 8;   // every column stops once its quarter steps pass maxIt
    while (true) {
        if (use_graph) {
            if (!exec_ || graph_b_ != b) {
                drop_graph();
                SAP_HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
                try {
                    iteration(ops, b);
                } catch (...) {
                    hipGraph_t g = nullptr;
                    (void)hipStreamEndCapture(stream_, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                SAP_HIP_CHECK(hipStreamEndCapture(stream_, &graph_));
                SAP_HIP_CHECK(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
                graph_b_ = b;
            }
            SAP_HIP_CHECK(hipGraphLaunch(exec_, stream_));
        } else {
            iteration(ops, b);
        }
        SAP_HIP_CHECK(hipStreamSynchronize(stream_));
        ++launches;
        if (*count_host_ == 0.0) break;
        if (launches > cap) throw Error(SAP_ERR_HIP, "batched solve: a column did not stop");
    }
    return launches;
}

// the end of an inner solve for the columns that ran it: x_out = M^{-1} xx
// when converged, else the better of M^{-1} xx and M^{-1} x_min
// (sap/bicgstab2.h:316-353)
void BatchKrylov::finish_inner(BatchOps& ops, const double* b, double* x_out)
{
    step(BS_FINAL_SPLIT);
    ops.apply(xx_, x_out, g_.cnv);
    ops.apply(xx_, Pxx_, g_.ncv);
    ops.spmv(Pxx_, APxx_, g_.ncv);
    axpby(bconst(1.0), b, bconst(-1.0), APxx_, rc_, g_.ncv);
    ops.apply(xmin_, Pxmin_, g_.ncv);
    ops.spmv(Pxmin_, APxmin_, g_.ncv);
    axpby(bconst(1.0), b, bconst(-1.0), APxmin_, rcm_, g_.ncv);
    dot2(rc_, rc_, rcm_, rcm_, g_.ncv);
    step(BS_FINAL_PICK);
    launch_bcopy(bc_, Pxx_, x_out, g_.fa, stream_);
    launch_bcopy(bc_, Pxmin_, x_out, g_.fb, stream_);
}

void BatchKrylov::solve(BatchOps& ops, const double* B, double* X, const BatchSettings& s,
                        std::vector<BatchResult>& res)
{
    max_it_ = s.maxIt;
    drop_graph();   // the operators' buffers may have changed since the last solve
    // per-column settings
    std::vector<BCol> h(C_);
    for (int c = 0; c < C_; ++c) {
        BCol m{};
        m.relTol0 = s.relTol;
        m.absTol0 = s.absTol;
        m.refineTol = s.refineTol;
        m.refine = s.refine;
        m.maxIt = s.maxIt;
        m.maxm = 8;   // maxmSteps (sap/solver.h:314-319)
        h[c] = m;
    }
    SAP_HIP_CHECK(hipMemcpyAsync(cols_.get(), h.data(), sizeof(BCol) * C_, hipMemcpyHostToDevice, stream_));
    iterations_ = 0;
    if (s.refine <= 0) {
        // r0 = b - A x0, rr0 = r0, uu0 = 0, xx = x0, x_min = 0
        ops.spmv(X, r0_, nullptr);
        axpby(bconst(1.0), B, bconst(-1.0), r0_, r0_, nullptr);
        {
            double* cp[1] = {rr_[0]};
            double* zs[2] = {uu_[0], xmin_};
            launch_bspread(bc_, r0_, cp, 1, zs, 2, nullptr, stream_);
        }
        launch_bcopy(bc_, X, xx_, nullptr, stream_);
        dot2(B, B, r0_, r0_, nullptr);
        step(BS_INNER_INIT0);
        step(BS_INNER_START);
        iterations_ += run_inner(ops, B, s.graph);
        finish_inner(ops, B, X);
    } else {
        // outer FP64 refinement: x += d, A d = b - A x to refineTol
        launch_bcopy(bc_, X, xo_, nullptr, stream_);
        dot1(B, B, nullptr);
        step(BS_OUTER_INIT);
        for (int outer = 0; outer <= s.refine; ++outer) {
            ops.spmv(xo_, Ax_, g_.oact);
            axpby(bconst(1.0), B, bconst(-1.0), Ax_, r_, g_.oact);
            dot1(r_, r_, g_.oact);
            step(BS_OUTER_CHECK);
            step(BS_COUNT);
            SAP_HIP_CHECK(hipStreamSynchronize(stream_));
            if (*count_host_ == 0.0) break;
            // inner solve from d = 0: r0 = rr0 = r, uu0 = xx = x_min = 0
            {
                double* cp[2] = {r0_, rr_[0]};
                double* zs[3] = {uu_[0], xx_, xmin_};
                launch_bspread(bc_, r_, cp, 2, zs, 3, g_.run, stream_);
            }
            step(BS_INNER_START);
            iterations_ += run_inner(ops, r_, s.graph);
            finish_inner(ops, r_, d_);
            axpby(bconst(1.0), d_, bconst(1.0), xo_, xo_, g_.run);
            step(BS_OUTER_POST);
            step(BS_OUTER_ACT);
        }
        launch_bcopy(bc_, xo_, X, nullptr, stream_);
    }
    SAP_HIP_CHECK(hipMemcpyAsync(h.data(), cols_.get(), sizeof(BCol) * C_, hipMemcpyDeviceToHost, stream_));
    SAP_HIP_CHECK(hipStreamSynchronize(stream_));
    res.resize(C_);
    for (int c = 0; c < C_; ++c) {
        const BCol& m = h[c];
        BatchResult& r = res[c];
        if (s.refine <= 0) {
            r.code = m.code;
            r.iterations = m.iterations;
            r.rNorm = m.rNorm;
            r.rhsNorm = m.rhsNorm;
        } else {
            r.code = m.code_outer;
            r.iterations = m.its_total;
            r.rNorm = m.rn_outer;
            r.rhsNorm = m.bnorm;
        }
    }
}

}  // namespace sap_amd
