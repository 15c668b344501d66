// Batched BiCGStab(2) with device-resident scalars (batch.hip).
#pragma once

#include "sap_internal.hpp"

#include <vector>

namespace sap_amd {

// Per-column state of a batched solve, in device memory.  All doubles come
// first (the batched vector kernels address them as doubles).
struct BCol {
    // BiCGStab(2) (sap/bicgstab2.h:31-354)
    double rho0, rho1, alpha, omega, beta, gam;
    double tao12, sigma1, sigma2, gp1, gp2, g1, g2, gpp1;
    double dot[3];                        // the last batched dot products
    double r_norm, r_norm_act, r_norm_min;
    // BiCGStabLMonitor (sap/monitor.h:153-335) of the inner solve
    double rhsNorm, relTol, absTol, rNorm;
    // the solve's settings and the outer refinement
    double relTol0, absTol0, refineTol, bnorm, tol_outer, rn_outer;
    float iterations, its_total;
    int stag, moreSteps, code, maxIt, maxm;
    int outer, refine, code_outer;
};

// Per-column gate flags (device arrays of C ints each).
struct BGates {
    int* act;    // inner solve running
    int* chk;    // true-residual check due
    int* mcp;    // copy xx to x_min
    int* oact;   // outer refinement running
    int* run;    // ran an inner solve in this outer step
    int* cnv;    // ended converged
    int* ncv;    // ended not converged
    int* fa;     // not converged: M^{-1} xx is the better iterate
    int* fb;     // not converged: M^{-1} x_min is
};
constexpr int kNumGates = 9;

// scalar steps of batch_step_kernel
enum BatchStep {
    BS_OUTER_INIT, BS_OUTER_CHECK, BS_INNER_INIT0, BS_INNER_START, BS_ITER_START, BS_BETA, BS_ALPHA, BS_NORMS,
    BS_NEED_CHECK, BS_CHECK_DONE, BS_MIN, BS_TAO, BS_SIGMA1, BS_SIGMA2, BS_GAMMA, BS_STAG_G1, BS_STAG_GPP1,
    BS_MR_NORM, BS_MR2_NORM, BS_INC, BS_FINAL_SPLIT, BS_FINAL_PICK, BS_OUTER_POST, BS_OUTER_ACT, BS_COUNT
};

// The operators of a batched solve on batch vectors (column c at
// v + off[c]); gate may be null (all columns).  Both must queue work on the
// solve's stream only and allocate nothing (they are captured into a graph).
struct BatchOps {
    virtual ~BatchOps() = default;
    virtual void apply(const double* v, double* z, const int* gate) = 0;   // z = M^{-1} v
    virtual void spmv(const double* x, double* y, const int* gate) = 0;    // y = A x
};

struct BatchSettings {
    double relTol = 1e-6, absTol = 0, refineTol = 1e-6;
    int maxIt = 100;
    int refine = 0;
    bool graph = true;   // replay each iteration from a captured HIP graph
};

struct BatchResult {
    int code = 0;
    float iterations = 0.f;
    double rNorm = 0, rhsNorm = 0;
};

class BatchKrylov {
public:
    // columns: off[c], len[c] (doubles) within every batch vector
    BatchKrylov(hipStream_t st, const std::vector<int64_t>& off, const std::vector<int>& len);
    ~BatchKrylov();
    BatchKrylov(const BatchKrylov&) = delete;
    BatchKrylov& operator=(const BatchKrylov&) = delete;
    // Solves A X = B column by column (X holds the initial guesses; with
    // refine > 0 each column starts from its guess as the outer iterate).
    void solve(BatchOps& ops, const double* B, double* X, const BatchSettings& s, std::vector<BatchResult>& res);
    int columns() const { return C_; }
    hipStream_t stream() const { return stream_; }
    int64_t total() const { return total_; }
    int iteration_launches() const { return iterations_; }

private:
    hipStream_t stream_;
    int C_;
    int64_t total_ = 0;
    BatchCols bc_;
    DevBuf<int64_t> off_;
    DevBuf<int> len_;
    DevBuf<BCol> cols_;
    DevBuf<int> flags_;
    BGates g_{};
    DevBuf<double> partials_, pool_;
    double* count_host_ = nullptr;
    double* count_dev_ = nullptr;
    static constexpr int kVectors = 20;
    double *r0_, *xx_, *Pv_, *rr_[3], *uu_[3], *xmin_, *Pxx_, *APxx_, *Pxmin_, *APxmin_, *rc_, *rcm_;
    double *r_, *Ax_, *d_, *xo_;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    const double* graph_b_ = nullptr;
    int max_it_ = 100;
    int iterations_ = 0;

    void step(int s);
    void dot1(const double* a, const double* b, const int* gate);
    void dot2(const double* a0, const double* b0, const double* a1, const double* b1, const int* gate);
    void dot3(const double* a0, const double* a1, const double* a2, const int* gate);   // squared norms
    void axpby(BCoef ka, const double* x, BCoef kb, const double* y, double* z, const int* gate);
    void check_true(BatchOps& ops, const double* b);
    void iteration(BatchOps& ops, const double* b);
    int run_inner(BatchOps& ops, const double* b, bool use_graph);
    void finish_inner(BatchOps& ops, const double* b, double* x_out);
    void drop_graph();
};

}  // namespace sap_amd
