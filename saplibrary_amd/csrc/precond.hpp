// SaP preconditioner (the MI355X counterpart of sap::Precond,
// sap/precond.h:53-510) for the constant-bandwidth banded path.
#pragma once

#include "sap_internal.hpp"

namespace sap_amd {

struct PrecondConfig {
    int numPartitions = 1;    // requested P
    int factMethod = SAP_LU_only;
    int precondType = SAP_Spike;
    int safeFactorization = 0;
};

template <class T>
class Precond {
public:
    PrecondConfig cfg;
    hipStream_t stream = nullptr;
    Profiler* prof = nullptr;

    int64_t n = 0;
    int k = 0;
    int P = 1;                 // partitions actually used

    // timings of the last setup, milliseconds (sap::Stats fields)
    double time_toBanded = 0, time_offDiags = 0, time_bandLU = 0, time_bandUL = 0;
    double time_assembly = 0, time_fullLU = 0;
    int64_t actual_nnz = 0;

    // Factors the row-major FP64 band d_band (n x (2k+1)).  Returns 0 or a
    // negative sap_c.h code (SAP_ERR_ZERO_PIVOT).
    int setup(int64_t n_, int k_, const double* d_band);
    // z = M^{-1} v, FP64 in and out, preconditioner precision inside.
    void apply(const double* v, double* z);

    bool spike_active() const { return cfg.precondType == SAP_Spike && P > 1 && k > 0; }
    bool lu_ul() const { return cfg.factMethod == SAP_LU_UL; }

    // diagnostics
    const DevBuf<T>& factors() const { return B_; }
    const DevBuf<T>& factors_ul() const { return Bul_; }
    const DevBuf<T>& spikes() const { return WV_; }
    const DevBuf<T>& schur() const { return S_; }

private:
    DevBuf<T> B_, Bul_, offB_, offC_, WV_, S_, Sinv_, g_, corr_, scr0_, scr1_;
    DevBuf<FactorJob> jobs_;
    DevBuf<int> flag_;
    void factor(const std::vector<FactorJob>& jobs);
    bool zero_pivots(const T* band, int64_t row0, int64_t rows);
    double elapsed_ms(hipEvent_t a, hipEvent_t b);
};

}  // namespace sap_amd
