// SaP preconditioner (the MI355X counterpart of sap::Precond,
// sap/precond.h:53-510), constant-bandwidth band.
#pragma once

#include "autotune.hpp"
#include "graph.hpp"
#include "sap_internal.hpp"

namespace sap_amd {

struct PrecondConfig {
    int numPartitions = 1;    // requested P
    int factMethod = SAP_LU_only;
    int precondType = SAP_Spike;
    int safeFactorization = 0;
    int autotune = 0;         // time the tunable launch parameters on first use (autotune.hpp)
    int bcr = 0;              // useBCR: block cyclic reduction of the partitions (block-diagonal preconditioner)
};

// Reordering options of a CSR setup, as sap::Solver forwards them to
// sap::Precond (sap/solver.h:305-307).
struct ReorderConfig {
    bool reorder = true;           // performReorder
    bool doDB = true;              // performDB
    bool dbFirstStageOnly = false;
    bool scale = true;             // applyScaling
    int maxBandwidth = 0x7fffffff;
    double dropOffFraction = 0;
    bool track = false;            // trackReordering: keep what update() needs
    bool variable = false;         // variableBandwidth (LU_only, every pivot boosted)
    bool testDB = false;           // testDB: the DB statistics only, no preconditioner
};

// Setup timings: events recorded on the stream while a setup runs, turned
// into milliseconds once at its end (reading an event waits for it, which
// would stall the host between the setup's launches).
class SetupTimer {
public:
    SetupTimer() = default;
    ~SetupTimer()
    {
        for (hipEvent_t& e : ev_)
            if (e) (void)hipEventDestroy(e);
    }
    SetupTimer(const SetupTimer&) = delete;
    SetupTimer& operator=(const SetupTimer&) = delete;
    static constexpr int N = 8;
    // record event i (each index at most once per setup)
    void mark(int i, hipStream_t st)
    {
        if (!ev_[i]) SAP_HIP_CHECK(hipEventCreate(&ev_[i]));
        SAP_HIP_CHECK(hipEventRecord(ev_[i], st));
    }
    // *dst = time from event a to event b, filled in by resolve()
    void span(double* dst, int a, int b) { pend_.push_back({dst, a, b}); }
    void resolve()
    {
        for (const Span& p : pend_) {
            float ms = 0.f;
            SAP_HIP_CHECK(hipEventSynchronize(ev_[p.b]));
            SAP_HIP_CHECK(hipEventElapsedTime(&ms, ev_[p.a], ev_[p.b]));
            *p.dst = (double)ms;
        }
        pend_.clear();
    }
    void clear() { pend_.clear(); }

private:
    struct Span {
        double* dst;
        int a, b;
    };
    hipEvent_t ev_[N] = {};
    std::vector<Span> pend_;
};

// A second stream (default priority) for work that runs beside the main
// stream's, with the event that orders it after the main stream, and a
// pinned host word for a result read back from it.
struct SideStream {
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr;
    unsigned long long* host = nullptr;
    SideStream() = default;
    SideStream(const SideStream&) = delete;
    SideStream& operator=(const SideStream&) = delete;
    ~SideStream()
    {
        if (host) (void)hipHostFree(host);
        if (fork) (void)hipEventDestroy(fork);
        if (s) (void)hipStreamDestroy(s);
    }
    // the side stream, ordered after everything queued on main so far
    hipStream_t after(hipStream_t main)
    {
        if (!s) {
            SAP_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
            SAP_HIP_CHECK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
            SAP_HIP_CHECK(hipHostMalloc((void**)&host, sizeof(unsigned long long), hipHostMallocDefault));
        }
        SAP_HIP_CHECK(hipEventRecord(fork, main));
        SAP_HIP_CHECK(hipStreamWaitEvent(s, fork, 0));
        return s;
    }
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
    // host reordering of CSR setups; time_reorder accumulates over setups
    // as the reference's m_time_reorder does (sap/precond.h:1752)
    double time_DB = 0, time_reorder = 0, time_dropOff = 0, time_transfer = 0, time_cpu_assemble = 0;
    int k_db = 0, k_reorder = 0;    // stay 0 without reordering, as in the reference
    double dropoff_actual = 0;
    // testDB statistics (0 unless the last CSR setup ran the DB test, which
    // builds no preconditioner: tested_db_only)
    double diag_dom_ori = 0, d_p1_ori = 0, diag_dom = 0, d_p1 = 0;
    bool tested_db_only = false;
    BandProfile profile;            // ks_row / ks_col sums (empty without reordering)
    int64_t actual_nnz = 0;

    // Factors the row-major FP64 band d_band (n x (2k+1)).  Returns 0 or a
    // negative sap_c.h code (SAP_ERR_ZERO_PIVOT).
    int setup(int64_t n_, int k_, const double* d_band);
    // Factors the FP64 CSR matrix (device arrays): with rc.reorder the host
    // DB / RCM / drop-off pipeline of transformToBandedMatrix
    // (sap/precond.h:1656-2101), else the plain conversion of
    // convertToBandedMatrix(CSR) (sap/precond.h:2240-2363).  Throws
    // Error(SAP_ERR_SINGULAR / SAP_ERR_NEG_DB_WEIGHT) from DB.
    int setup_csr(int64_t n_, int64_t nnz, const int* d_rp, const int* d_ci, const double* d_val,
                  const ReorderConfig& rc);
    // z = M^{-1} v, FP64 in and out, preconditioner precision inside
    // (solve, sap/precond.h:1393-1406: left permutation and scaling, band
    // solve, right permutation and scaling when the setup reordered).
    void apply(const double* v, double* z);
    // Z(:, r) = M^{-1} V(:, r) for nr columns (column r at V + r*ldv, Z +
    // r*ldz), every kernel family in one launch for all columns; columns
    // whose gate flag (device array of nr ints, may be null) is 0 are
    // skipped.  Each column gets the arithmetic of apply(), so the same
    // bits.  Work buffers grow to nr columns on first use.
    void apply_multi(const double* V, int64_t ldv, double* Z, int64_t ldz, int nr, const int* gate = nullptr);
    // Allocates apply_multi's work buffers for nr columns (so that a later
    // apply_multi, e.g. under graph capture, allocates nothing).
    void reserve_multi(int nr);
    // Appends the full-sweep jobs of this preconditioner for nr columns
    // (input at voff + r*n, output at ooff + r*n, gates gate0 + r) to a
    // launch that covers several systems; false (nothing appended) unless
    // the apply is that one sweep pass: Block / SaP-D or one partition,
    // no reordering, k > 0 and a width the staged sweep takes.
    bool append_sweep_jobs(std::vector<SweepJob>& jobs, size_t& lds, int64_t voff, int64_t ooff, int nr, int gate0);
    bool reordered() const { return reorder_; }
    // Refactors with new values of the matrix given to the last setup, same
    // pattern (Precond::update, sap/precond.h:751-900): count doubles in
    // the layout of that setup (the row-major band, or the CSR values in
    // their order).  Returns false when count does not match; throws
    // Error(SAP_ERR_ILLEGAL_UPDATE) before a setup.
    bool update(const double* d_entries, int64_t count, int* rc);
    // A SaP-C setup of the constant band leaves the spike tips, the reduced
    // blocks and the nonzero count running on a second stream (the first
    // application's interface sweep needs only the factors and runs beside
    // them; its reduced solve waits for them).  finish_pending() waits for
    // that work and fills the statistics it produces (time_assembly,
    // time_fullLU, actual_nnz); every other user of those buffers calls it.
    void finish_pending();
    ~Precond();
    Precond() = default;
    Precond(const Precond&) = delete;
    Precond& operator=(const Precond&) = delete;

    bool spike_active() const
    {
        return cfg.precondType == SAP_Spike && P > 1 && k > 0 && !block_fallback_ && !cfg.bcr;
    }
    bool lu_ul() const { return cfg.factMethod == SAP_LU_UL && !var_mode_ && !block_fallback_ && !cfg.bcr; }
    // the last setup could not allocate the SaP-C buffers and built the
    // block-diagonal (SaP-D) preconditioner instead (sap/precond.h:1255-1265)
    bool block_fallback() const { return block_fallback_; }
    bool variable_mode() const { return var_mode_; }

    // diagnostics
    // the factored band: (2k+1) n entries, or on the variable-bandwidth path
    // the per-partition bands (band_offsets().back() entries)
    const DevBuf<T>& factors() const { return var_mode_ ? Bv_ : B_; }
    const std::vector<int64_t>& band_offsets() const { return vboff_; }
    const std::vector<int>& band_widths() const { return vks_; }
    const DevBuf<T>& factors_ul() const { return Bul_; }
    const DevBuf<T>& spikes() const { return WV_; }
    const DevBuf<T>& schur() const { return S_; }

private:
    DevBuf<T> B_, Bul_, offB_, offC_, WV_, S_, Sinv_, g_, corr_, scr0_, scr1_;
    DevBuf<FactorJob> jobs_;
    std::vector<FactorJob> jobs_host_;   // what jobs_ holds
    DevBuf<int> flag_;
    DevBuf<unsigned long long> cnt_;
    SideStream side_;
    SetupTimer tm_;
    // the deferred spike / reduced-block work (see finish_pending)
    SideStream side2_;
    SetupTimer tm2_;
    hipEvent_t spk_done_ = nullptr;   // recorded after the reduced blocks
    bool spk_wait_ = false;           // the main stream is not yet ordered after spk_done_
    bool spk_open_ = false;           // its statistics are not yet read
    void wait_spikes();
    // the conversion that preceded factor_band() also wrote Bul_ (the copy
    // the UL factorizations overwrite), so no device copy is needed
    bool bul_filled_ = false;
    // LU_UL with the register-resident factorization: the matrix is written
    // once, to Bin_, which the LU (into B_) and UL (into Bul_)
    // factorizations both read; otherwise both factor in place
    DevBuf<T> Bin_;
    bool sep_in_ = false;
    T* band_in() const { return var_mode_ ? Bv_.get() : sep_in_ ? Bin_.get() : B_.get(); }
    size_t band_in_bytes() const { return (var_mode_ ? Bv_.size() : sep_in_ ? Bin_.size() : B_.size()) * sizeof(T); }
    // reordering: row permutation (DB then RCM), column reordering, scalings
    bool reorder_ = false, scale_ = false;
    // variable-bandwidth setup: the preconditioner of the variable-bandwidth
    // path is factored LU_only with every pivot boosted
    bool var_mode_ = false;
    bool block_fallback_ = false;
    // allocates the SaP-C buffers with alloc(); on an out-of-memory error
    // releases them with release() and switches this setup to Block
    template <class A, class R>
    bool try_spike_buffers(A alloc, R release);
    // variable-bandwidth storage: the partitions' own bands in second-level
    // order (partition p at vboff_[p], half-bandwidth vks_[p]), then the
    // coupling blocks B_i and C_{i+1} (k x k column-major each, first-level
    // order); second-level permutation sp_ (first-level row -> row) and its
    // inverse sr_; the packed spike columns (offDiagPerms) and their counts
    DevBuf<T> Bv_;
    std::vector<int> vks_;
    std::vector<int64_t> vboff_;
    DevBuf<int> vks_d_, sp_, sr_, vperm_r_, vperm_l_, vwidth_;
    DevBuf<int64_t> vboff_d_;
    std::vector<VarSpikeCol> vcols_h_;   // the packed coupling columns, one spike solve each
    DevBuf<VarSpikeCol> vcols_;
    void var_spikes();
    std::vector<T> vcoupling_h_;   // host copy of the coupling blocks (uploaded by setup)
    T* voffB() const { return Bv_.get() + vboff_.back(); }
    T* voffC() const { return voffB() + (size_t)k * k * (P - 1); }
    DevBuf<double> gd_, vq_, gdm_, vqm_;   // first-pass result, purified right-hand side (second-level order)
    // the full-sweep job tables of the partition bands: [0] for one column,
    // [1] for apply_multi's columns (built by reserve_multi, so that a
    // captured apply_multi uploads nothing)
    struct VJobKey { int nr; int64_t ldv, ldz; const void* scr; };
    struct VJobs {
        DevBuf<SweepJob> d;
        size_t n = 0;                       // jobs in d (0: not built)
        size_t start[3] = {0, 0, 0}, count[3] = {0, 0, 0}, lds[3] = {0, 0, 0};   // per sweep variant
        VJobKey key{0, 0, 0, nullptr};
    } vjobs_[2];
    VJobs& var_jobs(int nr, int64_t ldv, int64_t ldz, T* scr);
    int var_factor_steps();
    void var_sweeps(const double* V, int64_t ldv, double* Z, int64_t ldz, int nr, const int* gate, T* scr);
    void var_solve(const double* V, int64_t ldv, double* Z, int64_t ldz, int nr, const int* gate);
    // block cyclic reduction (cfg.bcr): the blocks of every partition, the
    // eliminations of each level, each partition's last block, the
    // couplings (CL, CU, k x k per block, updated level by level) and the
    // neighbours' couplings each eliminated block saw (UA, LC)
    bool bcr_ = false;
    std::vector<BcrBlock> bblk_;
    DevBuf<BcrBlock> bblk_d_;
    struct BcrLevel {
        std::vector<BcrElim> el;
        DevBuf<BcrElim> d;
    };
    std::vector<BcrLevel> blev_;
    std::vector<int> bfinal_;
    DevBuf<T> bCL_, bCU_, bUA_, bLC_;
    DevBuf<double> bw_, bt_, bwm_, btm_;
    struct BcrJobs {
        DevBuf<SweepJob> d;
        std::vector<size_t> start, count;   // per stage: levels 0..L-1, then the last blocks
        size_t lds = 0;
        int mode = 0, nr = 0;
        const void* scr = nullptr;
    } bjobs_[2];
    int bcr_factor_steps();
    BcrJobs& bcr_jobs(int nr, T* scr);
    void bcr_solve(const double* V, int64_t ldv, double* Z, int64_t ldz, int nr, const int* gate);
    DevBuf<int> row_perm_, col_reord_;
    DevBuf<T> row_scale_, col_scale_;
    DevBuf<double> vp_, zp_;
    // update(): what the last setup was, and its value maps
    enum class Source { none, band, csr, csr_reordered } src_ = Source::none;
    int64_t src_count_ = 0;                 // doubles an update() takes
    const int* src_rp_ = nullptr;           // CSR pattern of a plain CSR setup (caller's)
    const int* src_ci_ = nullptr;
    DevBuf<int64_t> upd_map_;               // input entry -> band slot (-1: dropped)
    DevBuf<T> upd_scale_;                   // DB scale factor per input entry
    void begin_setup(int64_t n_, int k_, bool const_band = true);
    int factor_band();
    int factor_band_steps();
    void solve_band(const double* v, double* z);
    // tuned launch parameters (0: the untuned defaults)
    int tile_ = 0, nbuf_ = 0;
    // LU_UL with every pivot boosted: LU and UL factors of every partition,
    // two workgroups per partition in the full sweep
    bool split_ = false;
    void tune_sweeps();
    void solve_band_multi(const double* V, int64_t ldv, double* Z, int64_t ldz, int nr, const int* gate);
    // apply_multi's work buffers (nr_cap_ columns)
    DevBuf<T> gm_, corrm_, scr0m_, scr1m_;
    DevBuf<double> vpm_, zpm_;
    int nr_cap_ = 0;
    void factor(const std::vector<FactorJob>& jobs);
    // true when a pivot of rows [row0, row0 + rows) of band, or of rows
    // [row1, row1 + rows1) of band1 (if given), is below 1e-7 in magnitude
    bool zero_pivots(const T* band, int64_t row0, int64_t rows, const T* band1 = nullptr, int64_t row1 = 0,
                     int64_t rows1 = 0);
    double elapsed_ms(hipEvent_t a, hipEvent_t b);
};

}  // namespace sap_amd
