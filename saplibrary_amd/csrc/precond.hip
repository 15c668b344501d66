// Setup and apply of the SaP preconditioner for the constant-bandwidth
// banded path (sap/precond.h:910-1362 setup, :1370-1528 apply).
#include "precond.hpp"

#include <algorithm>
#include <chrono>
#include <cstdlib>

namespace sap_amd {

template <class T>
double Precond<T>::elapsed_ms(hipEvent_t a, hipEvent_t b)
{
    float ms = 0.f;
    SAP_HIP_CHECK(hipEventSynchronize(b));
    SAP_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    return (double)ms;
}

template <class T>
Precond<T>::~Precond()
{
    try {
        finish_pending();
    } catch (...) {
    }
    if (spk_done_) (void)hipEventDestroy(spk_done_);
}

template <class T>
void Precond<T>::finish_pending()
{
    if (!spk_open_) return;
    spk_open_ = spk_wait_ = false;
    SAP_HIP_CHECK(hipStreamSynchronize(side2_.s));
    tm2_.resolve();
    actual_nnz = (2 * (int64_t)k + 1) * n - (int64_t)*side2_.host;
}

template <class T>
void Precond<T>::wait_spikes()
{
    if (!spk_wait_) return;
    SAP_HIP_CHECK(hipStreamWaitEvent(stream, spk_done_, 0));
    spk_wait_ = false;
}

namespace {

// profile scopes of work queued on another stream record their events there
struct ProfOnStream {
    Profiler* p;
    hipStream_t saved = nullptr;
    ProfOnStream(Profiler* p_, hipStream_t s) : p(p_)
    {
        if (p) { saved = p->stream; p->stream = s; }
    }
    ~ProfOnStream()
    {
        if (p) p->stream = saved;
    }
};

}  // namespace

template <class T>
void Precond<T>::factor(const std::vector<FactorJob>& jobs)
{
    if (jobs.empty()) return;
    // the job list is uploaded only when it changed (repeated setups of one
    // size reuse it)
    const bool same = jobs.size() == jobs_host_.size() &&
                      std::equal(jobs.begin(), jobs.end(), jobs_host_.begin(), [](const FactorJob& a, const FactorJob& b) {
                          return a.start == b.start && a.len == b.len && a.dir == b.dir && a.boost == b.boost &&
                                 a.boost_last == b.boost_last && a.buf == b.buf && a.kb == b.kb && a.boff == b.boff;
                      });
    if (!same) {
        jobs_.resize(jobs.size());
        SAP_HIP_CHECK(hipMemcpyAsync(jobs_.get(), jobs.data(), jobs.size() * sizeof(FactorJob), hipMemcpyHostToDevice,
                                     stream));
        jobs_host_ = jobs;
    }
    ProfScope ps(prof, "band_factor");
    int64_t max_len = 0;
    for (const FactorJob& j : jobs) max_len = std::max(max_len, (int64_t)j.len);
    launch_band_factor<T>(B_.get(), Bul_.get(), sep_in_ ? Bin_.get() : nullptr, k, jobs_.get(), (int)jobs.size(),
                          max_len, stream);
}

template <class T>
bool Precond<T>::zero_pivots(const T* band, int64_t row0, int64_t rows, const T* band1, int64_t row1, int64_t rows1)
{
    flag_.resize(1);
    SAP_HIP_CHECK(hipMemsetAsync(flag_.get(), 0, sizeof(int), stream));
    launch_zero_pivot_check<T>(band, k, row0, rows, flag_.get(), stream);
    if (band1) launch_zero_pivot_check<T>(band1, k, row1, rows1, flag_.get(), stream);
    int h = 0;
    SAP_HIP_CHECK(hipMemcpyAsync(&h, flag_.get(), sizeof(int), hipMemcpyDeviceToHost, stream));
    SAP_HIP_CHECK(hipStreamSynchronize(stream));
    return h != 0;
}

namespace {

double wall_now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Events {
    hipEvent_t e[6];
    Events()
    {
        for (auto& x : e) SAP_HIP_CHECK(hipEventCreate(&x));
    }
    ~Events()
    {
        for (auto& x : e) (void)hipEventDestroy(x);
    }
};

template <class V>
void upload(DevBuf<V>& d, const std::vector<V>& h, hipStream_t st)
{
    d.resize(h.size());
    if (!h.empty()) SAP_HIP_CHECK(hipMemcpyAsync(d.get(), h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice, st));
}

}  // namespace

namespace {

// Out-of-memory for the SaP-C buffers can be forced from the host
// (SAP_AMD_FAIL_SPIKE_ALLOC=1 in the environment), so the fallback below
// is testable without exhausting the device.
bool spike_alloc_forced_to_fail()
{
    const char* e = std::getenv("SAP_AMD_FAIL_SPIKE_ALLOC");
    return e && e[0] == '1';
}

int clamp_partitions(int64_t n, int k, int requested)
{
    const int64_t rows = std::max<int64_t>(k + 1, 2 * (int64_t)k);
    const int64_t maxP = std::max<int64_t>(1, n / rows);
    return std::max(1, (int)std::min<int64_t>(requested, maxP));
}

}  // namespace

// The reference builds the truncated-SPIKE buffers inside try blocks and,
// on std::bad_alloc, factors the block-diagonal preconditioner instead
// (sap/precond.h:1242-1265, :1281-1291, :1325-1346); the same here for the
// device allocations of the SaP-C path.
template <class T>
template <class A, class R>
bool Precond<T>::try_spike_buffers(A alloc, R release)
{
    if (block_fallback_) return false;
    try {
        if (spike_alloc_forced_to_fail()) throw Error(SAP_ERR_OOM, "SaP-C buffers (forced)");
        alloc();
        return true;
    } catch (const Error& e) {
        if (e.code != SAP_ERR_OOM) throw;
        release();
        block_fallback_ = true;
        return false;
    }
}

template <class T>
void Precond<T>::begin_setup(int64_t n_, int k_, bool const_band)
{
    finish_pending();
    n = n_;
    k = k_;
    block_fallback_ = false;
    nr_cap_ = 0;   // apply_multi's buffers are sized again for this setup
    nbuf_ = 0;     // the sweeps' staging depth is tuned again for this k
    time_toBanded = time_offDiags = time_bandLU = time_bandUL = time_assembly = time_fullLU = 0;
    // Partition clamp of convertToBandedMatrix (sap/precond.h:2116-2117,
    // :2265-2266): each partition needs k+1 <= n_p and 2k <= n_p.  Every
    // path uses it: with shorter partitions the 2k-row interface windows of
    // neighbouring boundaries overlap, and the reduced solve and
    // purification of one boundary would change the other's values.
    P = clamp_partitions(n, k, cfg.numPartitions);
    const int64_t w = 2 * k + 1;
    flag_.resize(1);
    // 8 elements of slack: the sweeps copy 16-byte chunks covering each
    // column segment, which may reach a few elements past the last column
    auto alloc_band = [&](DevBuf<T>& b) { b.resize((size_t)w * n + 8); };
    if (const_band) {
        alloc_band(B_);
        Bv_.release();
    } else {
        B_.release();   // the variable-bandwidth bands replace it
    }
    // LU_UL factors the UL of partitions 1..P-1 on a second copy of the band
    // (kept across setups of the same size)
    bool need_ul = const_band && cfg.factMethod == SAP_LU_UL && cfg.precondType == SAP_Spike && P > 1 && k > 0 &&
                   !cfg.bcr;
    // the first partition is the longest (PartInfo gives the n mod P extra
    // rows to the first partitions)
    int64_t s0 = 0, len0 = n;
    PartInfo{n, P}.bounds(0, s0, len0);
    sep_in_ = need_ul && band_factor_resident_fits<T>(k, len0);
    // the SaP-C buffers of the constant-band path, before anything is
    // written to the bands (the variable-bandwidth path allocates its own)
    if (const_band && cfg.precondType == SAP_Spike && P > 1 && k > 0 && !cfg.bcr) {
        const size_t kk = (size_t)k * k;
        const bool ok = try_spike_buffers(
            [&] {
                if (need_ul) {
                    alloc_band(Bul_);
                    if (sep_in_) alloc_band(Bin_);
                    scr1_.resize(n);
                }
                offB_.resize(kk * (P - 1));
                offC_.resize(kk * (P - 1));
                WV_.resize(2 * kk * (P - 1));
                S_.resize(kk * (P - 1));
                Sinv_.resize(kk * (P - 1));
                corr_.resize((size_t)2 * k * (P - 1));
            },
            [&] {
                Bul_.release(); Bin_.release(); scr1_.release();
                offB_.release(); offC_.release(); WV_.release(); S_.release(); Sinv_.release(); corr_.release();
            });
        if (!ok) need_ul = sep_in_ = false;
    }
    if (!need_ul) Bul_.release();
    if (!sep_in_) Bin_.release();
}

template <class T>
int Precond<T>::setup(int64_t n_, int k_, const double* d_band)
{
    reorder_ = scale_ = var_mode_ = false;
    row_perm_.release(); col_reord_.release(); row_scale_.release(); col_scale_.release();
    vp_.release(); zp_.release();
    k_db = k_reorder = 0;
    dropoff_actual = 0;
    profile = BandProfile{};
    begin_setup(n_, k_);
    src_ = Source::band;
    src_count_ = n_ * (2 * (int64_t)k_ + 1);
    upd_map_.release();
    upd_scale_.release();
    tm_.clear();
    tm_.mark(0, stream);
    tile_ = 0;
    if (cfg.autotune) {
        // conversion tile width (a copy: every width writes the same band)
        const int tmax = band_to_colmajor_max_tile<T>(k);
        if (tmax >= 8) {
            // multiples of 8 from 16 up: narrower tiles let two or three
            // workgroups share a CU (read and write phases overlap), wider
            // ones read longer row segments (less overfetch at their ends)
            std::vector<int> cand;
            for (int t = 16; t <= tmax; t += 8) cand.push_back(t);
            if (cand.empty() || cand.back() != tmax) cand.push_back(tmax);
            T* b2 = sep_in_ ? nullptr : Bul_.get();
            tile_ = Autotuner::instance().tune(
                TUNE_B2C_TILE, k, (int)sizeof(T), tmax, cand,
                [&](int t) { launch_band_to_colmajor<T>(n, k, d_band, band_in(), stream, b2, t); }, stream);
            tm_.clear();
            tm_.mark(0, stream);
        }
    }
    {
        ProfScope ps(prof, "band_to_colmajor");
        // the UL factorizations' copy of the band, when they factor in
        // place, is written in the same pass
        launch_band_to_colmajor<T>(n, k, d_band, band_in(), stream, sep_in_ ? nullptr : Bul_.get(), tile_);
        bul_filled_ = !sep_in_ && Bul_.get() != nullptr;
    }
    tm_.mark(1, stream);
    tm_.span(&time_toBanded, 0, 1);
    const int rc = factor_band();
    if (rc == 0) tune_sweeps();
    return rc;
}

// Staging depth of the triangular sweeps, timed on full applications of
// this preconditioner (the depth does not change the arithmetic).
template <class T>
void Precond<T>::tune_sweeps()
{
    nbuf_ = 0;
    if (!cfg.autotune || k == 0 || cfg.precondType == SAP_None) return;
    const int nmax = sweep_max_nbuf<T>(k);
    if (nmax < 4) return;
    std::vector<int> cand;
    for (int b = 4; b <= nmax; ++b) cand.push_back(b);
    if (cand.size() < 2) return;
    // already tuned for this width: no timing runs (and no allocation, whose
    // release would wait for the setup work still on the second stream)
    int known = 0;
    if (Autotuner::instance().find(TUNE_SWEEP_NBUF, k, (int)sizeof(T), &known)) {
        nbuf_ = known;
        return;
    }
    DevBuf<double> tmp(2 * (size_t)n);
    launch_fill(n, 1.0, tmp.get(), stream);
    const bool on = prof && prof->enabled;
    if (on) prof->enabled = false;
    try {
        nbuf_ = Autotuner::instance().tune(
            TUNE_SWEEP_NBUF, k, (int)sizeof(T), nmax, cand,
            [&](int b) { nbuf_ = b; solve_band(tmp.get(), tmp.get() + n); }, stream);
    } catch (...) {
        if (on) prof->enabled = true;
        throw;
    }
    if (on) prof->enabled = true;
    SAP_HIP_CHECK(hipStreamSynchronize(stream));
}

template <class T>
int Precond<T>::setup_csr(int64_t n_, int64_t nnz, const int* d_rp, const int* d_ci, const double* d_val,
                          const ReorderConfig& rc)
{
    Events ev;
    var_mode_ = false;
    tm_.clear();
    if (!rc.reorder) {
        reorder_ = scale_ = false;
        row_perm_.release(); col_reord_.release(); row_scale_.release(); col_scale_.release();
        vp_.release(); zp_.release();
        k_db = k_reorder = 0;
        dropoff_actual = 0;
        profile = BandProfile{};
        // half-bandwidth over every stored entry, explicit zeros included
        flag_.resize(1);
        SAP_HIP_CHECK(hipMemsetAsync(flag_.get(), 0, sizeof(int), stream));
        launch_csr_bandwidth(n_, d_rp, d_ci, flag_.get(), stream);
        int kk = 0;
        SAP_HIP_CHECK(hipMemcpyAsync(&kk, flag_.get(), sizeof(int), hipMemcpyDeviceToHost, stream));
        SAP_HIP_CHECK(hipStreamSynchronize(stream));
        begin_setup(n_, kk);
        src_ = Source::csr;
        src_count_ = nnz;
        src_rp_ = d_rp;
        src_ci_ = d_ci;
        upd_map_.release();
        upd_scale_.release();
        tm_.mark(0, stream);
        {
            ProfScope ps(prof, "csr_to_band");
            SAP_HIP_CHECK(hipMemsetAsync(band_in(), 0, band_in_bytes(), stream));
            launch_csr_to_band<double, T>(n, k, d_rp, d_ci, d_val, band_in(), stream);
        }
        tm_.mark(1, stream);
        tm_.span(&time_toBanded, 0, 1);
        const int rc0 = factor_band();
        if (rc0 == 0) tune_sweeps();
        return rc0;
    }

    // (1) matrix to the host (transformToBandedMatrix, sap/precond.h:1656-1713)
    double t = wall_now_ms();
    std::vector<int> h_rp(n_ + 1), h_ci(nnz);
    std::vector<double> h_val(nnz);
    SAP_HIP_CHECK(hipMemcpyAsync(h_rp.data(), d_rp, (n_ + 1) * sizeof(int), hipMemcpyDeviceToHost, stream));
    if (nnz > 0) {
        SAP_HIP_CHECK(hipMemcpyAsync(h_ci.data(), d_ci, nnz * sizeof(int), hipMemcpyDeviceToHost, stream));
        SAP_HIP_CHECK(hipMemcpyAsync(h_val.data(), d_val, nnz * sizeof(double), hipMemcpyDeviceToHost, stream));
    }
    SAP_HIP_CHECK(hipStreamSynchronize(stream));
    if (h_rp[0] != 0 || h_rp[n_] != nnz) throw Error(SAP_ERR_BADARG, "CSR row offsets do not span nnz entries");
    time_transfer = wall_now_ms() - t;

    // (2) DB, RCM (when maxBandwidth > 64, sap/precond.h:1737-1738), drop-off
    GraphOptions go;
    go.doDB = rc.doDB;
    go.scale = rc.scale;
    go.dbFirstStageOnly = rc.dbFirstStageOnly;
    go.doRCM = rc.maxBandwidth > 64;
    go.maxBandwidth = rc.maxBandwidth;
    go.dropOffFrac = rc.dropOffFraction;
    go.testDB = rc.testDB;
    Reordering<T> R = reorder_csr<T>((int)n_, h_rp.data(), h_ci.data(), h_val.data(), go);
    time_DB = R.time_DB;
    time_reorder += R.time_reorder;
    time_dropOff = R.time_dropOff;
    k_db = R.k_db;
    k_reorder = R.k_reorder;
    dropoff_actual = R.dropoff_actual;
    diag_dom_ori = R.diag_dom_ori;
    d_p1_ori = R.d_p1_ori;
    diag_dom = R.diag_dom;
    d_p1 = R.d_p1;
    tested_db_only = rc.testDB;
    if (rc.testDB) {
        // the DB test ends the setup here (sap/precond.h:1755-1756)
        n = n_;
        k = R.k;
        src_ = Source::none;
        return 0;
    }

    // (3) variable bandwidth (sap/precond.h:1777-1800): with more than one
    //     partition (clamped to n / (k+1)) and k > 0 the reference builds the
    //     variable-bandwidth preconditioner.  It is the same truncated-SPIKE
    //     operator (diagonal blocks, their k x k spike tips, the reduced
    //     blocks, purification with the coupling blocks); this build factors
    //     the diagonal blocks on the constant-width band of the reordered
    //     matrix instead of after the per-partition second-level RCM, LU_only
    //     with every pivot boosted as the reference's variable-bandwidth
    //     kernels do.  The host assembly runs for its band extents
    //     (Stats::nuKf, flops_LU).
    //     The reference clamps the partition count to n / (k+1) on this
    //     path (sap/precond.h:1777-1778); partitions shorter than 2k rows
    //     make neighbouring interface windows overlap (the reference's
    //     in-place reduced sweeps, sap/device/sweep_band_var.cuh:21-191, then
    //     race), so the banded path's clamp n / max(k+1, 2k) is used here
    //     too, and the assembly below uses the same partition count.
    // (BCR runs on the constant-width band)
    var_mode_ = rc.variable && !cfg.bcr && R.k > 0 && clamp_partitions(n_, R.k, cfg.numPartitions) > 1;
    begin_setup(n_, R.k, !var_mode_);
    const PartInfo pi{n, P};

    // (4) band of the reordered matrix, vectors of the permutations
    t = wall_now_ms();
    VarBand<T> VB;
    if (var_mode_) {
        VB = assemble_var((int)n, R.k, P, R.rp, R.ci, R.val);
        profile = BandProfile{};
        for (int i = 0; i < (int)n; ++i) {
            profile.sum_row += VB.ks_row[i];
            profile.sum_col += VB.ks_col[i];
            profile.sum_prod += (double)VB.ks_row[i] * VB.ks_col[i];
        }
        // storage widths: a partition without off-diagonal entries keeps a
        // band of width 3 (zeros beside the diagonal) so every kernel sees
        // k >= 1.  Each band starts 256-byte aligned or better (the staged
        // sweeps copy 16-byte chunks aligned relative to the band's start).
        vks_.resize(P);
        vboff_.assign(P + 1, 0);
        for (int p = 0; p < P; ++p) {
            int64_t s, len;
            pi.bounds(p, s, len);
            vks_[p] = std::max(VB.ks[p], 1);
            vboff_[p + 1] = (vboff_[p] + len * (2 * (int64_t)vks_[p] + 1) + 63) / 64 * 64;
        }
    } else {
        profile = band_profile((int)n, R.rp.data(), R.ci.data(), P);
        vks_.clear();
        vboff_.clear();
    }
    time_cpu_assemble = wall_now_ms() - t;
    t = wall_now_ms();
    DevBuf<int> rp_d, ci_d;
    DevBuf<T> val_d;
    if (var_mode_) {
        // the second-level permutation folded into the vector permutations
        // (sap/precond.h:2070-2080): leftTrans lands in second-level order,
        // rightTrans reads from it
        std::vector<int> rperm(n), creord(n);
        for (int64_t i = 0; i < n; ++i) rperm[i] = VB.second_perm[R.row_perm[i]];
        for (int64_t i = 0; i < n; ++i) creord[i] = R.col_reord[VB.second_reorder[i]];
        upload(row_perm_, rperm, stream);
        upload(col_reord_, creord, stream);
        upload(rp_d, VB.drp, stream);
        upload(ci_d, VB.dci, stream);
        upload(val_d, VB.dval, stream);
        upload(vks_d_, vks_, stream);
        upload(vboff_d_, vboff_, stream);
        upload(sp_, VB.second_perm, stream);
        upload(sr_, VB.second_reorder, stream);
        upload(vperm_r_, VB.perms_right, stream);
        upload(vperm_l_, VB.perms_left, stream);
        // packed column counts per (boundary, side), and the spike columns
        std::vector<int> widths(2 * (P - 1));
        vcols_h_.clear();
        for (int i = 0; i < P - 1; ++i) {
            widths[2 * i] = VB.widths_right[i];
            widths[2 * i + 1] = VB.widths_left[i];
            for (int q = 0; q < VB.widths_right[i]; ++q) vcols_h_.push_back({i, 0, q});
            for (int q = 0; q < VB.widths_left[i]; ++q) vcols_h_.push_back({i, 1, q});
        }
        upload(vwidth_, widths, stream);
        upload(vcols_, vcols_h_, stream);
        // the coupling blocks, column-major k x k after the bands
        const size_t kk = (size_t)k * k;
        Bv_.resize((size_t)vboff_[P] + 2 * kk * (P - 1) + 8);
        std::vector<T> cb(2 * kk * (P - 1));
        for (int i = 0; i < P - 1; ++i)
            for (int a = 0; a < k; ++a)
                for (int b = 0; b < k; ++b) {
                    cb[i * kk + (size_t)b * k + a] = VB.offDiags[(size_t)2 * i * kk + (size_t)a * k + b];
                    cb[(P - 1) * kk + i * kk + (size_t)b * k + a] =
                        VB.offDiags[(size_t)(2 * i + 1) * kk + (size_t)a * k + b];
                }
        vcoupling_h_.swap(cb);
    } else {
        upload(rp_d, R.rp, stream);
        upload(ci_d, R.ci, stream);
        upload(val_d, R.val, stream);
        upload(row_perm_, R.row_perm, stream);
        upload(col_reord_, R.col_reord, stream);
        vcoupling_h_.clear();
    }
    upload(row_scale_, R.row_scale, stream);
    upload(col_scale_, R.col_scale, stream);
    src_ = Source::csr_reordered;
    src_count_ = nnz;
    std::vector<int64_t> map;
    if (rc.track) {
        // input entry -> band slot of where reordering put it (-1 when it
        // was dropped or was an explicit zero)
        const int64_t w = 2 * (int64_t)k + 1;
        map.assign(nnz, -1);
        const size_t kk = (size_t)k * k;
        for (int r = 0; r < (int)n; ++r)
            for (int e = R.rp[r]; e < R.rp[r + 1]; ++e) {
                const int64_t c = R.ci[e];
                if (!var_mode_) {
                    map[R.entry_orig[e]] = c * w + k + r - c;
                    continue;
                }
                // variable bandwidth: the partition's own band (second-level
                // order) or a coupling block (first-level order)
                const int pr = pi.part_of(r), pc = pi.part_of(c);
                int64_t slot;
                if (pr == pc) {
                    int64_t s, len;
                    pi.bounds(pr, s, len);
                    const int64_t r2 = VB.second_perm[r], c2 = VB.second_perm[c];
                    slot = vboff_[pr] + (c2 - s) * (2 * (int64_t)vks_[pr] + 1) + vks_[pr] + (r2 - c2);
                } else {
                    const int i = std::min(pr, pc);
                    int64_t s1, l1;
                    pi.bounds(i + 1, s1, l1);
                    if (pc > pr)   // B_i(a, b), a = r - (s1 - k), b = c - s1
                        slot = vboff_[P] + (int64_t)(i * kk) + (c - s1) * k + (r - (s1 - k));
                    else           // C_{i+1}(a, b), a = r - s1, b = c - (s1 - k)
                        slot = vboff_[P] + (int64_t)((P - 1 + i) * kk) + (c - (s1 - k)) * k + (r - s1);
                }
                map[R.entry_orig[e]] = slot;
            }
        upload(upd_map_, map, stream);
        if (rc.scale && !R.entry_scale.empty()) upload(upd_scale_, R.entry_scale, stream);
        else upd_scale_.release();
    } else {
        upd_map_.release();
        upd_scale_.release();
    }
    vp_.resize(n);
    zp_.resize(n);
    SAP_HIP_CHECK(hipStreamSynchronize(stream));
    time_transfer += wall_now_ms() - t;
    reorder_ = true;
    scale_ = rc.scale;
    SAP_HIP_CHECK(hipEventRecord(ev.e[0], stream));
    {
        ProfScope ps(prof, "csr_to_band");
        SAP_HIP_CHECK(hipMemsetAsync(band_in(), 0, band_in_bytes(), stream));
        if (var_mode_) {
            launch_csr_to_varband<T>(pi, rp_d.get(), ci_d.get(), val_d.get(), vks_d_.get(), vboff_d_.get(), Bv_.get(),
                                     stream);
            if (!vcoupling_h_.empty())
                SAP_HIP_CHECK(hipMemcpyAsync(voffB(), vcoupling_h_.data(), vcoupling_h_.size() * sizeof(T),
                                             hipMemcpyHostToDevice, stream));
        } else {
            launch_csr_to_band<T, T>(n, k, rp_d.get(), ci_d.get(), val_d.get(), band_in(), stream);
        }
    }
    SAP_HIP_CHECK(hipEventRecord(ev.e[1], stream));
    time_toBanded = elapsed_ms(ev.e[0], ev.e[1]);   // waits for the fill before the uploads are freed
    vcoupling_h_ = std::vector<T>();
    const int rc1 = factor_band();
    if (rc1 == 0) tune_sweeps();
    return rc1;
}

template <class T>
bool Precond<T>::update(const double* d_entries, int64_t count, int* rc)
{
    *rc = 0;
    if (src_ == Source::none) throw Error(SAP_ERR_ILLEGAL_UPDATE, "Illegal call to update() before setup().");
    if (src_ == Source::csr_reordered && !upd_map_.get())
        throw Error(SAP_ERR_ILLEGAL_UPDATE, "Illegal call to update() with reordering tracking disabled.");
    if (count != src_count_) return false;
    if (cfg.precondType == SAP_None) return true;
    finish_pending();
    const double t0 = std::chrono::duration<double, std::milli>(
        std::chrono::steady_clock::now().time_since_epoch()).count();
    time_transfer = 0;
    time_offDiags = time_bandLU = time_bandUL = time_assembly = time_fullLU = 0;
    tm_.clear();
    tm_.mark(0, stream);
    {
        ProfScope ps(prof, "update_band");
        if (src_ == Source::band) {
            launch_band_to_colmajor<T>(n, k, d_entries, band_in(), stream, sep_in_ ? nullptr : Bul_.get(), tile_);
            bul_filled_ = !sep_in_ && Bul_.get() != nullptr;
        } else {
            SAP_HIP_CHECK(hipMemsetAsync(band_in(), 0, band_in_bytes(), stream));
            if (src_ == Source::csr)
                launch_csr_to_band<double, T>(n, k, src_rp_, src_ci_, d_entries, band_in(), stream);
            else
                launch_update_band<T>(count, d_entries, upd_map_.get(), upd_scale_.get(), band_in(), stream);
        }
    }
    tm_.mark(1, stream);
    tm_.span(&time_toBanded, 0, 1);
    time_cpu_assemble = std::chrono::duration<double, std::milli>(
        std::chrono::steady_clock::now().time_since_epoch()).count() - t0;
    *rc = factor_band();
    return true;
}

template <class T>
int Precond<T>::factor_band()
{
    const int rc = factor_band_steps();
    bul_filled_ = false;
    if (side_.s) SAP_HIP_CHECK(hipStreamSynchronize(side_.s));   // also on the early returns
    tm_.resolve();
    return rc;
}

// Spike tips of the variable-bandwidth path: every packed coupling column
// a full-length right-hand side of its partition, solved by the staged
// job-table sweep on the partition's band (one job per column, in chunks of
// columns), tip rows gathered into WV.
template <class T>
void Precond<T>::var_spikes()
{
    const PartInfo pi{n, P};
    const int ncol = (int)vcols_h_.size();
    if (ncol == 0) return;
    int64_t maxlen = 0;
    for (int p = 0; p < P; ++p) {
        int64_t s, len;
        pi.bounds(p, s, len);
        maxlen = std::max(maxlen, len);
    }
    const int64_t stride = (maxlen + 63) / 64 * 64;
    // columns per chunk: about 1 GiB of FP64 right-hand sides, solutions and
    // T scratch at a time
    const int chunk = (int)std::max<int64_t>(1, std::min<int64_t>(ncol, ((int64_t)1 << 30) /
                                                                             (stride * (16 + (int64_t)sizeof(T)))));
    DevBuf<double> V((size_t)stride * chunk), Z((size_t)stride * chunk);
    DevBuf<T> scr((size_t)stride * chunk);
    DevBuf<SweepJob> jd(chunk);
    for (int c0 = 0; c0 < ncol; c0 += chunk) {
        const int nc = std::min(chunk, ncol - c0);
        launch_var_rhs<T>(pi, k, sp_.get(), voffB(), voffC(), vperm_r_.get(), vperm_l_.get(), vwidth_.get(),
                          vcols_.get() + c0, nc, stride, V.get(), stream);
        std::vector<SweepJob> by_mode[3];
        size_t lds[3] = {0, 0, 0};
        for (int j = 0; j < nc; ++j) {
            const VarSpikeCol& c = vcols_h_[c0 + j];
            const int p = c.i + c.side;
            int64_t s, len;
            pi.bounds(p, s, len);
            std::vector<SweepJob> one;
            size_t l = 0;
            const int mode = sweep_range_job_append<T>(one, l, Bv_.get() + vboff_[p], vks_[p], 0, len,
                                                       scr.get() + (int64_t)j * stride, len, 0, 0, 1);
            one[0].voff = one[0].ooff = (int64_t)j * stride;
            by_mode[mode].push_back(one[0]);
            lds[mode] = std::max(lds[mode], l);
        }
        std::vector<SweepJob> jobs;
        size_t start[3];
        for (int m = 0; m < 3; ++m) {
            start[m] = jobs.size();
            jobs.insert(jobs.end(), by_mode[m].begin(), by_mode[m].end());
        }
        SAP_HIP_CHECK(hipMemcpyAsync(jd.get(), jobs.data(), jobs.size() * sizeof(SweepJob), hipMemcpyHostToDevice,
                                     stream));
        for (int m = 2; m >= 0; --m)
            if (!by_mode[m].empty())
                launch_sweep_full_jobs<T>(jd.get() + start[m], (int)by_mode[m].size(), lds[m], V.get(), Z.get(),
                                          nullptr, stream, m);
        launch_var_tips<T>(pi, k, sp_.get(), vperm_r_.get(), vperm_l_.get(), vwidth_.get(), vcols_.get() + c0, nc,
                           stride, Z.get(), WV_.get(), stream);
        SAP_HIP_CHECK(hipStreamSynchronize(stream));   // the job table is rewritten, the buffers released
    }
}

// Variable-bandwidth setup after the fill (sap/precond.h:4546-4631 LU of the
// partitions in their own bands, :5540-5790 spikes, :2549-2611 reduced
// blocks): every pivot boosted, as in the reference's variable-bandwidth
// kernels.
template <class T>
int Precond<T>::var_factor_steps()
{
    const PartInfo pi{n, P};
    cnt_.resize(1);
    scr0_.resize(n);
    split_ = false;
    vjobs_[0].n = vjobs_[1].n = 0;   // the sweep job tables are built again for these bands
    offB_.release(); offC_.release(); Bul_.release(); scr1_.release();
    // one launch per factorization kernel; the jobs of a launch keep their
    // own widths (FactorJob::kb) and band offsets
    std::vector<std::pair<int, int>> order;   // (kernel class, partition)
    for (int p = 0; p < P; ++p) {
        int64_t s, len;
        pi.bounds(p, s, len);
        const int cls = band_factor_kernel_class<T>(vks_[p], len);
        if (cls < 0) throw Error(SAP_ERR_UNSUPPORTED, "half-bandwidth too large for the banded factorization kernel");
        order.push_back({cls, p});
    }
    std::stable_sort(order.begin(), order.end(),
                     [](const std::pair<int, int>& a, const std::pair<int, int>& b) { return a.first < b.first; });
    std::vector<FactorJob> jobs;
    for (const auto& o : order) {
        int64_t s, len;
        pi.bounds(o.second, s, len);
        FactorJob J{0, len, 1, 1, 1, 0, vks_[o.second], vboff_[o.second]};
        jobs.push_back(J);
    }
    const bool same = jobs.size() == jobs_host_.size() &&
                      std::equal(jobs.begin(), jobs.end(), jobs_host_.begin(), [](const FactorJob& a, const FactorJob& b) {
                          return a.start == b.start && a.len == b.len && a.dir == b.dir && a.boost == b.boost &&
                                 a.boost_last == b.boost_last && a.buf == b.buf && a.kb == b.kb && a.boff == b.boff;
                      });
    if (!same) {
        jobs_.resize(jobs.size());
        SAP_HIP_CHECK(hipMemcpyAsync(jobs_.get(), jobs.data(), jobs.size() * sizeof(FactorJob), hipMemcpyHostToDevice,
                                     stream));
        jobs_host_ = jobs;
    }
    tm_.mark(3, stream);
    {
        ProfScope ps(prof, "band_factor");
        size_t a = 0;
        while (a < order.size()) {
            size_t b = a;
            int kmax = 0;
            int64_t max_len = 0;
            while (b < order.size() && order[b].first == order[a].first) {
                kmax = std::max(kmax, jobs[b].kb);
                max_len = std::max(max_len, jobs[b].len);
                ++b;
            }
            launch_band_factor<T>(Bv_.get(), Bv_.get(), nullptr, kmax, jobs_.get() + a, (int)(b - a), max_len, stream);
            a = b;
        }
    }
    tm_.mark(4, stream);
    tm_.span(&time_bandLU, 3, 4);
    actual_nnz = vboff_[P] - count_band_zeros<T>(Bv_.get(), vboff_[P], (const T*)nullptr, 0, cnt_.get(), stream);

    const size_t kk = (size_t)k * k;
    if (spike_active())
        (void)try_spike_buffers(
            [&] {
                g_.resize(n);
                gd_.resize(n);
                vq_.resize(n);
                WV_.resize(2 * kk * (P - 1));
                S_.resize(kk * (P - 1));
                Sinv_.resize(kk * (P - 1));
                corr_.resize((size_t)2 * k * (P - 1));
            },
            [&] {
                WV_.release(); S_.release(); Sinv_.release(); corr_.release(); g_.release(); gd_.release();
                vq_.release();
            });
    if (!spike_active()) {
        WV_.release(); S_.release(); Sinv_.release(); corr_.release(); g_.release(); gd_.release(); vq_.release();
        return 0;
    }
    tm_.mark(5, stream);
    {
        ProfScope ps(prof, "spike_tips");
        SAP_HIP_CHECK(hipMemsetAsync(WV_.get(), 0, WV_.size() * sizeof(T), stream));
        var_spikes();
    }
    tm_.mark(6, stream);
    {
        ProfScope ps(prof, "reduced_factor");
        launch_reduced_factor<T>(WV_.get(), k, P - 1, S_.get(), Sinv_.get(), cfg.safeFactorization, stream);
    }
    tm_.mark(7, stream);
    tm_.span(&time_assembly, 5, 6);
    tm_.span(&time_fullLU, 6, 7);
    return 0;
}

// Full sweeps (L, D^{-1}, U') of every partition on its own band, second-
// level order in and out: one launch over all partitions per sweep variant
// their widths need.
template <class T>
typename Precond<T>::VJobs& Precond<T>::var_jobs(int nr, int64_t ldv, int64_t ldz, T* scr)
{
    VJobs& J = vjobs_[nr == 1 ? 0 : 1];
    const VJobKey key{nr, ldv, ldz, scr};
    if (key.nr == J.key.nr && key.ldv == J.key.ldv && key.ldz == J.key.ldz && key.scr == J.key.scr && J.n > 0)
        return J;
    const PartInfo pi{n, P};
    // the jobs grouped by sweep variant (one launch each)
    std::vector<SweepJob> by_mode[3];
    for (int p = 0; p < P; ++p) {
        int64_t s, len;
        pi.bounds(p, s, len);
        const T* Bp = Bv_.get() + vboff_[p];
        std::vector<SweepJob> one;
        size_t lds = 0;
        int mode = 0;
        if (!sweep_jobs_append<T>(one, lds, Bp, Bp, vks_[p], PartInfo{len, 1}, 0, nullptr, scr + nr * s, s, ldv, s,
                                  ldz, nr, 0, 0, &mode))
            throw Error(SAP_ERR_HIP, "variable-bandwidth sweep job table");
        by_mode[mode].insert(by_mode[mode].end(), one.begin(), one.end());
        J.lds[mode] = std::max(J.lds[mode], lds);
    }
    std::vector<SweepJob> jobs;
    for (int m = 0; m < 3; ++m) {
        J.start[m] = jobs.size();
        J.count[m] = by_mode[m].size();
        jobs.insert(jobs.end(), by_mode[m].begin(), by_mode[m].end());
    }
    J.n = jobs.size();
    J.d.resize(jobs.size());
    SAP_HIP_CHECK(hipMemcpyAsync(J.d.get(), jobs.data(), jobs.size() * sizeof(SweepJob), hipMemcpyHostToDevice,
                                 stream));
    SAP_HIP_CHECK(hipStreamSynchronize(stream));   // the host table is released on return
    J.key = key;
    return J;
}

template <class T>
void Precond<T>::var_sweeps(const double* V, int64_t ldv, double* Z, int64_t ldz, int nr, const int* gate, T* scr)
{
    const VJobs& J = var_jobs(nr, ldv, ldz, scr);
    for (int m = 2; m >= 0; --m)
        if (J.count[m])
            launch_sweep_full_jobs<T>(J.d.get() + J.start[m], (int)J.count[m], J.lds[m], V, Z, gate, stream, m);
}

// getSRev's variable-bandwidth branch (sap/precond.h:1457-1470, then the
// final sweeps :1525-1526): sweeps of every partition (second-level order),
// the interface values gathered in first-level order, the reduced solve and
// purification, the purified right-hand side back in second-level order,
// sweeps again.  Block: one sweep pass.
template <class T>
void Precond<T>::var_solve(const double* V, int64_t ldv, double* Z, int64_t ldz, int nr, const int* gate)
{
    const PartInfo pi{n, P};
    T* scr = nr == 1 ? scr0_.get() : scr0m_.get();
    if (!spike_active()) {
        ProfScope ps(prof, "sweep_full");
        var_sweeps(V, ldv, Z, ldz, nr, gate, scr);
        return;
    }
    double* gd = nr == 1 ? gd_.get() : gdm_.get();
    double* vq = nr == 1 ? vq_.get() : vqm_.get();
    T* g = nr == 1 ? g_.get() : gm_.get();
    T* corr = nr == 1 ? corr_.get() : corrm_.get();
    {
        ProfScope ps(prof, "sweep_interface");
        var_sweeps(V, ldv, gd, n, nr, gate, scr);
    }
    {
        ProfScope ps(prof, "interface_solve");
        launch_var_gather<T>(pi, k, sp_.get(), gd, n, g, nr, gate, stream);
        Rhs rh;
        rh.nr = nr;
        rh.gate = gate;
        launch_interface_solve<T>(WV_.get(), Sinv_.get(), voffB(), voffC(), k, pi, g, corr, stream, rh);
        launch_var_purify<T>(pi, k, sr_.get(), V, ldv, corr, vq, n, nr, gate, stream);
    }
    ProfScope ps(prof, "sweep_full");
    var_sweeps(vq, n, Z, ldz, nr, gate, scr);
}

// Block cyclic reduction of every partition (the reference's useBCR setup,
// sap/precond.h:981-1210): blocks of at least 2k rows, level l eliminating
// the blocks s, 3s, 5s, ... (s = 2^l) of every partition together, each
// partition's block 0 factored last.  Like the reference's BCR setup it
// builds the block-diagonal preconditioner (no spikes); pivots are boosted
// by the Block rules, and there is no zero-pivot check (the reference's BCR
// branch returns before it).
template <class T>
int Precond<T>::bcr_factor_steps()
{
    const PartInfo pi{n, P};
    cnt_.resize(1);
    scr0_.resize(n);
    split_ = false;
    offB_.release(); offC_.release(); WV_.release(); S_.release(); Sinv_.release(); corr_.release();
    scr1_.release();
    bblk_.clear();
    blev_.clear();
    bfinal_.clear();
    std::vector<int> first(P), nb(P);
    int64_t max_len = 0;
    for (int p = 0; p < P; ++p) {
        int64_t s, len;
        pi.bounds(p, s, len);
        nb[p] = (int)std::max<int64_t>(1, len / (2 * (int64_t)k));
        first[p] = (int)bblk_.size();
        const PartInfo bi{len, nb[p]};
        for (int j = 0; j < nb[p]; ++j) {
            int64_t bs, bl;
            bi.bounds(j, bs, bl);
            bblk_.push_back({s + bs, (int)bl, j > 0, j < nb[p] - 1});
            max_len = std::max(max_len, bl);
        }
        bfinal_.push_back(first[p]);
    }
    for (int st = 1;; st *= 2) {
        BcrLevel L;
        for (int p = 0; p < P; ++p)
            for (int e = st; e < nb[p]; e += 2 * st)
                L.el.push_back({first[p] + e, first[p] + e - st, e + st < nb[p] ? first[p] + e + st : -1});
        if (L.el.empty()) break;
        blev_.push_back(std::move(L));
    }
    const size_t nblk = bblk_.size(), kk = (size_t)k * k;
    upload(bblk_d_, bblk_, stream);
    for (BcrLevel& L : blev_) upload(L.d, L.el, stream);
    bCL_.resize(nblk * kk);
    bCU_.resize(nblk * kk);
    bUA_.resize(nblk * kk);
    bLC_.resize(nblk * kk);
    bw_.resize(n);
    bt_.resize(n);
    bjobs_[0].nr = bjobs_[1].nr = 0;   // the sweep job tables are built again for these blocks
    const int boost = (k >= 70) || cfg.safeFactorization;
    const int boost_last = (k >= 70);
    tm_.mark(3, stream);
    {
        ProfScope ps(prof, "band_factor");
        launch_bcr_extract<T>(B_.get(), k, bblk_d_.get(), (int)nblk, bCL_.get(), bCU_.get(), stream);
        for (const BcrLevel& L : blev_) {
            std::vector<FactorJob> jobs;
            for (const BcrElim& E : L.el) jobs.push_back({bblk_[E.e].start, bblk_[E.e].len, 1, boost, boost_last, 0});
            factor(jobs);
            launch_bcr_level_setup<T>(B_.get(), k, bblk_d_.get(), L.d.get(), (int)L.el.size(), max_len, bCL_.get(),
                                      bCU_.get(), bUA_.get(), bLC_.get(), stream);
        }
        std::vector<FactorJob> jobs;
        for (int b : bfinal_) jobs.push_back({bblk_[b].start, bblk_[b].len, 1, boost, boost_last, 0});
        factor(jobs);
    }
    tm_.mark(4, stream);
    tm_.span(&time_bandLU, 3, 4);
    const int64_t w = 2 * k + 1;
    actual_nnz = (int64_t)w * n - count_band_zeros<T>(B_.get(), (int64_t)w * n, (const T*)nullptr, 0, cnt_.get(),
                                                      stream);
    bcr_ = true;
    return 0;
}

// The sweep jobs of the BCR apply: stage l < L the blocks eliminated at
// level l, stage L every partition's last block; input and output columns n
// apart.
template <class T>
typename Precond<T>::BcrJobs& Precond<T>::bcr_jobs(int nr, T* scr)
{
    BcrJobs& J = bjobs_[nr == 1 ? 0 : 1];
    if (J.nr == nr && J.scr == scr) return J;
    std::vector<SweepJob> jobs;
    size_t lds = 0;
    int mode = 0;
    J.start.clear();
    J.count.clear();
    auto add = [&](int b) {
        mode = sweep_range_job_append<T>(jobs, lds, B_.get(), k, bblk_[b].start, bblk_[b].len, scr, n, n, n, nr);
    };
    for (const BcrLevel& L : blev_) {
        J.start.push_back(jobs.size());
        for (const BcrElim& E : L.el) add(E.e);
        J.count.push_back(jobs.size() - J.start.back());
    }
    J.start.push_back(jobs.size());
    for (int b : bfinal_) add(b);
    J.count.push_back(jobs.size() - J.start.back());
    J.d.resize(jobs.size());
    SAP_HIP_CHECK(hipMemcpyAsync(J.d.get(), jobs.data(), jobs.size() * sizeof(SweepJob), hipMemcpyHostToDevice,
                                 stream));
    SAP_HIP_CHECK(hipStreamSynchronize(stream));   // the host table is released on return
    J.lds = lds;
    J.mode = mode;
    J.nr = nr;
    J.scr = scr;
    return J;
}

// The BCR apply (updateRHS, sap/precond.h:6485-6612): forward through the
// levels, the last blocks, back-substitution through the levels.
template <class T>
void Precond<T>::bcr_solve(const double* V, int64_t ldv, double* Z, int64_t ldz, int nr, const int* gate)
{
    T* scr = nr == 1 ? scr0_.get() : scr0m_.get();
    double* w = nr == 1 ? bw_.get() : bwm_.get();
    double* t = nr == 1 ? bt_.get() : btm_.get();
    const BcrJobs& J = bcr_jobs(nr, scr);
    const int L = (int)blev_.size();
    auto sweep = [&](int stage, const double* in, double* out) {
        launch_sweep_full_jobs<T>(J.d.get() + J.start[stage], (int)J.count[stage], J.lds, in, out, gate, stream,
                                  J.mode);
    };
    ProfScope ps(prof, "sweep_full");
    for (int r = 0; r < nr; ++r) launch_copy(n, V + r * ldv, w + r * n, stream);
    for (int l = 0; l < L; ++l) {
        sweep(l, w, t);
        launch_bcr_gemv<T>(0, k, bblk_d_.get(), blev_[l].d.get(), (int)blev_[l].el.size(), bUA_.get(), bLC_.get(), t,
                           n, w, n, nr, gate, stream);
    }
    sweep(L, w, Z);
    for (int l = L - 1; l >= 0; --l) {
        launch_bcr_gemv<T>(1, k, bblk_d_.get(), blev_[l].d.get(), (int)blev_[l].el.size(), bCL_.get(), bCU_.get(), Z,
                           ldz, w, n, nr, gate, stream);
        sweep(l, w, Z);
    }
}

template <class T>
int Precond<T>::factor_band_steps()
{
    bcr_ = false;
    if (var_mode_) return var_factor_steps();
    if (cfg.bcr && k > 0 && cfg.precondType != SAP_None) return bcr_factor_steps();
    const PartInfo pi{n, P};
    const int64_t w = 2 * k + 1;
    // timing events: 2 before the coupling blocks, 3 before and 4 after the
    // factorization, 5 and 6 around the spike tips, 7 after the reduced blocks
    cnt_.resize(1);
    g_.resize(n);
    scr0_.resize(n);
    split_ = false;

    if (!spike_active()) {
        // buffers of the SaP-C steps are not needed (and not reported)
        offB_.release(); offC_.release(); WV_.release(); S_.release(); Sinv_.release(); corr_.release();
        scr1_.release();
    }
    if (k == 0 || cfg.precondType == SAP_None) {
        if (cfg.precondType != SAP_None)
            actual_nnz = n - count_band_zeros<T>(B_.get(), n, (const T*)nullptr, 0, cnt_.get(), stream);
        return 0;
    }

    // sap/precond.h:4363-4397; the variable-bandwidth kernels boost every
    // pivot (sap/device/factor_band_var.cuh)
    const int boost = (k >= 70) || cfg.safeFactorization || var_mode_;
    const int boost_last = (k >= 70) || var_mode_;
    // With every pivot boosted (to +-1e-4 when |d| <= 1e-7, NaN included)
    // no stored pivot can be below 1e-7, so the check is skipped.
    const bool check_pivots = !cfg.safeFactorization && !(boost && boost_last);

    const size_t kk = (size_t)k * k;
    if (!spike_active()) {
        std::vector<FactorJob> jobs;
        for (int p = 0; p < P; ++p) {
            int64_t s, len;
            pi.bounds(p, s, len);
            jobs.push_back({s, len, 1, boost, boost_last, 0});
        }
        tm_.mark(3, stream);
        factor(jobs);
        tm_.mark(4, stream);
        tm_.span(&time_bandLU, 3, 4);
        if (check_pivots && zero_pivots(B_.get(), 0, n)) return SAP_ERR_ZERO_PIVOT;
        actual_nnz = (int64_t)w * n - count_band_zeros<T>(B_.get(), (int64_t)w * n, (const T*)nullptr, 0,
                                                          cnt_.get(), stream);
        return 0;
    }

    // (2) coupling blocks, extracted before factorization
    tm_.mark(2, stream);
    {
        ProfScope ps(prof, "extract_coupling");
        launch_extract_coupling<T>(band_in(), k, pi, offB_.get(), offC_.get(), stream);
    }
    tm_.mark(3, stream);

    // (3) factorization: LU of partitions 0..P-2 (LU_UL) or 0..P-1 (LU_only);
    //     for LU_UL also UL of partitions 1..P-1 on a copy of the band, in the
    //     same launch (or both from the one copy in Bin_).
    std::vector<FactorJob> jobs;
    int64_t s_last, l_last;
    pi.bounds(P - 1, s_last, l_last);
    if (lu_ul()) {
        if (sep_in_) {
            // Band entries no factorization writes keep the matrix, as with
            // in-place factorization: the last partition in B_, the first in
            // Bul_, and in the others the entries coupling a partition to
            // its neighbours (rows outside it, in its first and last k
            // columns; the factorization then overwrites the rest of those
            // columns).
            int64_t s0, l0;
            pi.bounds(0, s0, l0);
            SAP_HIP_CHECK(hipMemcpyAsync(B_.get() + (size_t)w * s_last, Bin_.get() + (size_t)w * s_last,
                                         sizeof(T) * w * l_last, hipMemcpyDeviceToDevice, stream));
            SAP_HIP_CHECK(hipMemcpyAsync(Bul_.get(), Bin_.get(), sizeof(T) * w * l0, hipMemcpyDeviceToDevice, stream));
            launch_copy_partition_edges<T>(Bin_.get(), B_.get(), k, pi, 0, P - 1, stream);
            launch_copy_partition_edges<T>(Bin_.get(), Bul_.get(), k, pi, 1, P, stream);
        } else if (!bul_filled_) {
            SAP_HIP_CHECK(hipMemcpyAsync(Bul_.get(), B_.get(), sizeof(T) * w * n, hipMemcpyDeviceToDevice, stream));
        }
        for (int p = 0; p < P - 1; ++p) {
            int64_t s, len;
            pi.bounds(p, s, len);
            jobs.push_back({s, len, 1, boost, boost_last, 0});
        }
        for (int p = 1; p < P; ++p) {
            int64_t s, len;
            pi.bounds(p, s, len);
            jobs.push_back({s, len, -1, boost, boost_last, 1});
        }
        // With every pivot boosted (no factorization can fail), also the LU
        // of the last partition and the UL of the first (two more jobs
        // beside P - 1 + P - 1 on a chip of 256 CUs): the full sweep then
        // gives each partition two workgroups, rows below its middle from
        // the LU factors and rows above from the UL factors.
        split_ = boost && boost_last;
        if (split_) {
            int64_t s0, l0;
            pi.bounds(0, s0, l0);
            jobs.push_back({s_last, l_last, 1, boost, boost_last, 0});
            jobs.push_back({s0, l0, -1, boost, boost_last, 1});
        }
        scr1_.resize(n);
    } else {
        split_ = false;
        scr1_.release();
        for (int p = 0; p < P; ++p) {
            int64_t s, len;
            pi.bounds(p, s, len);
            jobs.push_back({s, len, 1, boost, boost_last, 0});
        }
    }
    factor(jobs);
    tm_.mark(4, stream);
    tm_.span(&time_offDiags, 2, 3);
    tm_.span(&time_bandLU, 3, 4);
    if (check_pivots) {
        if (lu_ul()) {
            int64_t s1, l1;
            pi.bounds(1, s1, l1);
            if (zero_pivots(B_.get(), 0, s_last, Bul_.get(), s1, n - s1)) return SAP_ERR_ZERO_PIVOT;
        } else if (zero_pivots(B_.get(), 0, n)) {
            return SAP_ERR_ZERO_PIVOT;
        }
    }

    // (4) spike tips and (5) reduced blocks S_i = I - W V_i (LU, inverse),
    // then the non-zeros of the factored band (sap/precond.h:1352: the LU
    // part and, for LU_UL, the UL-factored last partition), on a second
    // stream after the factorization; the setup returns without waiting
    // (finish_pending, wait_spikes)
    hipStream_t s2 = side2_.after(stream);
    if (!spk_done_) SAP_HIP_CHECK(hipEventCreateWithFlags(&spk_done_, hipEventDisableTiming));
    tm2_.clear();
    tm2_.mark(5, s2);
    {
        ProfOnStream po(prof, s2);
        ProfScope ps(prof, "spike_tips");
        launch_spike_tips<T>(B_.get(), Bul_.get(), k, pi, lu_ul() ? 1 : 0, offB_.get(), offC_.get(), WV_.get(), s2);
    }
    tm2_.mark(6, s2);
    {
        ProfOnStream po(prof, s2);
        ProfScope ps(prof, "reduced_factor");
        launch_reduced_factor<T>(WV_.get(), k, P - 1, S_.get(), Sinv_.get(), cfg.safeFactorization, s2);
    }
    tm2_.mark(7, s2);
    SAP_HIP_CHECK(hipEventRecord(spk_done_, s2));
    if (lu_ul())
        count_band_zeros_async<T>(B_.get(), (int64_t)w * s_last, Bul_.get() + (size_t)w * s_last,
                                  (int64_t)w * l_last, cnt_.get(), side2_.host, s2, 1024);
    else
        count_band_zeros_async<T>(B_.get(), (int64_t)w * n, (const T*)nullptr, 0, cnt_.get(), side2_.host, s2,
                                  1024);
    tm2_.span(&time_assembly, 5, 6);
    tm2_.span(&time_fullLU, 6, 7);
    time_assembly = time_fullLU = 0;
    actual_nnz = 0;
    spk_wait_ = spk_open_ = true;
    return 0;
}

template <class T>
void Precond<T>::apply(const double* v, double* z)
{
    if (cfg.precondType == SAP_None) {
        launch_copy(n, v, z, stream);
        return;
    }
    if (!reorder_) {
        solve_band(v, z);
        return;
    }
    // leftTrans (sap/precond.h:1535-1544), getSRev, rightTrans (:1551-1559)
    {
        ProfScope ps(prof, "permute");
        launch_permute_scale<T>(n, v, row_perm_.get(), scale_ ? row_scale_.get() : nullptr, false, vp_.get(), stream);
    }
    if (var_mode_) var_solve(vp_.get(), n, zp_.get(), n, 1, nullptr);
    else solve_band(vp_.get(), zp_.get());
    ProfScope ps(prof, "permute");
    launch_permute_scale<T>(n, zp_.get(), col_reord_.get(), scale_ ? col_scale_.get() : nullptr, true, z, stream);
}

// getSRev (sap/precond.h:1411-1528) for the constant-bandwidth band
template <class T>
void Precond<T>::solve_band(const double* v, double* z)
{
    const PartInfo pi{n, P};
    Rhs rh;
    rh.nbuf = nbuf_;
    if (k == 0) {
        ProfScope ps(prof, "diag_divide");
        launch_diag_divide<T>(B_.get(), n, v, z, stream);
        return;
    }
    if (bcr_) {
        bcr_solve(v, n, z, n, 1, nullptr);
        return;
    }
    if (spike_active()) {
        if (lu_ul()) {
            ProfScope ps(prof, "sweep_interface");
            launch_sweep_interface<T>(B_.get(), Bul_.get(), k, pi, v, g_.get(), scr0_.get(), scr1_.get(), stream, rh);
        } else {
            ProfScope ps(prof, "sweep_interface");
            launch_sweep_full<T>(B_.get(), B_.get(), k, pi, 0, v, nullptr, nullptr, g_.get(), scr0_.get(), stream, rh);
        }
        wait_spikes();
        {
            ProfScope ps(prof, "interface_solve");
            launch_interface_solve<T>(WV_.get(), Sinv_.get(), offB_.get(), offC_.get(), k, pi, g_.get(), corr_.get(),
                                      stream);
        }
        ProfScope ps(prof, "sweep_full");
        Rhs r2 = rh;
        if (split_ && lu_ul()) r2.split_scr = scr1_.get();
        launch_sweep_full<T>(B_.get(), Bul_.get(), k, pi, lu_ul() ? 1 : 0, v, corr_.get(), z, nullptr, scr0_.get(),
                             stream, r2);
    } else {
        ProfScope ps(prof, "sweep_full");
        launch_sweep_full<T>(B_.get(), B_.get(), k, pi, 0, v, nullptr, z, nullptr, scr0_.get(), stream, rh);
    }
}

template <class T>
void Precond<T>::reserve_multi(int nr)
{
    // (a batched solve may capture apply_multi into a graph: the main stream
    // is ordered after the deferred setup work here, outside the capture)
    wait_spikes();
    if (nr > nr_cap_) {
        const size_t m = (size_t)nr;
        gm_.resize(m * n);
        scr0m_.resize(m * n);
        if (spike_active()) {
            corrm_.resize(m * 2 * k * (P - 1));
            if (lu_ul()) scr1m_.resize(m * n);
        }
        if (reorder_) {
            vpm_.resize(m * n);
            zpm_.resize(m * n);
        }
        if (var_mode_ && spike_active()) {
            gdm_.resize(m * n);
            vqm_.resize(m * n);
        }
        nr_cap_ = nr;
    }
    // the job table apply_multi's variable-bandwidth sweeps use (input and
    // output strides n), built here, before any graph capture
    if (var_mode_ && nr > 1) (void)var_jobs(nr, n, n, scr0m_.get());
    if (bcr_ && nr > 1) {
        if (bwm_.size() < (size_t)nr * n) {
            bwm_.resize((size_t)nr * n);
            btm_.resize((size_t)nr * n);
        }
        (void)bcr_jobs(nr, scr0m_.get());
    }
}

template <class T>
bool Precond<T>::append_sweep_jobs(std::vector<SweepJob>& jobs, size_t& lds, int64_t voff, int64_t ooff, int nr,
                                   int gate0)
{
    if (cfg.precondType == SAP_None || reorder_ || spike_active() || k == 0 || bcr_) return false;
    reserve_multi(nr);
    return sweep_jobs_append<T>(jobs, lds, B_.get(), B_.get(), k, PartInfo{n, P}, 0, nullptr, scr0m_.get(), voff, n,
                                ooff, n, nr, gate0, nbuf_);
}

template <class T>
void Precond<T>::apply_multi(const double* V, int64_t ldv, double* Z, int64_t ldz, int nr, const int* gate)
{
    if (nr < 1) return;
    if (cfg.precondType == SAP_None) {
        if (gate) throw Error(SAP_ERR_UNSUPPORTED, "gated apply without a preconditioner");
        for (int r = 0; r < nr; ++r) launch_copy(n, V + r * ldv, Z + r * ldz, stream);
        return;
    }
    reserve_multi(nr);
    if (!reorder_) {
        solve_band_multi(V, ldv, Z, ldz, nr, gate);
        return;
    }
    {
        ProfScope ps(prof, "permute");
        launch_permute_scale<T>(n, V, row_perm_.get(), scale_ ? row_scale_.get() : nullptr, false, vpm_.get(), stream,
                                nr, ldv, n, gate);
    }
    if (var_mode_) var_solve(vpm_.get(), n, zpm_.get(), n, nr, gate);
    else solve_band_multi(vpm_.get(), n, zpm_.get(), n, nr, gate);
    ProfScope ps(prof, "permute");
    launch_permute_scale<T>(n, zpm_.get(), col_reord_.get(), scale_ ? col_scale_.get() : nullptr, true, Z, stream, nr,
                            n, ldz, gate);
}

template <class T>
void Precond<T>::solve_band_multi(const double* V, int64_t ldv, double* Z, int64_t ldz, int nr, const int* gate)
{
    const PartInfo pi{n, P};
    Rhs rh;
    rh.nr = nr;
    rh.ldv = ldv;
    rh.ldo = ldz;
    rh.gate = gate;
    rh.nbuf = nbuf_;
    if (k == 0) {
        ProfScope ps(prof, "diag_divide");
        launch_diag_divide<T>(B_.get(), n, V, Z, stream, rh);
        return;
    }
    if (bcr_) {
        if (ldz != n) throw Error(SAP_ERR_BADARG, "block cyclic reduction: output columns must be n apart");
        bcr_solve(V, ldv, Z, ldz, nr, gate);
        return;
    }
    if (spike_active()) {
        if (lu_ul()) {
            ProfScope ps(prof, "sweep_interface");
            launch_sweep_interface<T>(B_.get(), Bul_.get(), k, pi, V, gm_.get(), scr0m_.get(), scr1m_.get(), stream, rh);
        } else {
            ProfScope ps(prof, "sweep_interface");
            launch_sweep_full<T>(B_.get(), B_.get(), k, pi, 0, V, nullptr, nullptr, gm_.get(), scr0m_.get(), stream, rh);
        }
        wait_spikes();
        {
            ProfScope ps(prof, "interface_solve");
            launch_interface_solve<T>(WV_.get(), Sinv_.get(), offB_.get(), offC_.get(), k, pi, gm_.get(), corrm_.get(),
                                      stream, rh);
        }
        ProfScope ps(prof, "sweep_full");
        Rhs r2 = rh;
        if (split_ && lu_ul()) r2.split_scr = scr1m_.get();
        launch_sweep_full<T>(B_.get(), Bul_.get(), k, pi, lu_ul() ? 1 : 0, V, corrm_.get(), Z, nullptr, scr0m_.get(),
                             stream, r2);
    } else {
        ProfScope ps(prof, "sweep_full");
        launch_sweep_full<T>(B_.get(), B_.get(), k, pi, 0, V, nullptr, Z, nullptr, scr0m_.get(), stream, rh);
    }
}

template class Precond<float>;
template class Precond<double>;

}  // namespace sap_amd
