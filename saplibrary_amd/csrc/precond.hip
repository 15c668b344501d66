// Setup and apply of the SaP preconditioner for the constant-bandwidth
// banded path (sap/precond.h:910-1362 setup, :1370-1528 apply).
#include "precond.hpp"

#include <algorithm>
#include <chrono>

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
void Precond<T>::factor(const std::vector<FactorJob>& jobs)
{
    if (jobs.empty()) return;
    // the job list is uploaded only when it changed (repeated setups of one
    // size reuse it)
    const bool same = jobs.size() == jobs_host_.size() &&
                      std::equal(jobs.begin(), jobs.end(), jobs_host_.begin(), [](const FactorJob& a, const FactorJob& b) {
                          return a.start == b.start && a.len == b.len && a.dir == b.dir && a.boost == b.boost &&
                                 a.boost_last == b.boost_last && a.buf == b.buf;
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

template <class T>
void Precond<T>::begin_setup(int64_t n_, int k_)
{
    n = n_;
    k = k_;
    nr_cap_ = 0;   // apply_multi's buffers are sized again for this setup
    nbuf_ = 0;     // the sweeps' staging depth is tuned again for this k
    time_toBanded = time_offDiags = time_bandLU = time_bandUL = time_assembly = time_fullLU = 0;
    // Partition clamp of convertToBandedMatrix (sap/precond.h:2116-2117,
    // :2265-2266): each partition needs k+1 <= n_p and 2k <= n_p.  Every
    // path uses it: with shorter partitions the 2k-row interface windows of
    // neighbouring boundaries overlap, and the reduced solve and
    // purification of one boundary would change the other's values.
    const int64_t rows = std::max<int64_t>(k + 1, 2 * (int64_t)k);
    const int64_t maxP = std::max<int64_t>(1, n / rows);
    P = (int)std::min<int64_t>(cfg.numPartitions, maxP);
    if (P < 1) P = 1;
    const int64_t w = 2 * k + 1;
    flag_.resize(1);
    // 8 elements of slack: the sweeps copy 16-byte chunks covering each
    // column segment, which may reach a few elements past the last column
    auto alloc_band = [&](DevBuf<T>& b) { b.resize((size_t)w * n + 8); };
    alloc_band(B_);
    // LU_UL factors the UL of partitions 1..P-1 on a second copy of the band
    // (kept across setups of the same size)
    const bool need_ul = lu_ul() && cfg.precondType == SAP_Spike && P > 1 && k > 0;
    if (need_ul) alloc_band(Bul_);
    else Bul_.release();
    // the first partition is the longest (PartInfo gives the n mod P extra
    // rows to the first partitions)
    int64_t s0 = 0, len0 = n;
    PartInfo{n, P}.bounds(0, s0, len0);
    sep_in_ = need_ul && band_factor_resident_fits<T>(k, len0);
    if (sep_in_) alloc_band(Bin_);
    else Bin_.release();
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
    begin_setup(n_, R.k);
    const int Pvar = P;
    var_mode_ = rc.variable && Pvar > 1 && R.k > 0;

    // (4) band of the reordered matrix, vectors of the permutations
    t = wall_now_ms();
    if (var_mode_) {
        const VarBand<T> VB = assemble_var((int)n, R.k, Pvar, R.rp, R.ci, R.val);
        profile = BandProfile{};
        for (int i = 0; i < (int)n; ++i) {
            profile.sum_row += VB.ks_row[i];
            profile.sum_col += VB.ks_col[i];
            profile.sum_prod += (double)VB.ks_row[i] * VB.ks_col[i];
        }
    } else {
        profile = band_profile((int)n, R.rp.data(), R.ci.data(), P);
    }
    time_cpu_assemble = wall_now_ms() - t;
    t = wall_now_ms();
    DevBuf<int> rp_d, ci_d;
    DevBuf<T> val_d;
    upload(rp_d, R.rp, stream);
    upload(ci_d, R.ci, stream);
    upload(val_d, R.val, stream);
    upload(row_perm_, R.row_perm, stream);
    upload(col_reord_, R.col_reord, stream);
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
        for (int r = 0; r < (int)n; ++r)
            for (int e = R.rp[r]; e < R.rp[r + 1]; ++e) {
                const int64_t c = R.ci[e];
                map[R.entry_orig[e]] = c * w + k + r - c;
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
        launch_csr_to_band<T, T>(n, k, rp_d.get(), ci_d.get(), val_d.get(), band_in(), stream);
    }
    SAP_HIP_CHECK(hipEventRecord(ev.e[1], stream));
    time_toBanded = elapsed_ms(ev.e[0], ev.e[1]);   // waits for the fill before the uploads are freed
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

template <class T>
int Precond<T>::factor_band_steps()
{
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

    if (cfg.precondType == SAP_Block || P == 1) {
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
    const size_t kk = (size_t)k * k;
    offB_.resize(kk * (P - 1));
    offC_.resize(kk * (P - 1));
    WV_.resize(2 * kk * (P - 1));
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

    // non-zeros of the factored band (sap/precond.h:1352): the LU part and,
    // for LU_UL, the UL-factored last partition.  The count reads the factors
    // only, so it runs on a second stream beside the spike tips and the
    // reduced blocks.
    {
        hipStream_t side = side_.after(stream);
        if (lu_ul())
            count_band_zeros_async<T>(B_.get(), (int64_t)w * s_last, Bul_.get() + (size_t)w * s_last,
                                      (int64_t)w * l_last, cnt_.get(), side_.host, side, 1024);
        else
            count_band_zeros_async<T>(B_.get(), (int64_t)w * n, (const T*)nullptr, 0, cnt_.get(), side_.host, side,
                                      1024);
    }

    // (4) spike tips
    tm_.mark(5, stream);
    {
        ProfScope ps(prof, "spike_tips");
        launch_spike_tips<T>(B_.get(), Bul_.get(), k, pi, lu_ul() ? 1 : 0, offB_.get(), offC_.get(), WV_.get(),
                             stream);
    }
    tm_.mark(6, stream);

    // (5) reduced blocks: S_i = I - W V_i, LU, inverse
    S_.resize(kk * (P - 1));
    Sinv_.resize(kk * (P - 1));
    corr_.resize((size_t)2 * k * (P - 1));
    {
        ProfScope ps(prof, "reduced_factor");
        launch_reduced_factor<T>(WV_.get(), k, P - 1, S_.get(), Sinv_.get(), cfg.safeFactorization, stream);
    }
    tm_.mark(7, stream);
    tm_.span(&time_assembly, 5, 6);
    tm_.span(&time_fullLU, 6, 7);
    SAP_HIP_CHECK(hipStreamSynchronize(side_.s));
    actual_nnz = (int64_t)w * n - (int64_t)*side_.host;
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
    solve_band(vp_.get(), zp_.get());
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
    if (spike_active()) {
        if (lu_ul()) {
            ProfScope ps(prof, "sweep_interface");
            launch_sweep_interface<T>(B_.get(), Bul_.get(), k, pi, v, g_.get(), scr0_.get(), scr1_.get(), stream, rh);
        } else {
            ProfScope ps(prof, "sweep_interface");
            launch_sweep_full<T>(B_.get(), B_.get(), k, pi, 0, v, nullptr, nullptr, g_.get(), scr0_.get(), stream, rh);
        }
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
    if (nr <= nr_cap_) return;
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
    nr_cap_ = nr;
}

template <class T>
bool Precond<T>::append_sweep_jobs(std::vector<SweepJob>& jobs, size_t& lds, int64_t voff, int64_t ooff, int nr,
                                   int gate0)
{
    if (cfg.precondType == SAP_None || reorder_ || spike_active() || k == 0) return false;
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
    solve_band_multi(vpm_.get(), n, zpm_.get(), n, nr, gate);
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
    if (spike_active()) {
        if (lu_ul()) {
            ProfScope ps(prof, "sweep_interface");
            launch_sweep_interface<T>(B_.get(), Bul_.get(), k, pi, V, gm_.get(), scr0m_.get(), scr1m_.get(), stream, rh);
        } else {
            ProfScope ps(prof, "sweep_interface");
            launch_sweep_full<T>(B_.get(), B_.get(), k, pi, 0, V, nullptr, nullptr, gm_.get(), scr0m_.get(), stream, rh);
        }
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
