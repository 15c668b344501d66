// Internal declarations shared by the HIP translation units of libsap_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sap_c.h"

namespace sap_amd {

// Error carried from a failing call to the C ABI, which turns it into a code.
struct Error : public std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define SAP_HIP_CHECK(expr)                                                              \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess) {                                                          \
            throw ::sap_amd::Error(_e == hipErrorOutOfMemory ? SAP_ERR_OOM : SAP_ERR_HIP, \
                                   std::string(#expr) + ": " + hipGetErrorString(_e));    \
        }                                                                                \
    } while (0)

#define SAP_LAUNCH_CHECK() SAP_HIP_CHECK(hipGetLastError())

// Owning device allocation.
template <class T>
class DevBuf {
public:
    DevBuf() = default;
    explicit DevBuf(size_t n) { resize(n); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
        return *this;
    }
    void resize(size_t n)
    {
        if (n == n_) return;
        release();
        if (n) SAP_HIP_CHECK(hipMalloc(&p_, n * sizeof(T)));
        n_ = n;
    }
    void release()
    {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T* get() const { return p_; }
    size_t size() const { return n_; }

private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

// Partition p of P over n rows: the first n mod P partitions get one more row
// (sap/device/factor_band_const.cuh:1432-1438).
struct PartInfo {
    int64_t n;
    int P;
    __host__ __device__ inline void bounds(int p, int64_t& s, int64_t& len) const
    {
        const int64_t q = n / P, r = n % P;
        if (p < r) { s = p * (q + 1); len = q + 1; }
        else       { s = p * q + r;   len = q; }
    }
    // the partition holding row j
    __host__ __device__ inline int part_of(int64_t j) const
    {
        const int64_t q = n / P, r = n % P;
        return j < r * (q + 1) ? (int)(j / (q + 1)) : (int)(r + (j - r * (q + 1)) / q);
    }
};

// Kernel-family timer (HIP events on the solver stream), used when profiling.
class Profiler {
public:
    struct Entry { std::string name; int launches = 0; double ms = 0; };
    bool enabled = false;
    hipStream_t stream = nullptr;
    void begin(const char* name);
    void end();
    void flush();          // resolve pending event pairs into totals (syncs)
    void reset();
    std::vector<Entry> entries;
    ~Profiler();

private:
    struct Pending { int idx; hipEvent_t a, b; };
    std::vector<Pending> pending_;
    std::vector<hipEvent_t> pool_;
    int cur_ = -1;
    hipEvent_t cur_a_ = nullptr;
    hipEvent_t get_event();
};

struct ProfScope {
    Profiler* p;
    ProfScope(Profiler* p_, const char* name) : p(p_) { if (p && p->enabled) p->begin(name); }
    ~ProfScope() { if (p && p->enabled) p->end(); }
};

// Pivot boosting (sap/device/factor_band_const.cuh:29-40, thresholds
// BURST_VALUE = 1e-7 and BURST_NEW_VALUE = 1e-4 of sap/common.h:46-47).
template <class T>
__device__ inline T boost_pivot(T v)
{
    const T thr = (T)1e-7, nv = (T)1e-4;
    if (v > thr || v < -thr) return v;
    return v < (T)0 ? -nv : nv;
}

// Workgroup barrier for data exchanged through LDS only.  __syncthreads()
// also waits for every outstanding global load and store of the wavefront
// (vmcnt(0)); kernels whose wavefronts share data only through LDS use this
// instead, so prefetches and result stores stay in flight across it.
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// ---------------------------------------------------------------- launchers
// band.hip
void launch_gen_banded(int64_t n, int k, double dom, uint64_t seed, double* band, hipStream_t st);
void launch_gen_vector(int64_t n, uint64_t seed, double* x, hipStream_t st);
void launch_band_spmv(int64_t n, int k, const double* band, const double* x, double* y, hipStream_t st);
// Y(:, c) = A X(:, c), c < nr (column c at X + c*ldx, Y + c*ldy), each
// column bit-identical to launch_band_spmv / launch_csr_spmv of it alone;
// gate (device, nr flags, may be null) skips the columns whose flag is 0
void launch_band_spmm(int64_t n, int k, const double* band, const double* X, int64_t ldx, double* Y, int64_t ldy,
                      int nr, const int* gate, hipStream_t st);
void launch_csr_spmm(int64_t n, int64_t nnz, const int* rp, const int* ci, const double* val, const double* X,
                     int64_t ldx, double* Y, int64_t ldy, int nr, const int* gate, hipStream_t st);
// tile: columns per workgroup (0: the widest that fits, at most 64; the
// arithmetic is a copy, so any tile gives the same band)
template <class T>
void launch_band_to_colmajor(int64_t n, int k, const double* band_rm, T* B, hipStream_t st, T* B2 = nullptr,
                             int tile = 0);
template <class T>
int band_to_colmajor_max_tile(int k);
// columns [s, s + k) and [s + len - k, s + len) of partitions p0 .. p1-1
// (each partition has at least 2k rows), in -> out
template <class T>
void launch_copy_partition_edges(const T* in, T* out, int k, PartInfo pi, int p0, int p1, hipStream_t st);
template <class Tin, class T>   // B zero-filled beforehand
void launch_csr_to_band(int64_t n, int k, const int* rp, const int* ci, const Tin* vals, T* B, hipStream_t st);
// B[map[e]] = (T)entries[e] * scale[e] for map[e] >= 0 (scale may be null)
template <class T>
void launch_update_band(int64_t nnz, const double* entries, const int64_t* map, const T* scale, T* B,
                        hipStream_t st);
// atomic max of max|row - col| into *out (zeroed beforehand)
void launch_csr_bandwidth(int64_t n, const int* rp, const int* ci, int* out, hipStream_t st);
void launch_csr_spmv(int64_t n, int64_t nnz, const int* rp, const int* ci, const double* val, const double* x,
                     double* y, hipStream_t st);
// w[perm[i]] = (T)v[i] * scale[scale_at_target ? perm[i] : i]  (scale may be null)
// (nr columns at v + r*ldv -> w + r*ldw; gate as for the sweeps)
template <class T>
void launch_permute_scale(int64_t n, const double* v, const int* perm, const T* scale, bool scale_at_target,
                          double* w, hipStream_t st, int nr = 1, int64_t ldv = 0, int64_t ldw = 0,
                          const int* gate = nullptr);

// blas.hip (FP64 vectors, results of reductions left in device memory)
void launch_dot(int64_t n, const double* a, const double* b, double* d_partials, double* d_out, hipStream_t st);
// out[q] = a[q] . b[q] for q < nd (1..3), one pass; out may be mapped host memory
void launch_dots(int64_t n, int nd, const double* const* a, const double* const* b, double* d_partials,
                 double* out, hipStream_t st);
void launch_axpby(int64_t n, double alpha, const double* x, double beta, const double* y, double* z, hipStream_t st);
void launch_copy(int64_t n, const double* x, double* y, hipStream_t st);
// y += sign * (num / *den) x with *den in device memory (as launch_axpby with
// alpha = sign * (num / den) and beta = 1)
void launch_axpy_ratio(int64_t n, double num, const double* den, double sign, const double* x, double* y,
                       hipStream_t st);
void launch_fill(int64_t n, double v, double* x, hipStream_t st);
// dst[q] = src (q < nd <= 4) and zero[q] = 0 (q < nz <= 4) in one launch
void launch_spread(int64_t n, const double* src, double* const* dst, int nd, double* const* zero, int nz,
                   hipStream_t st);
int  dot_partials_size();

// Batched vectors: column c of a batch vector spans [off[c], off[c] +
// len[c]) (device arrays of cols entries); maxlen = the longest column.
// Every batched kernel gives each column the arithmetic of the
// single-vector kernel on that column alone (the same bits); columns whose
// gate flag (device array of cols ints, may be null) is 0 are skipped.
struct BatchCols {
    const int64_t* off = nullptr;
    const int* len = nullptr;
    int cols = 0;
    int64_t maxlen = 0;
};
// coefficient of column c: sign * S[c * s_stride + field] for field >= 0,
// else cst
struct BCoef {
    int field = -1;
    double sign = 1.0;
    double cst = 0.0;
};
inline BCoef bconst(double v) { BCoef k; k.cst = v; return k; }
inline BCoef bfield(int f, double sign = 1.0) { BCoef k; k.field = f; k.sign = sign; return k; }
int  bdot_partials_size(int cols);
// best of reps timed streaming reads of n doubles (n even, 16-byte aligned)
double stream_read_best_ms(const double* x, int64_t n, int reps, hipStream_t st);
// out[c * out_stride + q] = a[q](:, c) . b[q](:, c) for q < nd (1..3)
void launch_bdots(const BatchCols& bc, int nd, const double* const* a, const double* const* b, const int* gate,
                  double* d_partials, double* out, int out_stride, hipStream_t st);
void launch_baxpby(const BatchCols& bc, BCoef ka, const double* x, BCoef kb, const double* y, double* z,
                   const double* S, int s_stride, const int* gate, hipStream_t st);
void launch_bcopy(const BatchCols& bc, const double* x, double* y, const int* gate, hipStream_t st);
void launch_bspread(const BatchCols& bc, const double* src, double* const* dst, int nd, double* const* zero, int nz,
                    const int* gate, hipStream_t st);

// factor.hip
struct FactorJob {          // one banded LU (dir = +1) or UL (dir = -1) of a partition
    int64_t start;          // first row of the partition
    int64_t len;            // rows
    int     dir;
    int     boost;          // boost pivots
    int     boost_last;     // boost also the partition's final pivot
    int     buf;            // 0: main factor band, 1: UL band
    int     kb;             // the job's half-bandwidth (0: the launch's k)
    int64_t boff;           // element offset of the job's band in its buffer
                            // (a partition stored in a band of its own width)
};
// max_len: longest partition among the jobs.  Bin: the unfactored band the
// jobs read (the factors go to B0 / B1), or null for in-place factorization;
// a non-null Bin needs band_factor_resident_fits(k, max_len).
template <class T>
void launch_band_factor(T* B0, T* B1, const T* Bin, int k, const FactorJob* d_jobs, int njobs, int64_t max_len,
                        hipStream_t st);
// factor_resident.hip: register-resident MFMA factorization; false if k is
// too wide for it or a partition too long (the caller then uses the
// LDS-blocked kernels)
template <class T>
bool band_factor_resident_fits(int k, int64_t max_len);
template <class T>
bool launch_band_factor_resident(T* B0, T* B1, const T* Bin, int k, const FactorJob* d_jobs, int njobs,
                                 int64_t max_len, hipStream_t st);
// Which factorization kernel launch_band_factor uses for half-bandwidth k
// (jobs of one launch with their own widths, FactorJob::kb, must share it;
// the launch then takes the largest of their widths as its k).
template <class T>
int band_factor_kernel_class(int k, int64_t max_len);
template <class T>
void launch_zero_pivot_check(const T* B, int k, int64_t row0, int64_t rows, int* d_flag, hipStream_t st);
template <class T>
void launch_extract_coupling(const T* B, int k, PartInfo pi, T* offB, T* offC, hipStream_t st);
template <class T>
void launch_spike_tips(const T* B, const T* Bul, int k, PartInfo pi, int fact_lu_ul,
                       const T* offB, const T* offC, T* WV, hipStream_t st);
// bcr.hip: block cyclic reduction of the partitions (see there)
struct BcrBlock {
    int64_t start;           // first row
    int len;                 // rows (>= 2k, or the whole partition)
    int has_left, has_right; // neighbours inside the partition
};
struct BcrElim {
    int e, a, c;             // eliminated block, its left and right neighbours (c = -1: none)
};
template <class T>
void launch_bcr_extract(const T* B, int k, const BcrBlock* blk, int nblk, T* CL, T* CU, hipStream_t st);
template <class T>
void launch_bcr_level_setup(T* B, int k, const BcrBlock* blk, const BcrElim* el, int nel, int64_t max_len, T* CL,
                            T* CU, T* UA, T* LC, hipStream_t st);
template <class T>
void launch_bcr_gemv(int mode, int k, const BcrBlock* blk, const BcrElim* el, int nel, const T* M0, const T* M1,
                     const double* src, int64_t lds, double* wv, int64_t ldw, int nr, const int* gate,
                     hipStream_t st);
// var.hip: the variable-bandwidth path (per-partition bands in second-level
// order; see there)
template <class T>
void launch_csr_to_varband(PartInfo pi, const int* rp, const int* ci, const T* val, const int* ks,
                           const int64_t* boff, T* Bv, hipStream_t st);
struct VarSpikeCol {
    int i, side, q;   // boundary, 0: V_i (right spike) / 1: W_{i+1} (left spike), packed column
};
template <class T>
void launch_var_rhs(PartInfo pi, int k, const int* sp, const T* offB, const T* offC, const int* perms_r,
                    const int* perms_l, const int* widths, const VarSpikeCol* cols, int ncol, int64_t stride,
                    double* V, hipStream_t st);
template <class T>
void launch_var_tips(PartInfo pi, int k, const int* sp, const int* perms_r, const int* perms_l, const int* widths,
                     const VarSpikeCol* cols, int ncol, int64_t stride, const double* Z, T* WV, hipStream_t st);
template <class T>
void launch_var_gather(PartInfo pi, int k, const int* sp, const double* gd, int64_t ldgd, T* g, int nr,
                       const int* gate, hipStream_t st);
template <class T>
void launch_var_purify(PartInfo pi, int k, const int* sr, const double* v, int64_t ldv, const T* corr, double* out,
                       int64_t ldo, int nr, const int* gate, hipStream_t st);
template <class T>
void launch_reduced_factor(const T* WV, int k, int nb, T* S, T* Sinv, int safe, hipStream_t st);

// sweep.hip
// Right-hand sides of one launch: nr columns; column r of the input v is at
// v + r*ldv and of the FP64 output at out + r*ldo; the work vectors (g,
// scr*, outt: n per column; corr: 2k(P-1) per column) hold nr consecutive
// blocks.  gate (device, nr flags, may be null) skips the columns whose flag
// is 0.
struct Rhs {
    int nr = 1;
    int64_t ldv = 0, ldo = 0;
    const int* gate = nullptr;
    int nbuf = 0;   // LDS block buffers of the staged sweeps (0: as many as fit)
    // full sweeps only: when set, every partition gets two workgroups, the
    // bottom half of its rows from the LU band and the top half from the UL
    // band (both must hold factors of every partition); the top half's
    // stage-1 values go to this second scratch (same layout as scr)
    void* split_scr = nullptr;
};
// the staged sweeps' largest LDS block-buffer count for k (0: not staged)
template <class T>
int sweep_max_nbuf(int k);
// Full sweeps (L, D^{-1}, U' of every partition) of several systems in one
// launch: one job per (system, partition, right-hand side) in a device
// table.  The input column of a job is V + voff and its output Z + ooff,
// V and Z given at launch; gate[job.gate] switches a column off.
struct SweepJob {
    const void* B;        // the system's factor band (null: a padding job)
    const void* Bul;      // its UL band (last partition under LU_UL)
    const void* corr;     // reduced-system corrections of this column, or null
    void* scr;            // n-element work vector of this column
    int64_t voff, ooff;
    int64_t n;
    int P, k, part, gate, last_is_ul, ring_mask, nbuf, pad_;
    int64_t s0, len0;     // len0 > 0: band rows [s0, s0 + len0) instead of partition `part`
};
// Appends nr jobs (column r: input V + r*ldv, output Z + r*ldo, scratch
// scr + r*n, gate r) sweeping band rows [s0, s0 + len0) of B (half-bandwidth
// k, n rows) as one system; returns the sweep variant (jobs of one launch
// share it) and raises lds to what it needs.
template <class T>
int sweep_range_job_append(std::vector<SweepJob>& jobs, size_t& lds, const T* B, int k, int64_t s0, int64_t len0,
                           T* scr, int64_t n, int64_t ldv, int64_t ldo, int nr);
// Appends the jobs of one system (nr columns at voff + r*ldv / ooff +
// r*ldo, scr and corr holding nr consecutive blocks, gates gate0 + r) and
// raises lds to what they need.  Without mode_out: false when its width
// does not take the staged sweep (nothing appended); with mode_out: any
// width, *mode_out the sweep variant its jobs need (jobs of one launch
// share it).  False for k = 0.
template <class T>
bool sweep_jobs_append(std::vector<SweepJob>& jobs, size_t& lds, const T* B, const T* Bul, int k, PartInfo pi,
                       int last_is_ul, const T* corr, T* scr, int64_t voff, int64_t ldv, int64_t ooff, int64_t ldo,
                       int nr, int gate0, int nbuf_req, int* mode_out = nullptr);
template <class T>
void launch_sweep_full_jobs(const SweepJob* d_jobs, int njobs, size_t lds, const double* V, double* Z,
                            const int* gate, hipStream_t st, int mode = 2);
template <class T>
void launch_sweep_interface(const T* B, const T* Bul, int k, PartInfo pi, const double* v, T* g, T* scr0, T* scr1,
                            hipStream_t st, const Rhs& rh = Rhs());
template <class T>
void launch_interface_solve(const T* WV, const T* Sinv, const T* offB, const T* offC, int k, PartInfo pi,
                            const T* g, T* corr, hipStream_t st, const Rhs& rh = Rhs());
template <class T>
void launch_sweep_full(const T* B, const T* Bul, int k, PartInfo pi, int last_is_ul, const double* v, const T* corr,
                       double* outd, T* outt, T* scr, hipStream_t st, const Rhs& rh = Rhs());
template <class T>
void launch_diag_divide(const T* B, int64_t n, const double* v, double* out, hipStream_t st, const Rhs& rh = Rhs());
// zeros in A[0, na) and B[0, nb) (B may be null when nb = 0)
template <class T>
int64_t count_band_zeros(const T* A, int64_t na, const T* B, int64_t nb, unsigned long long* d_cnt, hipStream_t st);
// the same queued on st, the count copied to *h_out (pinned) when st reaches
// it; at most max_groups workgroups (fewer leave room for kernels running
// beside it)
template <class T>
void count_band_zeros_async(const T* A, int64_t na, const T* B, int64_t nb, unsigned long long* d_cnt,
                            unsigned long long* h_out, hipStream_t st, int max_groups);

}  // namespace sap_amd
