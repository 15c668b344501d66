/*
 * CPU oracle for the SaP solver -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-loop restatement of the reference algorithm, used to check the HIP
 * path (tests/, __graft_entry__.smoke()) and timed as the CPU baseline
 * (bench.py, cpu_baseline leg).  The shipped library never links or loads it.
 *
 * Every routine cites the reference code it restates.  Floating-point
 * summation orders follow the reference's per-thread loops where that is
 * cheap; results are compared to the GPU path within stated tolerances.
 *
 * Pinned by tests/test_oracle.py against: numpy/scipy dense and banded
 * solves (closed forms), a dense truncated-SPIKE preconditioner built from
 * its definition in numpy, the special cases P=1 (M = A) and Block
 * (M = blockdiag), textbook BiCGStab(l) convergence, MINRES on SPD systems
 * and the monitor state machine (exhaustive host tests).
 */
#include "sap_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* ------------------------------------------------------------------------ */
/* Workload generator: xoshiro256** (Blackman & Vigna 2018), one generator   */
/* per row, seeded through splitmix64 from (seed, row+1).                    */
/* ------------------------------------------------------------------------ */
inline uint64_t splitmix64(uint64_t& x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct Xoshiro256ss {
    uint64_t s[4];
    Xoshiro256ss(uint64_t seed, uint64_t stream)
    {
        uint64_t x = seed ^ (stream * 0xD1B54A32D192ED03ULL);
        for (int i = 0; i < 4; ++i) s[i] = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    inline uint64_t next()
    {
        const uint64_t result = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    /* uniform on [-1, 1): 53 random bits scaled by 2^-52, minus one (exact) */
    inline double uniform_pm1() { return (double)(next() >> 11) * 0x1.0p-52 - 1.0; }
};

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

/* Deterministic reductions: fixed chunking, chunks summed in order. */
constexpr int64_t RED_CHUNKS = 64;

double dot(const std::vector<double>& a, const std::vector<double>& b)
{
    const int64_t n = (int64_t)a.size();
    double part[RED_CHUNKS];
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < RED_CHUNKS; ++c) {
        int64_t lo = n * c / RED_CHUNKS, hi = n * (c + 1) / RED_CHUNKS;
        double s = 0.0;
        for (int64_t i = lo; i < hi; ++i) s += a[i] * b[i];
        part[c] = s;
    }
    double s = 0.0;
    for (int c = 0; c < RED_CHUNKS; ++c) s += part[c];
    return s;
}
double nrm2(const std::vector<double>& a) { return std::sqrt(dot(a, a)); }

/* y = alpha*x + beta*y written to z  (cusp::blas::axpby(x, y, z, alpha, beta)) */
void axpby(const std::vector<double>& x, const std::vector<double>& y, std::vector<double>& z,
           double alpha, double beta)
{
    const int64_t n = (int64_t)x.size();
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) z[i] = alpha * x[i] + beta * y[i];
}
/* y += alpha*x  (cusp::blas::axpy(x, y, alpha)) */
void axpy(const std::vector<double>& x, std::vector<double>& y, double alpha)
{
    const int64_t n = (int64_t)x.size();
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) y[i] += alpha * x[i];
}

/* ------------------------------------------------------------------------ */
/* Monitors: restatement of sap/monitor.h                                   */
/* ------------------------------------------------------------------------ */
struct BiCGStabLMonitor {          /* sap/monitor.h:153-335 */
    int    maxIterations;
    float  iterations = 0.f;
    int    maxmSteps;
    int    stag = 0, moreSteps = 0, maxStagSteps = 3;
    double relTol, absTol, rhsNorm = 0, rNorm = std::numeric_limits<double>::max();
    int    code = 0;
    std::string message;

    BiCGStabLMonitor(int maxIt, int maxm, double rt, double at)
        : maxIterations(maxIt), maxmSteps(maxm), relTol(rt), absTol(at) {}
    void   init(double rhs_norm) { rhsNorm = rhs_norm; iterations = 0; code = 0; message = ""; } /* :271-277 */
    double tolerance() const { return absTol + relTol * rhsNorm; }
    void   stop(int c, const char* m) { code = c; message = m; }
    void   increment(float x) { iterations += x; }
    void   incrementStag() { stag++; }
    void   resetStag() { stag = 0; }
    bool   finished()                               /* :177-184 */
    {
        if (code != 0) return true;
        if (iterations > maxIterations) stop(-1, "Maximum number of iterations was reached");
        return code != 0;
    }
    bool   finished(double r)                       /* :291-314 */
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
    bool   needCheckConvergence(double r)           /* :322-335 */
    {
        rNorm = r;
        if (code != 0) return false;
        if (std::isnan(rNorm)) stop(-2, "Residual norm is NaN");
        return code == 0 && (rNorm <= tolerance() || stag >= maxStagSteps || moreSteps);
    }
    bool   converged() const { return code > 0; }
};

struct Monitor {                   /* sap/monitor.h:25-147 */
    int    maxIterations;
    float  iterations = 0.f;
    double relTol, absTol, rhsNorm = 0, rNorm = std::numeric_limits<double>::max();
    int    code = 0;
    std::string message;
    Monitor(int maxIt, double rt, double at) : maxIterations(maxIt), relTol(rt), absTol(at) {}
    void   init(double rhs_norm) { rhsNorm = rhs_norm; iterations = 0; code = 0; message = ""; }
    double tolerance() const { return absTol + relTol * rhsNorm; }
    void   stop(int c, const char* m) { code = c; message = m; }
    size_t iteration_count() const { return (size_t)(iterations + 0.5f); }
    bool   finished(double r)                       /* :133-147 */
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

/* ------------------------------------------------------------------------ */
/* Pivot boosting: sap/device/factor_band_const.cuh:29-40 with               */
/* BURST_VALUE = 1e-7, BURST_NEW_VALUE = 1e-4 (sap/common.h:46-47).          */
/* ------------------------------------------------------------------------ */
template <class T>
inline T boost(T v)
{
    const T thr = (T)1e-7, nv = (T)1e-4;
    if (v > thr || v < -thr) return v;
    return v < 0 ? -nv : nv;
}

inline void part_bounds(int64_t n, int P, int p, int64_t& s, int64_t& len)
{
    /* sap/device/factor_band_const.cuh:1432-1438 */
    const int64_t q = n / P, r = n % P;
    if (p < r) { s = p * (q + 1); len = q + 1; }
    else       { s = p * q + r;   len = q; }
}

/* ------------------------------------------------------------------------ */
/* The preconditioner: restatement of sap::Precond for the constant-band     */
/* path (sap/precond.h:910-1362 setup, :1370-1528 apply).                    */
/* ------------------------------------------------------------------------ */
template <class T>
struct Precond {
    int  P_req, fact, ptype;
    bool safe;
    int64_t n = 0;
    int  k = 0, P = 1;
    std::vector<T> B;        /* column-major band: A(r,c) at c*(2k+1) + k + r - c   */
    std::vector<T> offB;     /* per boundary i: B_i(a,b) row-major (k x k)          */
    std::vector<T> offC;     /* per boundary i: C_{i+1}(a,b) row-major (k x k)      */
    std::vector<T> WV;       /* per boundary: V_i then W_{i+1}, each (a,r) at r*k+a */
    std::vector<T> R;        /* per boundary 2k x 2k, column-major                  */
    double t_lu = 0, t_ul = 0, t_asm = 0, t_full = 0;

    inline T& at(int64_t r, int64_t c) { return B[c * (2 * k + 1) + k + (r - c)]; }
    static inline T& at2(std::vector<T>& M, int k, int64_t r, int64_t c)
    {
        return M[c * (2 * k + 1) + k + (r - c)];
    }

    /* ---- no-pivot banded LU of rows/cols [s, s+len) in place.
     * Restates blockedBandLU_critical_const_phase1/2/3 (factor_band_const.cuh:
     * 1425-1566; K>=70, every pivot boosted) and bandLU_g32[_safe]
     * (:231-400; K<70, pivots boosted only with safeFactorization and never
     * the partition's last one).  Result: L (unit, multipliers below the
     * diagonal) and U (pivots on the diagonal, not yet divided). */
    static void band_lu(std::vector<T>& M, int k, int64_t s, int64_t len, bool do_boost, bool boost_last)
    {
        for (int64_t c = 0; c < len; ++c) {
            const int64_t col = s + c;
            T& piv = at2(M, k, col, col);
            if (do_boost && (c < len - 1 || boost_last)) piv = boost(piv);
            const int m = (int)std::min<int64_t>(k, len - 1 - c);
            if (m == 0) continue;
            const T p = piv;
            for (int t = 1; t <= m; ++t) at2(M, k, col + t, col) /= p;
            for (int cc = 1; cc <= m; ++cc) {
                const T u = at2(M, k, col, col + cc);
                T* colp = &at2(M, k, col, col + cc);     /* rows col.. of column col+cc */
                const T* lp = &at2(M, k, col, col);
                for (int t = 1; t <= m; ++t) colp[t] -= lp[t] * u;
            }
        }
    }

    /* ---- no-pivot banded UL (bottom-up) of [s, s+len).  Restates
     * blockedBandUL_critical_const_phase1/2/3 (factor_band_const.cuh:1570-1710)
     * and bandUL_g32[_safe].  Result: U unit upper (entries above the
     * diagonal divided by the pivot), L lower with the pivots, undivided. */
    static void band_ul(std::vector<T>& M, int k, int64_t s, int64_t len, bool do_boost, bool boost_last)
    {
        for (int64_t c = len - 1; c >= 0; --c) {
            const int64_t col = s + c;
            T& piv = at2(M, k, col, col);
            if (do_boost && (c > 0 || boost_last)) piv = boost(piv);
            const int m = (int)std::min<int64_t>(k, c);
            if (m == 0) continue;
            const T p = piv;
            for (int t = 1; t <= m; ++t) at2(M, k, col - t, col) /= p;
            for (int cc = 1; cc <= m; ++cc) {
                const T l = at2(M, k, col, col - cc);
                for (int t = 1; t <= m; ++t) at2(M, k, col - t, col - cc) -= at2(M, k, col - t, col) * l;
            }
        }
    }

    /* bandLU_post_divide (factor_band_const.cuh:1715-1736): U(r,c) /= U(r,r) */
    static void post_divide_lu(std::vector<T>& M, int k, int64_t s, int64_t len)
    {
        for (int64_t c = 0; c < len; ++c)
            for (int t = 1; t <= std::min<int64_t>(k, c); ++t)
                at2(M, k, s + c - t, s + c) /= at2(M, k, s + c - t, s + c - t);
    }
    /* bandLUUL_post_divide, last-partition branch (:1740-1760): L(r,c) /= L(r,r) */
    static void post_divide_ul(std::vector<T>& M, int k, int64_t s, int64_t len)
    {
        for (int64_t c = 0; c < len; ++c)
            for (int t = 1; t <= std::min<int64_t>(k, len - 1 - c); ++t)
                at2(M, k, s + c + t, s + c) /= at2(M, k, s + c + t, s + c + t);
    }

    /* hasZeroPivots (precond.h:6337-6350) with SmallerThan (:526-534) */
    bool has_zero_pivots(const std::vector<T>& M, int64_t s, int64_t len) const
    {
        for (int64_t i = s; i < s + len; ++i)
            if (std::abs(M[i * (2 * k + 1) + k]) < (T)1e-7) return true;
        return false;
    }

    void clear()
    {
        B.clear(); offB.clear(); offC.clear(); WV.clear(); R.clear();
    }

    /* returns 0 or -1 (Zero_pivoting) */
    int setup(int64_t n_, int k_, const double* band)
    {
        clear();
        n = n_; k = k_;
        /* convertToBandedMatrix (precond.h:2103-2146): P clamp and transpose */
        P = std::min(P_req, (int)std::max<int64_t>(1, n / std::max(k + 1, 2 * k)));
        if (P < 1) P = 1;
        const int w = 2 * k + 1;
        B.assign((size_t)w * n, T(0));
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < n; ++c)
            for (int d = 0; d < w; ++d) {
                const int64_t r = c + d - k;
                if (r < 0 || r >= n) continue;
                B[c * w + d] = (T)band[r * w + k + (c - r)];
            }
        if (k == 0 || ptype == 2) return 0;

        const bool do_boost = (k >= 70) || safe;   /* precond.h:4363-4397 */
        const bool boost_last = (k >= 70);

        if (ptype == 1 || P == 1) {               /* precond.h:1212-1235 */
            double t0 = now_s();
#pragma omp parallel for schedule(dynamic, 1)
            for (int p = 0; p < P; ++p) {
                int64_t s, len; part_bounds(n, P, p, s, len);
                band_lu(B, k, s, len, do_boost, boost_last);
            }
            if (!safe && has_zero_pivots(B, 0, n)) return -1;
#pragma omp parallel for schedule(dynamic, 1)
            for (int p = 0; p < P; ++p) {
                int64_t s, len; part_bounds(n, P, p, s, len);
                post_divide_lu(B, k, s, len);
            }
            t_lu = now_s() - t0;
            return 0;
        }

        /* extractOffDiagonal + copydWV (precond.h:2371-2400, data_transfer.cuh:127-199) */
        const size_t kk = (size_t)k * k;
        offB.assign(kk * (P - 1), T(0));
        offC.assign(kk * (P - 1), T(0));
        for (int i = 0; i < P - 1; ++i) {
            int64_t s1, l1; part_bounds(n, P, i + 1, s1, l1);
            for (int a = 0; a < k; ++a)
                for (int b = 0; b < k; ++b) {
                    offB[i * kk + a * k + b] = (b <= a) ? at(s1 - k + a, s1 + b) : T(0);
                    offC[i * kk + a * k + b] = (a <= b) ? at(s1 + a, s1 - k + b) : T(0);
                }
        }

        WV.assign(2 * kk * (P - 1), T(0));
        double t0 = now_s();
        if (fact == 1) {
            /* LU_only (precond.h:1269-1292).  Every partition LU factored; the
             * spike tips come from the LU factors: V_i from the bottom blocks,
             * W_{i+1} from full sweeps (calculateSpikes_const, :5787-5949).
             * The factors are post-divided before the sweeps, as the GPU
             * library does on every path: the reference leaves U undivided
             * here (bandLU_post_divide runs only for Block or P = 1,
             * precond.h:4408-4417) while its sweeps assume a unit U'
             * (:5152-5164). */
#pragma omp parallel for schedule(dynamic, 1)
            for (int p = 0; p < P; ++p) {
                int64_t s, len; part_bounds(n, P, p, s, len);
                band_lu(B, k, s, len, do_boost, boost_last);
            }
            if (!safe && has_zero_pivots(B, 0, n)) return -1;
#pragma omp parallel for schedule(dynamic, 1)
            for (int p = 0; p < P; ++p) {
                int64_t s, len; part_bounds(n, P, p, s, len);
                post_divide_lu(B, k, s, len);
            }
            t_lu = now_s() - t0;
            t0 = now_s();
#pragma omp parallel for schedule(dynamic, 1)
            for (int i = 0; i < P - 1; ++i) {
                int64_t s0, l0, s1, l1;
                part_bounds(n, P, i, s0, l0);
                part_bounds(n, P, i + 1, s1, l1);
                const int64_t e0 = s0 + l0;
                std::vector<T> y(k), full(l1);
                for (int r = 0; r < k; ++r) {
                    /* V_i(:, r): bottom K rows of A_i^{-1} [0; B_i(:, r)] */
                    for (int a = 0; a < k; ++a) y[a] = offB[i * kk + a * k + r];
                    for (int j = 0; j < k; ++j)
                        for (int t = j + 1; t < k; ++t) y[t] -= y[j] * at(e0 - k + t, e0 - k + j);
                    for (int j = 0; j < k; ++j) y[j] /= at(e0 - k + j, e0 - k + j);
                    for (int j = k - 1; j >= 0; --j)
                        for (int t = 0; t < j; ++t) y[t] -= y[j] * at(e0 - k + t, e0 - k + j);
                    for (int a = 0; a < k; ++a) WV[2 * kk * i + (size_t)r * k + a] = y[a];
                    /* W_{i+1}(:, r): top K rows of A_{i+1}^{-1} [C_{i+1}(:, r); 0] */
                    std::fill(full.begin(), full.end(), T(0));
                    for (int a = 0; a < k; ++a) full[a] = offC[i * kk + a * k + r];
                    sweep_lu_range(full.data(), s1, l1);
                    for (int a = 0; a < k; ++a) WV[2 * kk * i + kk + (size_t)r * k + a] = full[a];
                }
            }
            t_asm = now_s() - t0;
        } else {
            /* LU_UL (precond.h:1294-1347) */
            std::vector<T> B2 = B;
#pragma omp parallel for schedule(dynamic, 1)
            for (int p = 0; p < P - 1; ++p) {
                int64_t s, len; part_bounds(n, P, p, s, len);
                band_lu(B, k, s, len, do_boost, boost_last);
            }
            int64_t sl, ll; part_bounds(n, P, P - 1, sl, ll);
            if (!safe && has_zero_pivots(B, 0, sl)) return -1;
            t_lu = now_s() - t0;
            t0 = now_s();
            /* partBlockedBandedUL (precond.h:4782-4846): partitions 1..P-1 */
#pragma omp parallel for schedule(dynamic, 1)
            for (int p = 1; p < P; ++p) {
                int64_t s, len; part_bounds(n, P, p, s, len);
                band_ul(B2, k, s, len, do_boost, boost_last);
            }
            int64_t s1_, l1_; part_bounds(n, P, 1, s1_, l1_);
            if (!safe && has_zero_pivots(B2, s1_, n - s1_)) return -1;
            t_ul = now_s() - t0;
            t0 = now_s();
            /* calculateSpikes(B2, WV) (precond.h:6230-6287): short sweeps on the
             * K x K corner blocks.  forwardElimLdWV/backwardElimUdWV
             * (sweep_band_const.cuh:589-787). */
#pragma omp parallel for schedule(dynamic, 1)
            for (int i = 0; i < P - 1; ++i) {
                int64_t s0, l0, s1, l1;
                part_bounds(n, P, i, s0, l0);
                part_bounds(n, P, i + 1, s1, l1);
                const int64_t e0 = s0 + l0;
                std::vector<T> y(k);
                for (int r = 0; r < k; ++r) {
                    /* V_i(:, r) with the LU of partition i: unit L, then U with division */
                    for (int a = 0; a < k; ++a) y[a] = offB[i * kk + a * k + r];
                    for (int j = 0; j < k - 1; ++j)
                        for (int t = j + 1; t < k; ++t) y[t] -= y[j] * at(e0 - k + t, e0 - k + j);
                    for (int j = k - 1; j >= 0; --j) {
                        y[j] /= at(e0 - k + j, e0 - k + j);
                        for (int t = 0; t < j; ++t) y[t] -= y[j] * at(e0 - k + t, e0 - k + j);
                    }
                    for (int a = 0; a < k; ++a) WV[2 * kk * i + (size_t)r * k + a] = y[a];
                    /* W_{i+1}(:, r) with the UL of partition i+1: unit U bottom-up,
                     * then L (with pivots) top-down with division */
                    for (int a = 0; a < k; ++a) y[a] = offC[i * kk + a * k + r];
                    for (int j = k - 1; j >= 1; --j)
                        for (int t = 0; t < j; ++t) y[t] -= y[j] * at2(B2, k, s1 + t, s1 + j);
                    for (int j = 0; j < k; ++j) {
                        y[j] /= at2(B2, k, s1 + j, s1 + j);
                        for (int t = j + 1; t < k; ++t) y[t] -= y[j] * at2(B2, k, s1 + t, s1 + j);
                    }
                    for (int a = 0; a < k; ++a) WV[2 * kk * i + kk + (size_t)r * k + a] = y[a];
                }
            }
            /* copyLastPartition (precond.h:6326-6330) */
            {
                const int64_t lo = (int64_t)(2 * k + 1) * sl;
                std::copy(B2.begin() + lo, B2.end(), B.begin() + lo);
            }
            /* bandLUUL_post_divide (precond.h:1328-1339) */
#pragma omp parallel for schedule(dynamic, 1)
            for (int p = 0; p < P; ++p) {
                int64_t s, len; part_bounds(n, P, p, s, len);
                if (p < P - 1) post_divide_lu(B, k, s, len);
                else           post_divide_ul(B, k, s, len);
            }
            t_asm = now_s() - t0;
        }

        /* assembleReducedMat (data_transfer.cuh:14-80):
         * R_i = [[I, V_i], [W_{i+1}, I]] (2k x 2k, column-major) */
        t0 = now_s();
        const int k2 = 2 * k;
        R.assign((size_t)k2 * k2 * (P - 1), T(0));
#pragma omp parallel for schedule(dynamic, 1)
        for (int i = 0; i < P - 1; ++i) {
            T* Ri = &R[(size_t)k2 * k2 * i];
            auto Rat = [&](int r, int c) -> T& { return Ri[(size_t)c * k2 + r]; };
            const T* V = &WV[2 * kk * i];
            const T* W = &WV[2 * kk * i + kk];
            for (int a = 0; a < k; ++a) {
                Rat(a, a) = T(1);
                Rat(k + a, k + a) = T(1);
                for (int b = 0; b < k; ++b) {
                    Rat(a, k + b) = V[(size_t)b * k + a];
                    Rat(k + a, b) = W[(size_t)b * k + a];
                }
            }
            /* partBlockedFullLU_var (precond.h:2551-2611):
             * fullLU_sub_spec: S = I - W V (factor_band_var.cuh:734-748) */
            for (int b = 0; b < k; ++b)
                for (int a = 0; a < k; ++a) {
                    T tmp = Rat(k + a, k + b);
                    for (int c = 0; c < k; ++c) tmp -= Rat(k + a, c) * Rat(c, k + b);
                    Rat(k + a, k + b) = tmp;
                }
            /* blocked LU of S (factor_band_var.cuh:964-1092), boosting only with
             * safeFactorization */
            for (int j = k; j < k2; ++j) {
                if (safe) Rat(j, j) = boost(Rat(j, j));
                const T p = Rat(j, j);
                for (int t = j + 1; t < k2; ++t) Rat(t, j) /= p;
                for (int c = j + 1; c < k2; ++c) {
                    const T u = Rat(j, c);
                    for (int t = j + 1; t < k2; ++t) Rat(t, c) -= Rat(t, j) * u;
                }
            }
            /* fullLU_post_divide (factor_band_var.cuh:771-790) */
            for (int c = k + 1; c < k2; ++c)
                for (int r = k; r < c; ++r) Rat(r, c) /= Rat(r, r);
        }
        t_full = now_s() - t0;
        return 0;
    }

    /* L, D^{-1}, U' sweep on one LU-factored partition (forwardElimL +
     * D^{-1} transform + backwardElimU; sweep_band_const.cuh:154-584,
     * precond.h:5152-5155) applied to v[0..len) = x[s..s+len). */
    void sweep_lu_range(T* v, int64_t s, int64_t len)
    {
        for (int64_t c = 0; c < len; ++c) {
            const int m = (int)std::min<int64_t>(k, len - 1 - c);
            const T x = v[c];
            const T* l = &at(s + c, s + c);
            for (int t = 1; t <= m; ++t) v[c + t] -= x * l[t];
        }
        for (int64_t c = 0; c < len; ++c) v[c] /= at(s + c, s + c);
        for (int64_t c = len - 1; c >= 0; --c) {
            const int m = (int)std::min<int64_t>(k, c);
            const T x = v[c];
            const T* u = &at(s + c, s + c);
            for (int t = 1; t <= m; ++t) v[c - t] -= x * u[-t];
        }
    }
    /* the UL-factored last partition (forwardElimL_LU_UL / backwardElimU_LU_UL,
     * sweep_band_const.cuh:1004-1284): U' bottom-up, D^{-1}, L' top-down */
    void sweep_ul_range(T* v, int64_t s, int64_t len)
    {
        for (int64_t c = len - 1; c >= 0; --c) {
            const int m = (int)std::min<int64_t>(k, c);
            const T x = v[c];
            const T* u = &at(s + c, s + c);
            for (int t = 1; t <= m; ++t) v[c - t] -= x * u[-t];
        }
        for (int64_t c = 0; c < len; ++c) v[c] /= at(s + c, s + c);
        for (int64_t c = 0; c < len; ++c) {
            const int m = (int)std::min<int64_t>(k, len - 1 - c);
            const T x = v[c];
            const T* l = &at(s + c, s + c);
            for (int t = 1; t <= m; ++t) v[c + t] -= x * l[t];
        }
    }

    void sweeps(std::vector<T>& v)      /* partBandedFwdSweep + partBandedBckSweep */
    {
        const bool spike = (ptype == 0 && P > 1);
#pragma omp parallel for schedule(dynamic, 1)
        for (int p = 0; p < P; ++p) {
            int64_t s, len; part_bounds(n, P, p, s, len);
            if (spike && fact == 0 && p == P - 1) sweep_ul_range(&v[s], s, len);
            else                                  sweep_lu_range(&v[s], s, len);
        }
    }

    /* partFullFwdSweep / partFullBckSweep (precond.h:5418-5487,
     * sweep_band_var.cuh:21-191) on the 2k interface entries */
    void reduced_solve(std::vector<T>& g)
    {
        const int k2 = 2 * k;
#pragma omp parallel for schedule(dynamic, 1)
        for (int i = 0; i < P - 1; ++i) {
            int64_t s1, l1; part_bounds(n, P, i + 1, s1, l1);
            T* z = &g[s1 - k];
            const T* Ri = &R[(size_t)k2 * k2 * i];
            auto Rat = [&](int r, int c) -> T { return Ri[(size_t)c * k2 + r]; };
            for (int a = 0; a < k; ++a) {
                T tmp = z[k + a];
                for (int b = 0; b < k; ++b) tmp -= z[b] * Rat(k + a, b);
                z[k + a] = tmp;
            }
            for (int j = k; j < k2 - 1; ++j)
                for (int t = j + 1; t < k2; ++t) z[t] -= z[j] * Rat(t, j);
            for (int a = 0; a < k; ++a) z[k + a] /= Rat(k + a, k + a);
            for (int j = k2 - 1; j >= k; --j)
                for (int t = 0; t < j; ++t) z[t] -= z[j] * Rat(t, j);
        }
    }

    /* purifyRHS + innerProductBCX (precond.h:5494-5530, inner_product.cuh:8-293) */
    void purify(const std::vector<T>& x, std::vector<T>& sol)
    {
        const size_t kk = (size_t)k * k;
#pragma omp parallel for schedule(dynamic, 1)
        for (int i = 0; i < P - 1; ++i) {
            int64_t s1, l1; part_bounds(n, P, i + 1, s1, l1);
            for (int a = 0; a < k; ++a) {
                T s_b = 0, s_c = 0;
                for (int b = 0; b < k; ++b) {
                    s_b += offB[i * kk + (size_t)a * k + b] * x[s1 + b];
                    s_c += offC[i * kk + (size_t)a * k + b] * x[s1 - k + b];
                }
                sol[s1 - k + a] -= s_b;
                sol[s1 + a] -= s_c;
            }
        }
    }

    /* Precond::operator() + solve + getSRev (precond.h:1370-1528) */
    void apply(const double* vin, double* zout)
    {
        if (ptype == 2) { std::memcpy(zout, vin, sizeof(double) * n); return; }
        std::vector<T> rhs(n);
        for (int64_t i = 0; i < n; ++i) rhs[i] = (T)vin[i];
        if (k == 0) {
            for (int64_t i = 0; i < n; ++i) zout[i] = (double)(rhs[i] / B[i]);
            return;
        }
        std::vector<T> sol;
        if (ptype == 0 && P > 1) {
            sol = rhs;
            sweeps(rhs);
            reduced_solve(rhs);
            purify(rhs, sol);
        } else {
            sol.swap(rhs);
        }
        sweeps(sol);
        for (int64_t i = 0; i < n; ++i) zout[i] = (double)sol[i];
    }
};

/* ------------------------------------------------------------------------ */
struct Solver {
    orc_options opt;
    int P_req;
    Precond<float>  pf;
    Precond<double> pd;
    bool setup_done = false;
    int64_t n = 0;
    int k = 0;
    std::vector<double> band;     /* the FP64 row-major band used by the SpMV */
    BiCGStabLMonitor bmon;
    Monitor mmon;
    orc_stats st{};

    Solver(int P, const orc_options& o)
        : opt(o), P_req(P), bmon(o.max_iters, 8, o.rel_tol, o.abs_tol), mmon(o.max_iters, o.rel_tol, o.abs_tol)
    {
        pf.P_req = pd.P_req = P;
        pf.fact = pd.fact = o.fact_method;
        pf.ptype = pd.ptype = o.precond_type;
        pf.safe = pd.safe = o.safe_fact != 0;
    }

    void spmv(const std::vector<double>& x, std::vector<double>& y)
    {
        st.num_spmv++;
        orc_band_spmv(n, k, band.data(), x.data(), y.data());
    }
    void prec(const std::vector<double>& v, std::vector<double>& z)
    {
        st.num_applies++;
        if (opt.prec_double) pd.apply(v.data(), z.data());
        else                 pf.apply(v.data(), z.data());
    }

    /* sap::bicgstabl<..., L> (sap/bicgstab2.h:31-354) */
    template <int L>
    void bicgstabl(std::vector<double>& x, const std::vector<double>& b, BiCGStabLMonitor& monitor)
    {
        const double eps = 1e-20;
        double rho0 = 1, alpha = 0, omega = 1, rho1;
        std::vector<double> r0(n), r(n), u(n, 0.0), xx(n), Pv(n);
        std::vector<std::vector<double>> rr(L + 1, std::vector<double>(n, 0.0)), uu(L + 1, std::vector<double>(n, 0.0));
        double tao[L + 1][L + 1] = {};
        double gamma[L + 2] = {}, gamma_prime[L + 2] = {}, gamma_primeprime[L + 2] = {}, sigma[L + 2] = {};
        std::vector<double> Pxx(n), APxx(n);

        spmv(x, r0);
        axpby(b, r0, r0, 1.0, -1.0);
        r = r0;
        uu[0] = u; xx = x; rr[0] = r;
        double r_norm_min = nrm2(r);
        double r_norm = r_norm_min;
        double r_norm_act = r_norm;
        std::vector<double> x_min(n, 0.0);

        auto check_true = [&]() -> bool {
            prec(xx, Pxx);
            spmv(Pxx, APxx);
            axpby(b, APxx, rr[0], 1.0, -1.0);
            r_norm_act = nrm2(rr[0]);
            return monitor.finished(r_norm_act);
        };

        while (true) {
            rho0 = -omega * rho0;
            monitor.increment(0.25f);
            for (int j = 0; j < L; ++j) {
                rho1 = dot(rr[j], r0);
                if (rho0 == 0) { monitor.stop(-10, "rho0 is zero"); break; }
                double beta = alpha * rho1 / rho0;
                rho0 = rho1;
                for (int i = 0; i <= j; ++i) axpby(rr[i], uu[i], uu[i], 1.0, -beta);
                prec(uu[j], Pv);
                spmv(Pv, uu[j + 1]);
                double gam = dot(uu[j + 1], r0);
                if (gam == 0) { monitor.stop(-11, "gamma is zero"); break; }
                alpha = rho0 / gam;
                for (int i = 0; i <= j; ++i) axpy(uu[i + 1], rr[i], -alpha);
                r_norm_act = r_norm = nrm2(rr[0]);
                prec(rr[j], Pv);
                spmv(Pv, rr[j + 1]);
                if (std::fabs(alpha) * nrm2(uu[0]) < eps * nrm2(xx)) monitor.incrementStag();
                else monitor.resetStag();
                axpy(uu[0], xx, alpha);
                if (monitor.needCheckConvergence(r_norm)) {
                    if (check_true()) break;
                }
                if (r_norm_act < r_norm_min) { r_norm_min = r_norm_act; x_min = xx; }
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
                gamma_primeprime[j] = gamma[j + 1];
                for (int i = j + 1; i < L; ++i) gamma_primeprime[j] += tao[j][i] * gamma[i + 1];
            }
            if (std::fabs(gamma[1]) * nrm2(rr[0]) < eps * nrm2(xx)) monitor.incrementStag();
            else monitor.resetStag();

            axpy(rr[0], xx, gamma[1]);
            axpy(rr[L], rr[0], -gamma_prime[L]);
            axpy(uu[L], uu[0], -gamma[L]);
            r_norm_act = r_norm = nrm2(rr[0]);
            monitor.increment(0.25f);
            if (monitor.needCheckConvergence(r_norm)) {
                if (check_true()) break;
            }
            if (r_norm_act < r_norm_min) { r_norm_min = r_norm_act; x_min = xx; }
            if (monitor.finished()) break;
            monitor.increment(0.25f);

            bool brk = false;
            for (int j = 1; j < L; ++j) {
                axpy(uu[j], uu[0], -gamma[j]);
                if (std::fabs(gamma_primeprime[j]) * nrm2(rr[j]) < eps * nrm2(xx)) monitor.incrementStag();
                else monitor.resetStag();
                axpy(rr[j], xx, gamma_primeprime[j]);
                axpy(rr[j], rr[0], -gamma_prime[j]);
                r_norm_act = r_norm = nrm2(rr[0]);
                if (monitor.needCheckConvergence(r_norm)) {
                    if (check_true()) { brk = true; break; }
                }
                if (r_norm_act < r_norm_min) { r_norm_min = r_norm_act; x_min = xx; }
                if (monitor.finished()) { brk = true; break; }
            }
            if (brk || monitor.finished()) break;
            u = uu[0]; x = xx; r = rr[0];
            monitor.increment(0.25f);
        }

        if (monitor.converged()) {
            prec(xx, x);
        } else {
            std::vector<double> Pxmin(n), APxmin(n), rc(n), rcm(n);
            prec(xx, Pxx);
            spmv(Pxx, APxx);
            axpby(b, APxx, rc, 1.0, -1.0);
            prec(x_min, Pxmin);
            spmv(Pxmin, APxmin);
            axpby(b, APxmin, rcm, 1.0, -1.0);
            double rcn = nrm2(rc), rcmn = nrm2(rcm);
            if (rcn < rcmn) { x = Pxx; monitor.rNorm = rcn; }
            else            { x = Pxmin; monitor.rNorm = rcmn; }
        }
    }

    /* sap::minres (sap/minres.h:38-181) */
    void minres(std::vector<double>& x, const std::vector<double>& b, Monitor& monitor)
    {
        const double eps = std::numeric_limits<double>::epsilon();
        std::vector<double> y(n), r1(n);
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
        std::vector<double> v(n), w(n, 0.0), w1(n, 0.0), w2(n, 0.0), r2 = r1;
        while (!monitor.finished(phibar)) {
            const double s = 1 / beta;
            for (int64_t i = 0; i < n; ++i) v[i] = s * y[i];
            spmv(v, y);
            if (monitor.iteration_count() > 0) axpby(y, r1, y, 1.0, -beta / oldb);
            alpha = dot(v, y);
            axpby(y, r2, y, 1.0, -alpha / beta);
            r1 = r2;
            r2 = y;
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
            w1 = w2;
            w2 = w;
            for (int64_t i = 0; i < n; ++i) w[i] = denom * v[i] + (-oldeps * denom) * w1[i] + (-delta * denom) * w2[i];
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

    /* Inner Krylov solve of A d = rhs from d = 0 with relative tolerance rtol;
     * returns the monitor code and adds the iterations to *its. */
    int krylov(const std::vector<double>& rhs, std::vector<double>& d, double rtol, float* its, double* rnorm)
    {
        std::fill(d.begin(), d.end(), 0.0);
        if (opt.solver == 2) {
            mmon = Monitor(opt.max_iters, rtol, 0.0);
            mmon.init(nrm2(rhs));
            minres(d, rhs, mmon);
            *its += mmon.iterations; *rnorm = mmon.rNorm;
            return mmon.code;
        }
        bmon = BiCGStabLMonitor(opt.max_iters, 8, rtol, 0.0);
        bmon.init(nrm2(rhs));
        if (opt.solver == 1) bicgstabl<1>(d, rhs, bmon);
        else                 bicgstabl<2>(d, rhs, bmon);
        *its += bmon.iterations; *rnorm = bmon.rNorm;
        return bmon.code;
    }

    /* Outer FP64 iterative refinement around the Krylov solve: x += d with
     * A d = b - A x solved by the preconditioned Krylov method to refine_tol.
     * Needed when the preconditioner runs in FP32: the right-preconditioned
     * BiCGStab(l) of sap/bicgstab2.h returns x = M^{-1} xx, whose FP32
     * rounding bounds the attainable relative residual near 1e-7. */
    int refine_solve(const std::vector<double>& b, std::vector<double>& x)
    {
        std::vector<double> r(n), d(n), Ax(n);
        const double bnorm = nrm2(b);
        const double tol = opt.abs_tol + opt.rel_tol * bnorm;
        float its = 0.f;
        int code = -1;
        double rn = 0;
        for (int outer = 0; ; ++outer) {
            spmv(x, Ax);
            axpby(b, Ax, r, 1.0, -1.0);
            rn = nrm2(r);
            if (std::isnan(rn)) { code = -2; break; }
            if (rn <= tol) { code = 1; break; }
            if (outer >= opt.refine) { code = -1; break; }
            double rt = std::max(opt.refine_tol, 0.5 * tol / rn);
            double inner_r;
            int c = krylov(r, d, rt, &its, &inner_r);
            st.num_refine++;
            axpy(d, x, 1.0);
            if (c == -2) { code = -2; break; }
        }
        st.rhs_norm = bnorm;
        st.residual_norm = rn;
        st.num_iterations = its;
        bmon.code = mmon.code = code;
        bmon.message = mmon.message = (code == 1 ? "Converged" : code == -2 ? "Residual norm is NaN"
                                                 : "Maximum number of refinement steps was reached");
        return code;
    }

    int setup(int64_t n_, int k_, const double* b)
    {
        double t0 = now_s();
        n = n_; k = k_;
        band.assign(b, b + (size_t)n * (2 * k + 1));
        int rc = opt.prec_double ? pd.setup(n, k, b) : pf.setup(n, k, b);
        st = orc_stats{};
        st.time_setup = now_s() - t0;
        st.num_partitions = opt.prec_double ? pd.P : pf.P;
        st.bandwidth = k;
        st.time_bandLU = opt.prec_double ? pd.t_lu : pf.t_lu;
        st.time_bandUL = opt.prec_double ? pd.t_ul : pf.t_ul;
        st.time_assembly = opt.prec_double ? pd.t_asm : pf.t_asm;
        st.time_fullLU = opt.prec_double ? pd.t_full : pf.t_full;
        setup_done = (rc == 0);
        return rc;
    }

    int solve(const double* b, double* xio)
    {
        if (!setup_done) return -4;
        std::vector<double> bv(b, b + n), xv(xio, xio + n);
        st.num_applies = 0; st.num_spmv = 0; st.num_refine = 0;
        double t0 = now_s();
        int code;
        if (opt.refine > 0) {
            code = refine_solve(bv, xv);
        } else if (opt.solver == 2) {
            mmon = Monitor(opt.max_iters, opt.rel_tol, opt.abs_tol);
            mmon.init(nrm2(bv));
            minres(xv, bv, mmon);
            code = mmon.code;
            st.rhs_norm = mmon.rhsNorm; st.residual_norm = mmon.rNorm;
            st.num_iterations = mmon.iterations;
        } else {
            bmon = BiCGStabLMonitor(opt.max_iters, 8, opt.rel_tol, opt.abs_tol);
            bmon.init(nrm2(bv));
            if (opt.solver == 1) bicgstabl<1>(xv, bv, bmon);
            else                 bicgstabl<2>(xv, bv, bmon);
            code = bmon.code;
            st.rhs_norm = bmon.rhsNorm; st.residual_norm = bmon.rNorm;
            st.num_iterations = bmon.iterations;
        }
        if (opt.refine > 0) { /* stats already filled by refine_solve */ }
        st.time_solve = now_s() - t0;
        st.code = code;
        st.rel_residual_norm = st.residual_norm / st.rhs_norm;
        std::memcpy(xio, xv.data(), sizeof(double) * n);
        return code > 0 ? 1 : 0;
    }
};

} // namespace

/* ======================================================================== */
extern "C" {

void orc_gen_banded(int64_t n, int k, double dominance, uint64_t seed, double* band)
{
    const int w = 2 * k + 1;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        Xoshiro256ss rng(seed, (uint64_t)i + 1);
        double* row = band + i * w;
        double s = 0.0;
        for (int d = 0; d < w; ++d) {
            if (d == k) continue;
            const int64_t j = i + d - k;
            if (j < 0 || j >= n) { row[d] = 0.0; continue; }
            const double v = rng.uniform_pm1();
            row[d] = v;
            s += std::fabs(v);
        }
        row[k] = dominance * s;
    }
}

void orc_gen_vector(int64_t n, uint64_t seed, double* x)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        Xoshiro256ss rng(seed, (uint64_t)i + 1);
        x[i] = rng.uniform_pm1();
    }
}

/* MVBanded / bandedMatrixVecMul (sap/spmv.h:27-74, matrix_multiply.cuh:301-350):
 * y_i = sum_j A(i,j) x_j, columns in increasing order. */
void orc_band_spmv(int64_t n, int k, const double* band, const double* x, double* y)
{
    const int w = 2 * k + 1;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const double* row = band + i * w;
        double s = 0.0;
        const int dlo = (int)std::max<int64_t>(0, k - i);
        const int dhi = (int)std::min<int64_t>(w - 1, n - 1 - i + k);
        for (int d = dlo; d <= dhi; ++d) s += row[d] * x[i + d - k];
        y[i] = s;
    }
}

/* FP64 band LU with partial pivoting, column storage of height 3k+1
 * (rows j-2k..j+k of column j), as in LAPACK dgbtrf/dgbtrs. */
int orc_band_solve_pp(int64_t n, int k, const double* band, const double* b, double* x)
{
    const int h = 3 * k + 1;       /* rows r with j-2k <= r <= j+k */
    std::vector<double> a((size_t)h * n, 0.0);
    auto A = [&](int64_t r, int64_t c) -> double& { return a[(size_t)c * h + (r - c + 2 * k)]; };
    const int w = 2 * k + 1;
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < w; ++d) {
            const int64_t j = i + d - k;
            if (j >= 0 && j < n) A(i, j) = band[i * w + d];
        }
    std::vector<int64_t> piv(n);
    for (int64_t j = 0; j < n; ++j) {
        const int64_t rmax = std::min<int64_t>(n - 1, j + k);
        int64_t p = j;
        double best = std::fabs(A(j, j));
        for (int64_t r = j + 1; r <= rmax; ++r)
            if (std::fabs(A(r, j)) > best) { best = std::fabs(A(r, j)); p = r; }
        piv[j] = p;
        if (best == 0.0) return -1;
        const int64_t cmax = std::min<int64_t>(n - 1, j + 2 * k);
        if (p != j)
            for (int64_t c = j; c <= cmax; ++c) std::swap(A(j, c), A(p, c));
        const double d = A(j, j);
        for (int64_t r = j + 1; r <= rmax; ++r) A(r, j) /= d;
        for (int64_t c = j + 1; c <= cmax; ++c) {
            const double u = A(j, c);
            if (u == 0.0) continue;
            for (int64_t r = j + 1; r <= rmax; ++r) A(r, c) -= A(r, j) * u;
        }
    }
    std::vector<double> y(b, b + n);
    for (int64_t j = 0; j < n; ++j) {
        if (piv[j] != j) std::swap(y[j], y[piv[j]]);
        const int64_t rmax = std::min<int64_t>(n - 1, j + k);
        for (int64_t r = j + 1; r <= rmax; ++r) y[r] -= A(r, j) * y[j];
    }
    for (int64_t j = n - 1; j >= 0; --j) {
        y[j] /= A(j, j);
        const int64_t rmin = std::max<int64_t>(0, j - 2 * k);
        for (int64_t r = rmin; r < j; ++r) y[r] -= A(r, j) * y[j];
    }
    std::memcpy(x, y.data(), sizeof(double) * n);
    return 0;
}

void* orc_create(int num_partitions, const orc_options* opts)
{
    return new Solver(num_partitions, *opts);
}
void orc_destroy(void* s) { delete static_cast<Solver*>(s); }
int orc_setup_banded(void* s, int64_t n, int k, const double* band)
{
    return static_cast<Solver*>(s)->setup(n, k, band);
}
int orc_solve(void* s, const double* b, double* x) { return static_cast<Solver*>(s)->solve(b, x); }
void orc_get_stats(void* s, orc_stats* st) { *st = static_cast<Solver*>(s)->st; }
const char* orc_monitor_message(void* s)
{
    Solver* S = static_cast<Solver*>(s);
    return S->opt.solver == 2 ? S->mmon.message.c_str() : S->bmon.message.c_str();
}
int orc_precond_apply(void* s, const double* v, double* z)
{
    Solver* S = static_cast<Solver*>(s);
    if (!S->setup_done) return -4;
    if (S->opt.prec_double) S->pd.apply(v, z);
    else                    S->pf.apply(v, z);
    return 0;
}
int orc_num_partitions(void* s)
{
    Solver* S = static_cast<Solver*>(s);
    return S->opt.prec_double ? S->pd.P : S->pf.P;
}
int orc_get_factors(void* s, double* out)
{
    Solver* S = static_cast<Solver*>(s);
    if (S->opt.prec_double) std::copy(S->pd.B.begin(), S->pd.B.end(), out);
    else                    std::copy(S->pf.B.begin(), S->pf.B.end(), out);
    return 0;
}
int orc_get_reduced(void* s, double* R_out, double* WV_out)
{
    Solver* S = static_cast<Solver*>(s);
    if (S->opt.prec_double) {
        std::copy(S->pd.R.begin(), S->pd.R.end(), R_out);
        std::copy(S->pd.WV.begin(), S->pd.WV.end(), WV_out);
    } else {
        std::copy(S->pf.R.begin(), S->pf.R.end(), R_out);
        std::copy(S->pf.WV.begin(), S->pf.WV.end(), WV_out);
    }
    return 0;
}

} // extern "C"
