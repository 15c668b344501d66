// Setup kernels of the SaP preconditioner: banded LU/UL of the partitions,
// coupling-block extraction, spike tips and the reduced (Schur) blocks.
//
// Band layout (column-major, as the reference's m_B, sap/device/
// data_transfer.cuh:440-447): A(r, c) at B[c*(2k+1) + k + r - c].
// A UL factorization of a partition is computed as the LU factorization of
// the same storage read backwards: with X = (s+len)*(2k+1) - 1, the reversed
// block's element (r, c) lives at X - (c*(2k+1) + k + r - c).  So one kernel
// serves both directions.
#include "mfma.hpp"
#include "sap_internal.hpp"

namespace sap_amd {

namespace {

// ---------------------------------------------------------------------------
// Blocked right-looking banded LU of one partition per workgroup.
// Panel of NB columns staged in LDS together with the NB x k block of U rows
// to its right; the panel is factored in LDS, written back with the U rows
// divided by their pivots (so the stored factors are L, D, U' with unit L and
// U'), and the trailing k x k window is updated in global memory (L2) with
// the rank-NB product from LDS.
// ---------------------------------------------------------------------------
template <class T, int NB>
__global__ __launch_bounds__(256) void band_factor_kernel(T* __restrict__ B0, T* __restrict__ B1, int k_all,
                                                          const FactorJob* __restrict__ jobs)
{
    const FactorJob job = jobs[blockIdx.x];
    const int k = job.kb > 0 ? job.kb : k_all;
    T* __restrict__ base = (job.buf ? B1 : B0) + job.boff;
    const int64_t w = 2 * k + 1;
    const int64_t len = job.len;
    const int dir = job.dir;
    const int64_t origin = dir > 0 ? job.start * w : (job.start + len) * w - 1;
    auto aidx = [&](int64_t r, int64_t c) -> int64_t { return origin + dir * (c * w + k + r - c); };

    extern __shared__ unsigned char smem_raw[];
    T* P = reinterpret_cast<T*>(smem_raw);   // [(NB + k) rows][PS]
    constexpr int PS = NB + 1;
    T* Ub = P + (size_t)(NB + k) * PS;       // [NB rows][k cols]
    __shared__ T s_piv;

    for (int64_t j0 = 0; j0 < len; j0 += NB) {
        const int nb = (int)min((int64_t)NB, len - j0);
        const int H = (int)min(len - j0, (int64_t)nb + k);
        const int M = (int)min(len - j0 - nb, (int64_t)k);

        for (int e = threadIdx.x; e < H * nb; e += blockDim.x) {
            const int c = e / H, r = e % H;
            T v = (T)0;
            if (r - c <= k && c - r <= k) v = base[aidx(j0 + r, j0 + c)];
            P[r * PS + c] = v;
        }
        for (int e = threadIdx.x; e < nb * M; e += blockDim.x) {
            const int c = e / nb, r = e % nb;
            T v = (T)0;
            if (nb + c - r <= k) v = base[aidx(j0 + r, j0 + nb + c)];
            Ub[r * k + c] = v;
        }
        __syncthreads();

        for (int jj = 0; jj < nb; ++jj) {
            if (threadIdx.x == 0) {
                T p = P[jj * PS + jj];
                if (job.boost && (j0 + jj < len - 1 || job.boost_last)) p = boost_pivot(p);
                P[jj * PS + jj] = p;
                s_piv = p;
            }
            __syncthreads();
            const T piv = s_piv;
            const int rmax = min(H - 1, jj + k);
            for (int r = jj + 1 + (int)threadIdx.x; r <= rmax; r += blockDim.x) P[r * PS + jj] /= piv;
            __syncthreads();
            const int ncol = nb - 1 - jj;
            const int nrow = rmax - jj;
            for (int e = threadIdx.x; e < ncol * nrow; e += blockDim.x) {
                const int c = jj + 1 + e / nrow, r = jj + 1 + e % nrow;
                P[r * PS + c] -= P[r * PS + jj] * P[jj * PS + c];
            }
            const int nrow_u = nb - 1 - jj;
            for (int e = threadIdx.x; e < nrow_u * M; e += blockDim.x) {
                const int c = e / nrow_u, r = jj + 1 + e % nrow_u;
                Ub[r * k + c] -= P[r * PS + jj] * Ub[jj * k + c];
            }
            __syncthreads();
        }

        // write back the factored panel (U entries divided by their pivot)
        for (int e = threadIdx.x; e < H * nb; e += blockDim.x) {
            const int c = e / H, r = e % H;
            if (r - c <= k && c - r <= k) {
                T v = P[r * PS + c];
                if (r < c) v /= P[r * PS + r];
                base[aidx(j0 + r, j0 + c)] = v;
            }
        }
        for (int e = threadIdx.x; e < nb * M; e += blockDim.x) {
            const int c = e / nb, r = e % nb;
            if (nb + c - r <= k) base[aidx(j0 + r, j0 + nb + c)] = Ub[r * k + c] / P[r * PS + r];
        }

        // trailing update of the M x M window (band entries only)
        const int cq_count = (M + 3) / 4;
        for (int e = threadIdx.x; e < M * cq_count; e += blockDim.x) {
            const int i = e % M;
            const int c = 4 * (e / M);
            T acc0 = 0, acc1 = 0, acc2 = 0, acc3 = 0;
            const T* prow = P + (size_t)(nb + i) * PS;
            for (int jj = 0; jj < nb; ++jj) {
                const T l = prow[jj];
                const T* u = Ub + jj * k + c;
                acc0 += l * u[0];
                if (c + 1 < M) acc1 += l * u[1];
                if (c + 2 < M) acc2 += l * u[2];
                if (c + 3 < M) acc3 += l * u[3];
            }
            const T accs[4] = {acc0, acc1, acc2, acc3};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int cc = c + q;
                if (cc < M && i - cc <= k && cc - i <= k) base[aidx(j0 + nb + i, j0 + nb + cc)] -= accs[q];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// MFMA variant of the blocked banded LU (32-column panels), for the bands the
// register-resident kernel does not take, up to k = 480 (256 or 512
// threads, one panel row each).  Per panel:
//   1. every thread holds one row of the (32 + k) x 32 panel in
//      registers and the panel is factored column by column (boosting at
//      pivot time); the pivot row is broadcast through LDS, one barrier per
//      column;
//   2. each thread also holds one column of the 32 x k block of U rows to
//      the right of the panel and applies the panel's row operations to it,
//      one column step behind (multipliers shared through LDS);
//   3. the k x k trailing window in global memory (L2) receives the rank-32
//      update -L_below * U_right on the matrix cores
//      (v_mfma_f32_32x32x2f32, or v_mfma_f64_16x16x4f64 in FP64), tiles
//      held transposed so that their loads and stores run along columns,
//      four tiles per wavefront in flight; -L_below comes from LDS or, for
//      the widest FP64 bands, from the band itself.
// Factors are written as L (unit), D, U' (divided), like the LDS variant.
// ---------------------------------------------------------------------------
template <class T>
struct Mfma;

template <>
struct Mfma<float> {
    static constexpr int TILE = 32, KSTEP = 2, NACC = 16;
    typedef float acc_t __attribute__((ext_vector_type(16)));
    __device__ static inline acc_t mma(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
    }
    // accumulator register r of lane l holds D[row(l, r)][col(l)]
    __device__ static inline int row(int l, int r) { return (r & 3) + 4 * (l >> 5) + 8 * (r >> 2); }
    __device__ static inline int col(int l) { return l & 31; }
    // A (TILE x KSTEP): lane l holds A[l % 32][l / 32]; B (KSTEP x TILE): B[l / 32][l % 32]
    __device__ static inline int ab_mn(int l) { return l & 31; }
    __device__ static inline int ab_k(int l) { return l >> 5; }
};

template <>
struct Mfma<double> {
    static constexpr int TILE = 16, KSTEP = 4, NACC = 4;
    typedef double acc_t __attribute__((ext_vector_type(4)));
    __device__ static inline acc_t mma(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // FP64 accumulators interleave rows: register r of lane l is row l/16 + 4r
    __device__ static inline int row(int l, int r) { return (l >> 4) + 4 * r; }
    __device__ static inline int col(int l) { return l & 15; }
    __device__ static inline int ab_mn(int l) { return l & 15; }
    __device__ static inline int ab_k(int l) { return l >> 4; }
};

constexpr int MNB = 32;   // panel width of the MFMA factorization

// lg: the -L_below operand is read from the band (written one step
// earlier by the same workgroup) instead of being staged in LDS, which
// leaves the LDS to the U block alone (the wide bands).
template <class T>
struct MfmaLds {
    int kp, ls_stride, us_stride;
    size_t ls_off, us_off, ld_off, prow_off, piv_off, bytes;
    __host__ __device__ MfmaLds(int k, bool lg = false)
    {
        constexpr int TILE = sizeof(T) == 4 ? 32 : 16;
        kp = ((k + TILE - 1) / TILE) * TILE;
        ls_stride = MNB + 1;
        us_stride = kp + 4;
        ls_off = 0;
        us_off = ls_off + (lg ? 0 : (size_t)kp * ls_stride);
        ld_off = us_off + (size_t)MNB * us_stride;
        prow_off = ld_off + (size_t)2 * MNB;
        piv_off = prow_off + 2 * MNB;
        bytes = (piv_off + MNB) * sizeof(T);
    }
};

// NT threads (256 or 512), one panel row and one U-block column each, so
// k + 32 <= NT.  LG: -L_below read from the band in the trailing update
// (MfmaLds lg).
template <class T, int NT, bool LG>
__global__ __launch_bounds__(NT) void band_factor_mfma_kernel(T* __restrict__ B0, T* __restrict__ B1, int k_all,
                                                               const FactorJob* __restrict__ jobs)
{
    using MF = Mfma<T>;
    constexpr int TILE = MF::TILE;
    const FactorJob job = jobs[blockIdx.x];
    const int k = job.kb > 0 ? job.kb : k_all;
    T* __restrict__ base = (job.buf ? B1 : B0) + job.boff;
    const int64_t w = 2 * k + 1;
    const int64_t len = job.len;
    const int dir = job.dir;
    const int64_t origin = dir > 0 ? job.start * w : (job.start + len) * w - 1;
    auto aidx = [&](int64_t r, int64_t c) -> int64_t { return origin + dir * (c * w + k + r - c); };
    const int64_t cstep = (int64_t)dir * (w - 1);   // next column, same row

    extern __shared__ unsigned char smem_raw[];
    T* sm = reinterpret_cast<T*>(smem_raw);
    const MfmaLds<T> lay(k, LG);
    T* Ls = sm + lay.ls_off;       // -L_below [kp][MNB+1]
    T* Us = sm + lay.us_off;       // U_right [MNB][kp+4]
    T* lcol = sm + lay.ld_off;     // multipliers of one panel column for rows < MNB, [2][MNB]
    T* prow = sm + lay.prow_off;   // pivot row, double-buffered [2][MNB]
    T* pivs = sm + lay.piv_off;    // pivots of the panel [MNB]
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    constexpr int RPT = 1, NW = NT / 64;

    for (int64_t j0 = 0; j0 < len; j0 += MNB) {
        const int nb = (int)min((int64_t)MNB, len - j0);
        const int H = (int)min(len - j0, (int64_t)nb + k);
        const int Mw = (int)min(len - j0 - nb, (int64_t)k);

        // (1) thread tid owns panel row r = tid (32 registers) and U-block
        //     column c = tid (rows j0..j0+31, 32 registers)
        T row[RPT][MNB];
        T ucol[RPT][MNB];
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int r = tid + NT * q;
            const T* rp = base + aidx(j0 + r, j0);
#pragma unroll
            for (int c = 0; c < MNB; ++c)
                row[q][c] = (r < H && c < nb && r - c <= k && c - r <= k) ? rp[c * cstep] : (T)0;
            const int c = r;
            const T* cp = base + aidx(j0, j0 + nb + c);
#pragma unroll
            for (int m = 0; m < MNB; ++m) ucol[q][m] = (c < Mw && m < nb && nb + c - m <= k) ? cp[dir * m] : (T)0;
        }

        // (2) panel LU, one barrier per column.  The U block receives the row
        //     operations of column jj one step later (multipliers via LDS).
#pragma unroll
        for (int jj = 0; jj < MNB; ++jj) {
            if (jj < nb) {
                T* pr = prow + (jj & 1) * MNB;
                if (tid == jj) {
                    T p = row[0][jj];
                    if (job.boost && (j0 + jj < len - 1 || job.boost_last)) p = boost_pivot(p);
                    row[0][jj] = p;
                    pivs[jj] = p;
#pragma unroll
                    for (int c = jj; c < MNB; ++c) pr[c] = row[0][c];
                }
                __syncthreads();
                const T piv = pr[jj];
                T* lc = lcol + (jj & 1) * MNB;
#pragma unroll
                for (int q = 0; q < RPT; ++q) {
                    const int r = tid + NT * q;
                    T l = (T)0;
                    if (r > jj && r < H && r - jj <= k) {
                        l = row[q][jj] / piv;
                        row[q][jj] = l;
#pragma unroll
                        for (int c = jj + 1; c < MNB; ++c) row[q][c] -= l * pr[c];
                    }
                    if (r < MNB) lc[r] = l;   // zero outside the band
                }
                if (jj >= 1) {
                    const T* lp = lcol + ((jj - 1) & 1) * MNB;
#pragma unroll
                    for (int q = 0; q < RPT; ++q) {
                        const T u = ucol[q][jj - 1];
#pragma unroll
                        for (int r = jj; r < MNB; ++r) ucol[q][r] -= lp[r] * u;
                    }
                }
            }
        }
        __syncthreads();

        // (3) write back: panel (U entries divided by the pivot), U block
        //     (divided); stage -L_below and the undivided U block in LDS
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int r = tid + NT * q;
            if (r < H) {
                const T pr_r = r < nb ? pivs[r] : (T)1;
                T* rp = base + aidx(j0 + r, j0);
#pragma unroll
                for (int c = 0; c < MNB; ++c) {
                    if (c < nb && r - c <= k && c - r <= k) {
                        T v = row[q][c];
                        if (r < c) v /= pr_r;
                        rp[c * cstep] = v;
                    }
                }
            }
            if (!LG && r >= nb && r - nb < lay.kp) {
#pragma unroll
                for (int c = 0; c < MNB; ++c) Ls[(r - nb) * lay.ls_stride + c] = (r < H) ? -row[q][c] : (T)0;
            }
            const int c = r;
            if (c < lay.kp) {
                T* cp = base + aidx(j0, j0 + nb + c);
#pragma unroll
                for (int m = 0; m < MNB; ++m) {
                    Us[m * lay.us_stride + c] = ucol[q][m];
                    if (c < Mw && m < nb && nb + c - m <= k) cp[dir * m] = ucol[q][m] / pivs[m];
                }
            }
        }
        // Ls rows not covered above (panel rows beyond the threads) are padding
        if (!LG) {
            for (int r = NT - nb + tid; r < lay.kp; r += NT) {
#pragma unroll
                for (int c = 0; c < MNB; ++c) Ls[r * lay.ls_stride + c] = (T)0;
            }
        }
        __syncthreads();

        // (4) trailing window -= L_below * U_right on the matrix cores, TG
        //     tiles per wavefront at a time (8 KB of FP64 or FP32 window): the
        //     window (and, LG, the
        //     -L_below) loads of the group are issued before its products, so
        //     their latencies overlap; each tile's products run in the same
        //     order as one tile at a time.
        const int nt = (Mw + TILE - 1) / TILE;
        constexpr int TG = sizeof(T) == 4 ? 2 : 4, NK = MNB / MF::KSTEP;   // 32 x 32 FP32 tiles: two
        for (int e0 = wave; e0 < nt * nt; e0 += NW * TG) {
            typename MF::acc_t acc[TG];
            T av[LG ? TG : 1][NK];
            int tis[TG], tjs[TG];
#pragma unroll
            for (int g = 0; g < TG; ++g) {
                const int e = e0 + NW * g;
                tis[g] = e < nt * nt ? e / nt : nt;   // nt: no rows (every entry masked)
                tjs[g] = e < nt * nt ? e % nt : 0;
                // the tile is held transposed (acc^T += U^T (-L)^T): register r
                // of lane l holds window row ii = col(l), column cc = row(l, r),
                // so each load reads whole runs of consecutive rows of a column
                const int ii = tis[g] * TILE + MF::col(lane);
                const T* tp = base + aidx(j0 + nb + ii, j0 + nb + tjs[g] * TILE);
#pragma unroll
                for (int r = 0; r < MF::NACC; ++r) {
                    const int cc = tjs[g] * TILE + MF::row(lane, r);
                    acc[g][r] = (ii < Mw && cc < Mw && ii - cc <= k && cc - ii <= k)
                                    ? tp[(int64_t)MF::row(lane, r) * cstep] : (T)0;
                }
                if constexpr (LG) {
                    // A = -L(j0 + nb + ia, j0 + kk + ka) from the band: row ia
                    // of L_below, panel column kk + ka (zero outside the band,
                    // past the partition's end and past the panel)
                    const int ia = tis[g] * TILE + MF::ab_mn(lane), ka = MF::ab_k(lane);
                    const T* lp = base + aidx(j0 + nb + min(ia, Mw - 1), j0 + ka);
#pragma unroll
                    for (int q = 0; q < NK; ++q) {
                        const int c = q * MF::KSTEP + ka;
                        const bool ok = ia < Mw && c < nb && nb + ia - c <= k;
                        av[g][q] = ok ? -lp[(int64_t)(q * MF::KSTEP) * cstep] : (T)0;
                    }
                }
            }
#pragma unroll
            for (int g = 0; g < TG; ++g) {
                if (e0 + NW * g >= nt * nt) break;   // uniform over the wavefront
                const T* ub = Us + MF::ab_k(lane) * lay.us_stride + tjs[g] * TILE + MF::ab_mn(lane);
                // A = U_right^T (from the staged U block), B = (-L_below)^T
                if constexpr (LG) {
#pragma unroll
                    for (int q = 0; q < NK; ++q)
                        acc[g] = MF::mma(ub[q * MF::KSTEP * lay.us_stride], av[g][q], acc[g]);
                } else {
                    const T* la = Ls + (tis[g] * TILE + MF::ab_mn(lane)) * lay.ls_stride + MF::ab_k(lane);
#pragma unroll
                    for (int q = 0; q < NK; ++q)
                        acc[g] = MF::mma(ub[q * MF::KSTEP * lay.us_stride], la[q * MF::KSTEP], acc[g]);
                }
                const int ii = tis[g] * TILE + MF::col(lane);
                T* tp = base + aidx(j0 + nb + ii, j0 + nb + tjs[g] * TILE);
#pragma unroll
                for (int r = 0; r < MF::NACC; ++r) {
                    const int cc = tjs[g] * TILE + MF::row(lane, r);
                    if (ii < Mw && cc < Mw && ii - cc <= k && cc - ii <= k) tp[(int64_t)MF::row(lane, r) * cstep] = acc[g][r];
                }
            }
        }
        __syncthreads();
    }
}

template <class T>
__global__ void zero_pivot_kernel(const T* __restrict__ B, int k, int64_t row0, int64_t rows, int* flag)
{
    const int64_t w = 2 * k + 1;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < rows; i += stride) {
        const T d = B[(row0 + i) * w + k];
        if (!(d >= (T)1e-7 || d <= (T)-1e-7)) atomicOr(flag, 1);   // |d| < 1e-7 (SmallerThan, precond.h:526-534)
    }
}

// Coupling blocks of boundary i (copydWV, sap/device/data_transfer.cuh:127-199):
// B_i(a,b) = A(e0-k+a, s1+b) (zero above the diagonal, b > a),
// C_{i+1}(a,b) = A(s1+a, e0-k+b) (zero below the diagonal, a > b).
// Stored column-major: offB[b*k + a].
template <class T>
__global__ void extract_coupling_kernel(const T* __restrict__ B, int k, PartInfo pi, T* __restrict__ offB,
                                        T* __restrict__ offC)
{
    // boundary blockIdx.x; columns b = blockIdx.y, blockIdx.y + gridDim.y, ...
    // of both blocks, rows a across the threads (runs along the columns)
    const int i = blockIdx.x;
    int64_t s1, l1;
    pi.bounds(i + 1, s1, l1);
    const int64_t w = 2 * k + 1;
    const size_t kk = (size_t)k * k;
    for (int b = blockIdx.y; b < k; b += gridDim.y) {
        for (int a = threadIdx.x; a < k; a += blockDim.x) {
            const size_t e = (size_t)b * k + a;
            const int64_t rb = s1 - k + a, cb = s1 + b;        // B_i(a, b)
            offB[i * kk + e] = (b <= a) ? B[cb * w + k + rb - cb] : (T)0;
            const int64_t rc = s1 + a, cc = s1 - k + b;        // C_{i+1}(a, b)
            offC[i * kk + e] = (a <= b) ? B[cc * w + k + rc - cc] : (T)0;
        }
    }
}

// ---------------------------------------------------------------------------
// Multi-RHS triangular solves on k x k corner blocks, RC right-hand sides per
// workgroup held in LDS as Y[a][r] (row stride RC+1).
//   lower_first = 1:  Y <- U'^{-1} D^{-1} L^{-1} Y   (LU factors)
//   lower_first = 0:  Y <- L'^{-1} D^{-1} U'^{-1} Y  (UL factors)
// Fl(t,j) = L(t,j) for t > j, Fu(t,j) = U(t,j) for t < j, Fd(t) = pivot.
// ---------------------------------------------------------------------------
constexpr int RC = 32;

template <class T, class FL, class FU, class FD>
__device__ void corner_solve(T* Y, int k, int nrhs, bool lower_first, FL Fl, FU Fu, FD Fd)
{
    constexpr int YS = RC + 1;
    auto lower = [&]() {
        for (int j = 0; j < k - 1; ++j) {
            const int nt = k - 1 - j;
            for (int e = threadIdx.x; e < nt * nrhs; e += blockDim.x) {
                const int t = j + 1 + e % nt, r = e / nt;
                Y[t * YS + r] -= Fl(t, j) * Y[j * YS + r];
            }
            __syncthreads();
        }
    };
    auto upper = [&]() {
        for (int j = k - 1; j > 0; --j) {
            const int nt = j;
            for (int e = threadIdx.x; e < nt * nrhs; e += blockDim.x) {
                const int t = e % nt, r = e / nt;
                Y[t * YS + r] -= Fu(t, j) * Y[j * YS + r];
            }
            __syncthreads();
        }
    };
    auto diag = [&]() {
        for (int e = threadIdx.x; e < k * nrhs; e += blockDim.x) {
            const int t = e % k, r = e / k;
            Y[t * YS + r] /= Fd(t);
        }
        __syncthreads();
    };
    if (lower_first) { lower(); diag(); upper(); }
    else             { upper(); diag(); lower(); }
}

// Spike tips (calculateSpikes, sap/precond.h:6230-6287) with L D U' / U' D L'
// factors.  blockIdx.x = boundary i, blockIdx.y = 0 (V_i) / 1 (W_{i+1}),
// blockIdx.z = RHS chunk.  V_i = bottom k x k block of A_i^{-1}[0; B_i] from
// the LU of partition i; W_{i+1} = top k x k block of A_{i+1}^{-1}[C_{i+1}; 0]
// from the UL of partition i+1.  WV[i][0|1][r*k + a].
template <class T>
__global__ __launch_bounds__(256) void spike_tips_kernel(const T* __restrict__ B, const T* __restrict__ Bul, int k,
                                                         PartInfo pi, const T* __restrict__ offB,
                                                         const T* __restrict__ offC, T* __restrict__ WV)
{
    extern __shared__ unsigned char smem_raw[];
    T* Y = reinterpret_cast<T*>(smem_raw);
    constexpr int YS = RC + 1;
    const int i = blockIdx.x;
    const int which = blockIdx.y;
    const int r0 = blockIdx.z * RC;
    const int nrhs = min(RC, k - r0);
    const int64_t w = 2 * k + 1;
    const size_t kk = (size_t)k * k;
    int64_t s0, l0, s1, l1;
    pi.bounds(i, s0, l0);
    pi.bounds(i + 1, s1, l1);
    const T* src = which == 0 ? offB + i * kk : offC + i * kk;
    for (int e = threadIdx.x; e < k * nrhs; e += blockDim.x) {
        const int a = e % k, r = e / k;
        Y[a * YS + r] = src[(size_t)(r0 + r) * k + a];
    }
    __syncthreads();
    if (which == 0) {
        const int64_t o = s0 + l0 - k;   // first row of the bottom block
        auto el = [&](int t, int j) -> T { const int64_t rr = o + t, cc = o + j; return B[cc * w + k + rr - cc]; };
        corner_solve<T>(Y, k, nrhs, true, el, el, [&](int t) -> T { return el(t, t); });
    } else {
        const int64_t o = s1;
        auto el = [&](int t, int j) -> T { const int64_t rr = o + t, cc = o + j; return Bul[cc * w + k + rr - cc]; };
        corner_solve<T>(Y, k, nrhs, false, el, el, [&](int t) -> T { return el(t, t); });
    }
    T* dst = WV + (2 * i + which) * kk;
    for (int e = threadIdx.x; e < k * nrhs; e += blockDim.x) {
        const int a = e % k, r = e / k;
        dst[(size_t)(r0 + r) * k + a] = Y[a * YS + r];
    }
}

// Spike tips on the matrix cores: the sweeps of corner_solve (unit L, D,
// unit U' in the factorization's order) blocked by 32 columns.  Per block,
// wavefront 0 solves the 32 x 32 unit-triangular diagonal block (one
// right-hand side per lane, the factor column block broadcast from LDS),
// then every wavefront applies the block to the remaining rows with
// 16 x 16 MFMA tiles, Y_rest -= F_blk Y_blk.  The factor's 32 columns are
// staged in LDS transposed (column jj, rows t contiguous), loaded from HBM
// one block ahead.  Same operator as corner_solve; the sums are grouped by
// block, so the values agree to rounding.
template <class T>
struct StkCfg;
template <>
struct StkCfg<float> {
    static constexpr int NR = 64;   // right-hand sides per workgroup
};
template <>
struct StkCfg<double> {
    static constexpr int NR = 32;
};
constexpr int STK_THREADS = 512;
constexpr int STK_IT = 13;          // rows per staged column per lane: k <= 32 * 13

template <class T>
struct StkLayout {
    static constexpr int NR = StkCfg<T>::NR;
    static constexpr int YS = NR + 16;   // Y row stride: the MFMA B operand reads four rows without bank conflicts
    int kpad, lts;
    __host__ __device__ explicit StkLayout(int k)
    {
        kpad = ((k + 15) / 16) * 16 + 32;   // rows a 32-row block or 16-row tile may touch past k (zero)
        // column stride of the staged factor block: the A operand reads four columns
        lts = sizeof(T) == 4 ? ((kpad + 47) / 64) * 64 + 16 : ((kpad + 15) / 32) * 32 + 16;
    }
    __host__ __device__ size_t y_elems() const { return (size_t)kpad * YS; }
    __host__ __device__ size_t bytes() const { return sizeof(T) * (y_elems() + (size_t)32 * lts); }
};

// The blocked sweeps on Y (kpad x YS in LDS, rows >= k zero; LT: 32 x lts,
// zero): lower_first: Y = U'^{-1} D^{-1} L^{-1} Y, else L^{-1} D^{-1} U'^{-1} Y,
// with el(t, j) the factor entry of row t, column j (t, j < k).  Called by
// all STK_THREADS threads after a barrier; ends with one.
template <class T, class El>
__device__ void stk_solve(T* Y, T* LT, const StkLayout<T>& L, int k, bool lower_first, El el)
{
    using MF = Mfma16<T>;
    using acc_t = typename MF::acc_t;
    constexpr int NR = StkCfg<T>::NR, YS = StkLayout<T>::YS, NCT = NR / 16;
    const int lts = L.lts;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // block sequence: the first pass then the second, each over ceil(k/32)
    // blocks (lower passes forward, upper passes backward)
    const int nblk = (k + 31) / 32;
    auto block_lower = [&](int q) -> bool { return lower_first != (q >= nblk); };
    auto block_jb = [&](int q) -> int {
        const int b = q >= nblk ? q - nblk : q;
        return block_lower(q) ? 32 * b : 32 * (nblk - 1 - b);
    };
    // staging: lanes 0-31 / 32-63 of wavefront v take columns 2v, 2v+1 (+16),
    // rows lane%32 + 32 it
    T pf[2][STK_IT];
    auto prefetch = [&](int q) {
        const int jb = block_jb(q);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int j = jb + 2 * wave + (lane >> 5) + 16 * h;
#pragma unroll
            for (int it = 0; it < STK_IT; ++it) {
                const int t = (lane & 31) + 32 * it;
                pf[h][it] = (j < k && t < k) ? el(t, j) : (T)0;
            }
        }
    };
    auto stage = [&]() {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int jj = 2 * wave + (lane >> 5) + 16 * h;
#pragma unroll
            for (int it = 0; it < STK_IT; ++it) {
                const int t = (lane & 31) + 32 * it;
                if (t < k) LT[jj * lts + t] = pf[h][it];
            }
        }
    };
    const int nseq = 2 * nblk;
    prefetch(0);
    for (int q = 0; q < nseq; ++q) {
        const int jb = block_jb(q);
        const bool lower = block_lower(q);
        stage();
        __syncthreads();
        if (q + 1 < nseq) prefetch(q + 1);
        if (wave == 0 && lane < NR) {
            // diagonal block, one right-hand side per lane
            T y[32];
#pragma unroll
            for (int tt = 0; tt < 32; ++tt) y[tt] = Y[(jb + tt) * YS + lane];
            if (lower) {
#pragma unroll
                for (int jj = 0; jj < 31; ++jj) {
                    const T yj = y[jj];
#pragma unroll
                    for (int tt = jj + 1; tt < 32; ++tt) y[tt] -= LT[jj * lts + jb + tt] * yj;
                }
            } else {
#pragma unroll
                for (int jj = 31; jj > 0; --jj) {
                    const T yj = y[jj];
#pragma unroll
                    for (int tt = 0; tt < jj; ++tt) y[tt] -= LT[jj * lts + jb + tt] * yj;
                }
            }
#pragma unroll
            for (int tt = 0; tt < 32; ++tt) Y[(jb + tt) * YS + lane] = y[tt];
        }
        __syncthreads();
        // the rest of the rows: below the block (lower) or above it (upper)
        const int t_lo = lower ? jb + 32 : 0;
        const int nrow = lower ? k - t_lo : jb;
        const int nrt = nrow > 0 ? (nrow + 15) / 16 : 0;
        const int g = lane >> 4, c = lane & 15;
        for (int f = wave; f < nrt * NCT; f += STK_THREADS / 64) {
            const int t0 = t_lo + 16 * (f / NCT), rc0 = 16 * (f % NCT);
            acc_t acc;
#pragma unroll
            for (int qq = 0; qq < 4; ++qq) acc[qq] = Y[(t0 + MF::row(lane, qq)) * YS + rc0 + c];
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                const T a = -LT[(4 * m + g) * lts + t0 + c];
                const T b = Y[(jb + 4 * m + g) * YS + rc0 + c];
                acc = MF::mma(a, b, acc);
            }
#pragma unroll
            for (int qq = 0; qq < 4; ++qq) Y[(t0 + MF::row(lane, qq)) * YS + rc0 + c] = acc[qq];
        }
        __syncthreads();
        if (q == nblk - 1) {
            // between the passes: Y /= D
            for (int e = tid; e < k * NR; e += STK_THREADS) {
                const int t = e / NR, r = e - t * NR;
                Y[t * YS + r] /= el(t, t);
            }
            __syncthreads();
        }
    }
}

template <class T>
__global__ __launch_bounds__(STK_THREADS) void spike_tips_mfma_kernel(const T* __restrict__ B,
                                                                      const T* __restrict__ Bul, int k, PartInfo pi,
                                                                      const T* __restrict__ offB,
                                                                      const T* __restrict__ offC,
                                                                      T* __restrict__ WV)
{
    constexpr int NR = StkCfg<T>::NR, YS = StkLayout<T>::YS;
    const StkLayout<T> L(k);
    extern __shared__ unsigned char smem_raw[];
    T* const Y = reinterpret_cast<T*>(smem_raw);   // Y[t * YS + r], t < kpad
    T* const LT = Y + L.y_elems();                  // LT[jj * lts + t] = F(t, jb + jj)
    const int tid = threadIdx.x;
    const int i = blockIdx.x, which = blockIdx.y, r0 = blockIdx.z * NR;
    const int nrhs = min(NR, k - r0);
    const int64_t w = 2 * k + 1;
    const size_t kk = (size_t)k * k;
    int64_t s0, l0, s1, l1;
    pi.bounds(i, s0, l0);
    pi.bounds(i + 1, s1, l1);
    const T* const F = which == 0 ? B : Bul;
    const int64_t o = which == 0 ? s0 + l0 - k : s1;   // first row / column of the corner block
    // Y = the coupling block's columns r0 .. r0 + nrhs - 1 (zero elsewhere), LT = 0
    // (neighbouring threads read neighbouring rows t of one column, four
    // loads in flight per thread)
    const T* src = (which == 0 ? offB : offC) + i * kk;
    const int ytot = L.kpad * YS;
    for (int e0 = 0; e0 < ytot; e0 += 4 * STK_THREADS) {
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = e0 + tid + u * STK_THREADS;
            const int r = e / L.kpad, t = e - r * L.kpad;
            v[u] = (e < ytot && t < k && r < nrhs) ? src[(size_t)(r0 + r) * k + t] : (T)0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = e0 + tid + u * STK_THREADS;
            const int r = e / L.kpad, t = e - r * L.kpad;
            if (e < ytot) Y[t * YS + r] = v[u];
        }
    }
    for (int e = tid; e < 32 * L.lts; e += STK_THREADS) LT[e] = (T)0;
    __syncthreads();
    // V: the LU's corner, lower sweep first; W: the UL's corner, upper first
    stk_solve<T>(Y, LT, L, k, which == 0, [&](int t, int j) -> T { return F[(o + j) * w + k + t - j]; });
    T* dst = WV + (2 * i + which) * kk;
    for (int e = tid; e < k * nrhs; e += STK_THREADS) {
        const int a = e % k, r = e / k;
        dst[(size_t)(r0 + r) * k + a] = Y[a * YS + r];
    }
}

// S_i^{-1} from the L D U' factors of S_i (column-major, k x k): the blocked
// sweeps on identity columns, NR per workgroup.
template <class T>
__global__ __launch_bounds__(STK_THREADS) void schur_inverse_mfma_kernel(const T* __restrict__ S, int k,
                                                                         T* __restrict__ Sinv)
{
    constexpr int NR = StkCfg<T>::NR, YS = StkLayout<T>::YS;
    const StkLayout<T> L(k);
    extern __shared__ unsigned char smem_raw[];
    T* const Y = reinterpret_cast<T*>(smem_raw);
    T* const LT = Y + L.y_elems();
    const int tid = threadIdx.x;
    const int i = blockIdx.x, r0 = blockIdx.y * NR;
    const int nrhs = min(NR, k - r0);
    const size_t kk = (size_t)k * k;
    const T* A = S + (size_t)i * kk;
    for (int e = tid; e < L.kpad * YS; e += STK_THREADS) {
        const int t = e / YS, r = e - t * YS;
        Y[e] = (t < k && r < nrhs && t == r0 + r) ? (T)1 : (T)0;
    }
    for (int e = tid; e < 32 * L.lts; e += STK_THREADS) LT[e] = (T)0;
    __syncthreads();
    stk_solve<T>(Y, LT, L, k, true, [&](int t, int j) -> T { return A[(size_t)j * k + t]; });
    T* dst = Sinv + (size_t)i * kk;
    for (int e = tid; e < k * nrhs; e += STK_THREADS) {
        const int a = e % k, r = e / k;
        dst[(size_t)(r0 + r) * k + a] = Y[a * YS + r];
    }
}

// W tips for LU_only: the full partition sweep with k right-hand sides
// (calculateSpikes_const, sap/precond.h:5910-5948), RC columns per
// workgroup, working vector in global scratch (len x RC, row-major).
template <class T>
__global__ __launch_bounds__(256) void spike_w_full_kernel(const T* __restrict__ B, int k, PartInfo pi,
                                                           const T* __restrict__ offC, T* __restrict__ WV,
                                                           T* __restrict__ scratch, int64_t scratch_stride)
{
    const int i = blockIdx.x;
    const int r0 = blockIdx.y * RC;
    const int nrhs = min(RC, k - r0);
    const int64_t w = 2 * k + 1;
    const size_t kk = (size_t)k * k;
    int64_t s1, l1;
    pi.bounds(i + 1, s1, l1);
    T* Y = scratch + (size_t)(i * gridDim.y + blockIdx.y) * scratch_stride;
    auto el = [&](int64_t t, int64_t j) -> T { const int64_t rr = s1 + t, cc = s1 + j; return B[cc * w + k + rr - cc]; };
    for (int64_t e = threadIdx.x; e < l1 * RC; e += blockDim.x) {
        const int64_t a = e / RC;
        const int r = (int)(e % RC);
        Y[e] = (a < k && r < nrhs) ? offC[i * kk + (size_t)(r0 + r) * k + a] : (T)0;
    }
    __syncthreads();
    for (int64_t j = 0; j < l1 - 1; ++j) {
        const int nt = (int)min((int64_t)k, l1 - 1 - j);
        for (int e = threadIdx.x; e < nt * nrhs; e += blockDim.x) {
            const int64_t t = j + 1 + e % nt;
            const int r = e / nt;
            Y[t * RC + r] -= el(t, j) * Y[j * RC + r];
        }
        __syncthreads();
    }
    for (int64_t e = threadIdx.x; e < l1 * nrhs; e += blockDim.x) {
        const int64_t t = e / nrhs;
        const int r = (int)(e % nrhs);
        Y[t * RC + r] /= el(t, t);
    }
    __syncthreads();
    for (int64_t j = l1 - 1; j > 0; --j) {
        const int nt = (int)min((int64_t)k, j);
        for (int e = threadIdx.x; e < nt * nrhs; e += blockDim.x) {
            const int64_t t = j - 1 - e % nt;
            const int r = e / nt;
            Y[t * RC + r] -= el(t, j) * Y[j * RC + r];
        }
        __syncthreads();
    }
    T* dst = WV + (2 * i + 1) * kk;
    for (int e = threadIdx.x; e < k * nrhs; e += blockDim.x) {
        const int a = e % k, r = e / k;
        dst[(size_t)(r0 + r) * k + a] = Y[(int64_t)a * RC + r];
    }
}

// ---------------------------------------------------------------------------
// Reduced blocks R_i = [[I, V_i], [W_{i+1}, I]] (assembleReducedMat,
// data_transfer.cuh:14-80) are never formed: their LU needs only the Schur
// complement S_i = I - W_{i+1} V_i (var::fullLU_sub_spec,
// factor_band_var.cuh:734-748), factored in place, and the apply uses
// S_i^{-1}.  All k x k blocks column-major.
// ---------------------------------------------------------------------------
constexpr int GT = 32;   // GEMM tile

template <class T>
__global__ __launch_bounds__(256) void schur_gemm_kernel(const T* __restrict__ WV, int k, T* __restrict__ S)
{
    __shared__ T Wt[GT][GT + 1];
    __shared__ T Vt[GT][GT + 1];
    const int i = blockIdx.x;
    const int a0 = blockIdx.y * GT, b0 = blockIdx.z * GT;
    const size_t kk = (size_t)k * k;
    const T* V = WV + 2 * i * kk;
    const T* W = WV + (2 * i + 1) * kk;
    // the 32 x 32 tile as four 16 x 16 quarters, one per wavefront, on the
    // matrix cores: D = V^T W^T, so accumulator q of lane l holds
    // (W V)(a0 + sa + l % 16, b0 + sb + row(l, q)) and the stores run along
    // the columns of S
    using MF = Mfma16<T>;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int sa = (wv & 1) * 16, sb = (wv >> 1) * 16;
    const int lr = lane & 15, lg = lane >> 4;
    typename MF::acc_t acc = {0, 0, 0, 0};
    for (int c0 = 0; c0 < k; c0 += GT) {
        // both tiles' loads issued together (GT * GT / 256 per thread), then stored
        constexpr int NE = GT * GT / 256;
        T wv[NE], vv[NE];
#pragma unroll
        for (int u = 0; u < NE; ++u) {
            const int e = threadIdx.x + 256 * u;
            const int rr = e % GT, cc = e / GT;
            const int a = a0 + rr, c = c0 + cc, b = b0 + rr;
            wv[u] = (a < k && c < k) ? W[(size_t)c * k + a] : (T)0;   // W(a, c)
            vv[u] = (c0 + cc < k && b < k) ? V[(size_t)b * k + (c0 + cc)] : (T)0;   // V(c, b)
        }
#pragma unroll
        for (int u = 0; u < NE; ++u) {
            const int e = threadIdx.x + 256 * u;
            const int rr = e % GT, cc = e / GT;
            Wt[rr][cc] = wv[u];
            Vt[cc][rr] = vv[u];   // V(c, b) at [c][b]
        }
        __syncthreads();
#pragma unroll
        for (int st = 0; st < GT / 4; ++st)
            acc = MF::mma(Vt[4 * st + lg][sb + lr], Wt[sa + lr][4 * st + lg], acc);
        __syncthreads();
    }
    T* Si = S + i * kk;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int a = a0 + sa + lr, b = b0 + sb + MF::row(lane, q);
        if (a < k && b < k) Si[(size_t)b * k + a] = (a == b ? (T)1 : (T)0) - acc[q];
    }
}

// In-place LU of S_i (Doolittle, no pivoting; boosted pivots only with
// safeFactorization as in var::blockedFullLU_phase1_safe_general), then the
// U rows divided by their pivots (var::fullLU_post_divide).
template <class T>
__global__ __launch_bounds__(256) void schur_lu_kernel(T* __restrict__ S, int k, int safe)
{
    T* A = S + (size_t)blockIdx.x * k * k;
    __shared__ T s_piv;
    for (int j = 0; j < k; ++j) {
        if (threadIdx.x == 0) {
            T p = A[(size_t)j * k + j];
            if (safe) p = boost_pivot(p);
            A[(size_t)j * k + j] = p;
            s_piv = p;
        }
        __syncthreads();
        const T piv = s_piv;
        for (int t = j + 1 + (int)threadIdx.x; t < k; t += blockDim.x) A[(size_t)j * k + t] /= piv;
        __syncthreads();
        const int m = k - 1 - j;
        for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
            const int t = j + 1 + e % m, c = j + 1 + e / m;
            A[(size_t)c * k + t] -= A[(size_t)j * k + t] * A[(size_t)c * k + j];
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < k * k; e += blockDim.x) {
        const int r = e % k, c = e / k;
        if (r < c) A[(size_t)c * k + r] /= A[(size_t)r * k + r];
    }
}

// LU of S_i and S_i^{-1} in one workgroup of 1024 threads with S_i held in
// LDS (k^2 elements: k <= 200 in FP32, 141 in FP64).  The same operations in
// the same order as schur_lu_kernel and schur_inverse_kernel (so the same
// results), without their per-step trips through global memory:
//   LU: per step j the pivot (boosted with safeFactorization), the
//       multipliers of column j, then the trailing update with wavefronts
//       over columns and lanes over rows (each lane keeps its rows'
//       multipliers in registers);
//   U rows divided by their pivots;
//   inverse: one wavefront per identity column, the column vector in
//       registers (lane l holds rows l, l + 64, ...), forward L, D^{-1},
//       backward U' with the pivot entries broadcast by readlane.
template <class T>
__device__ inline T wave_bcast(const T (&y)[4], int m)
{
    const T v = m < 64 ? y[0] : m < 128 ? y[1] : m < 192 ? y[2] : y[3];
    if constexpr (sizeof(T) == 4) {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), m & 63));
    } else {
        const long long b = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), m & 63);
        const int hi = __builtin_amdgcn_readlane((int)(b >> 32), m & 63);
        return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
    }
}

template <class T>
__global__ __launch_bounds__(1024) void schur_lu_inv_kernel(T* __restrict__ S, T* __restrict__ Sinv, int k, int safe,
                                                            int with_inverse)
{
    extern __shared__ unsigned char smem_raw[];
    T* M = reinterpret_cast<T*>(smem_raw);   // S_i, column-major, leading dimension k
    const size_t kk = (size_t)k * k;
    T* A = S + (size_t)blockIdx.x * kk;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nw = blockDim.x >> 6;
    for (size_t e = tid; e < kk; e += blockDim.x) M[e] = A[e];
    __syncthreads();
    for (int j = 0; j < k; ++j) {
        T piv = M[(size_t)j * k + j];
        if (safe) piv = boost_pivot(piv);
        for (int t = j + 1 + tid; t < k; t += blockDim.x) M[(size_t)j * k + t] /= piv;
        __syncthreads();   // multipliers written, every thread has read the pivot
        if (tid == 0) M[(size_t)j * k + j] = piv;
        T lm[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int t = j + 1 + lane + 64 * q;
            lm[q] = t < k ? M[(size_t)j * k + t] : (T)0;
        }
        for (int c = j + 1 + wave; c < k; c += nw) {
            const T u = M[(size_t)c * k + j];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int t = j + 1 + lane + 64 * q;
                if (t < k) M[(size_t)c * k + t] -= lm[q] * u;
            }
        }
        __syncthreads();
    }
    for (size_t e = tid; e < kk; e += blockDim.x) {
        const int r = (int)(e % k), c = (int)(e / k);
        if (r < c) M[e] /= M[(size_t)r * k + r];
    }
    __syncthreads();
    for (size_t e = tid; e < kk; e += blockDim.x) A[e] = M[e];
    if (!with_inverse) return;
    T* X = Sinv + (size_t)blockIdx.x * kk;
    for (int r = wave; r < k; r += nw) {
        T y[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) y[q] = (lane + 64 * q == r) ? (T)1 : (T)0;
        for (int m = 0; m < k - 1; ++m) {   // unit lower L
            const T ym = wave_bcast(y, m);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int t = lane + 64 * q;
                if (t > m && t < k) y[q] -= M[(size_t)m * k + t] * ym;
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int t = lane + 64 * q;
            if (t < k) y[q] /= M[(size_t)t * k + t];
        }
        for (int m = k - 1; m > 0; --m) {   // unit upper U'
            const T ym = wave_bcast(y, m);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int t = lane + 64 * q;
                if (t < m) y[q] -= M[(size_t)m * k + t] * ym;
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int t = lane + 64 * q;
            if (t < k) X[(size_t)r * k + t] = y[q];
        }
    }
}

// The same L D U' of S_i (the same operations on each entry, in the same
// order) with the matrix in registers, for k <= 32 R: thread (tr, tc) =
// (tid % 32, tid / 32) holds rows tr + 32 i and columns tc + 32 j (i, j < R).
// Step j: the half-wavefront holding column j divides it by the pivot (read
// from the pivot's lane) and publishes the multipliers, the threads holding
// row j publish it, and after one barrier every thread applies the rank-1
// update to its R x R entries from two LDS vectors; rows and columns that
// are done see a zero multiplier or a zero row entry, which leaves them
// unchanged.  The published vectors alternate between two buffers, so one
// barrier per step suffices.
template <class T>
__device__ inline T lane_value(T v, int src)
{
    if constexpr (sizeof(T) == 4) {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
    } else {
        const long long b = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), src);
        const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
        return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
    }
}

template <class T, int R>
__global__ __launch_bounds__(1024) void schur_lu_reg_kernel(T* __restrict__ S, int k, int safe)
{
    __shared__ __align__(16) T lbuf[2][32 * 8];
    __shared__ __align__(16) T ubuf[2][32 * 8];
    __shared__ T dpiv[32 * R];
    const int tid = threadIdx.x, tr = tid & 31, tc = tid >> 5;
    T* A = S + (size_t)blockIdx.x * k * k;
    T a[R][R];
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const int r = tr + 32 * i, c = tc + 32 * j;
            a[i][j] = (r < k && c < k) ? A[(size_t)c * k + r] : (T)0;
        }
#pragma unroll
    for (int jb = 0; jb < R; ++jb) {
        for (int jj = 0; jj < 32 && 32 * jb + jj < k; ++jj) {
            const int j = 32 * jb + jj, buf = j & 1;
            if (tc == jj) {   // column j (one half-wavefront)
                T piv = lane_value(a[jb][jb], ((jj & 1) << 5) | jj);
                if (safe) piv = boost_pivot(piv);
#pragma unroll
                for (int i = 0; i < R; ++i) {
                    const int r = tr + 32 * i;
                    T l = (T)0;
                    if (r > j && r < k) {
                        l = a[i][jb] / piv;
                        a[i][jb] = l;
                    }
                    lbuf[buf][tr * 8 + i] = l;
                }
                if (tr == jj) {
                    a[jb][jb] = piv;
                    dpiv[j] = piv;
                }
            }
            if (tr == jj) {   // row j
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    const int c = tc + 32 * q;
                    ubuf[buf][tc * 8 + q] = (c > j && c < k) ? a[jb][q] : (T)0;
                }
            }
            __syncthreads();
            T l[R], u[R];
#pragma unroll
            for (int i = 0; i < R; ++i) {
                l[i] = lbuf[buf][tr * 8 + i];
                u[i] = ubuf[buf][tc * 8 + i];
            }
#pragma unroll
            for (int i = 0; i < R; ++i)
#pragma unroll
                for (int q = 0; q < R; ++q) a[i][q] -= l[i] * u[q];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int r = tr + 32 * i, c = tc + 32 * q;
            if (r < k && c < k) {
                if (r < c) a[i][q] /= dpiv[r];
                A[(size_t)c * k + r] = a[i][q];
            }
        }
}

template <class T, int R>
bool launch_schur_lu_reg(T* S, int k, int nb, int safe, hipStream_t st)
{
    if constexpr (R > (sizeof(T) == 4 ? 7 : 5)) {
        return false;
    } else {
        if (k > 32 * R) return launch_schur_lu_reg<T, R + 1>(S, k, nb, safe, st);
        hipLaunchKernelGGL((schur_lu_reg_kernel<T, R>), dim3(nb), dim3(1024), 0, st, S, k, safe);
        SAP_LAUNCH_CHECK();
        return true;
    }
}

// S_i^{-1} from its L D U' factors: RC identity columns per workgroup.
template <class T>
__global__ __launch_bounds__(256) void schur_inverse_kernel(const T* __restrict__ S, int k, T* __restrict__ Sinv)
{
    extern __shared__ unsigned char smem_raw[];
    T* Y = reinterpret_cast<T*>(smem_raw);
    constexpr int YS = RC + 1;
    const int i = blockIdx.x;
    const int r0 = blockIdx.y * RC;
    const int nrhs = min(RC, k - r0);
    const T* A = S + (size_t)i * k * k;
    for (int e = threadIdx.x; e < k * nrhs; e += blockDim.x) {
        const int a = e % k, r = e / k;
        Y[a * YS + r] = (a == r0 + r) ? (T)1 : (T)0;
    }
    __syncthreads();
    auto el = [&](int t, int j) -> T { return A[(size_t)j * k + t]; };
    corner_solve<T>(Y, k, nrhs, true, el, el, [&](int t) -> T { return el(t, t); });
    T* dst = Sinv + (size_t)i * k * k;
    for (int e = threadIdx.x; e < k * nrhs; e += blockDim.x) {
        const int a = e % k, r = e / k;
        dst[(size_t)(r0 + r) * k + a] = Y[a * YS + r];
    }
}

// 1 for each +-0 among the words of v (FP32: each word; FP64: each pair)
template <class T>
__device__ inline unsigned zeros_in(uint4 v)
{
    if constexpr (sizeof(T) == 4) {
        return ((v.x << 1) == 0u) + ((v.y << 1) == 0u) + ((v.z << 1) == 0u) + ((v.w << 1) == 0u);
    } else {
        return ((v.x | (v.y << 1)) == 0u) + ((v.z | (v.w << 1)) == 0u);
    }
}

// zeros in p[0, n): elements up to 16-byte alignment and past the last full
// 16 bytes one at a time, the rest in 16-byte loads
template <class T>
__device__ inline unsigned count_zeros_range(const T* __restrict__ p, int64_t n, int64_t t, int64_t stride)
{
    constexpr int V = 16 / sizeof(T);
    int64_t head = (int64_t)(((16 - ((uintptr_t)p & 15)) & 15) / sizeof(T));
    if (head > n) head = n;
    const int64_t nv = (n - head) / V;
    unsigned c = 0;
    for (int64_t i = t; i < head; i += stride) c += p[i] == (T)0;
    const uint4* q = reinterpret_cast<const uint4*>(p + head);
    for (int64_t i = t; i < nv; i += stride) c += zeros_in<T>(q[i]);
    for (int64_t i = head + nv * V + t; i < n; i += stride) c += p[i] == (T)0;
    return c;
}

template <class T>
__global__ void count_zero_kernel(const T* __restrict__ A, int64_t na, const T* __restrict__ B, int64_t nb,
                                  unsigned long long* out)
{
    // zeros in A[0, na) and B[0, nb)
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned local = count_zeros_range<T>(A, na, t, stride);
    if (nb > 0) local += count_zeros_range<T>(B, nb, t, stride);
    for (int off = 32; off > 0; off >>= 1) local += __shfl_down(local, off, 64);
    if ((threadIdx.x & 63) == 0 && local) atomicAdd(out, (unsigned long long)local);
}

template <class T>
size_t factor_lds_bytes(int k, int nb)
{
    return sizeof(T) * ((size_t)(nb + k) * (nb + 1) + (size_t)nb * k);
}

template <class T, int NB>
void launch_factor_nb(T* B0, T* B1, int k, const FactorJob* jobs, int njobs, hipStream_t st)
{
    const size_t lds = factor_lds_bytes<T>(k, NB);
    SAP_HIP_CHECK(hipFuncSetAttribute((const void*)band_factor_kernel<T, NB>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((band_factor_kernel<T, NB>), dim3(njobs), dim3(256), lds, st, B0, B1, k, jobs);
    SAP_LAUNCH_CHECK();
}

constexpr size_t LDS_LIMIT = 160 * 1024 - 1024;   // leave room for static LDS

}  // namespace

template <class T, int NT, bool LG>
void launch_factor_mfma(T* B0, T* B1, int k, const FactorJob* jobs, int njobs, hipStream_t st)
{
    const size_t lds = MfmaLds<T>(k, LG).bytes;
    SAP_HIP_CHECK(hipFuncSetAttribute((const void*)band_factor_mfma_kernel<T, NT, LG>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((band_factor_mfma_kernel<T, NT, LG>), dim3(njobs), dim3(NT), lds, st, B0, B1, k, jobs);
    SAP_LAUNCH_CHECK();
}

// The matrix-core blocked kernel's variant for half-bandwidth k: 100 (256
// threads, -L_below in LDS), 101 (512 threads), 102 (512 threads, -L_below
// from the band); -1 when none applies.
template <class T>
int factor_mfma_variant(int k)
{
    if (k + MNB <= 256 && MfmaLds<T>(k).bytes <= LDS_LIMIT) return 100;
    if (k + MNB <= 512 && MfmaLds<T>(k).bytes <= LDS_LIMIT) return 101;
    if (k + MNB <= 512 && MfmaLds<T>(k, true).bytes <= LDS_LIMIT) return 102;
    return -1;
}

template <class T>
int band_factor_kernel_class(int k, int64_t max_len)
{
    if (band_factor_resident_fits<T>(k, max_len)) return (k + 32 + 15) / 16;   // resident, by window tiles
    const int mv = factor_mfma_variant<T>(k);
    if (mv > 0) return mv;
    if (factor_lds_bytes<T>(k, 32) <= LDS_LIMIT) return 132;
    if (factor_lds_bytes<T>(k, 16) <= LDS_LIMIT) return 116;
    if (factor_lds_bytes<T>(k, 8) <= LDS_LIMIT) return 108;
    return -1;
}

template <class T>
void launch_band_factor(T* B0, T* B1, const T* Bin, int k, const FactorJob* d_jobs, int njobs, int64_t max_len,
                        hipStream_t st)
{
    if (njobs <= 0) return;
    if (launch_band_factor_resident<T>(B0, B1, Bin, k, d_jobs, njobs, max_len, st)) return;
    if (Bin) throw Error(SAP_ERR_HIP, "separate-input factorization needs the register-resident kernel");
    switch (factor_mfma_variant<T>(k)) {
    case 100: launch_factor_mfma<T, 256, false>(B0, B1, k, d_jobs, njobs, st); return;
    case 101: launch_factor_mfma<T, 512, false>(B0, B1, k, d_jobs, njobs, st); return;
    case 102: launch_factor_mfma<T, 512, true>(B0, B1, k, d_jobs, njobs, st); return;
    default: break;
    }
    if (factor_lds_bytes<T>(k, 32) <= LDS_LIMIT)      launch_factor_nb<T, 32>(B0, B1, k, d_jobs, njobs, st);
    else if (factor_lds_bytes<T>(k, 16) <= LDS_LIMIT) launch_factor_nb<T, 16>(B0, B1, k, d_jobs, njobs, st);
    else if (factor_lds_bytes<T>(k, 8) <= LDS_LIMIT)  launch_factor_nb<T, 8>(B0, B1, k, d_jobs, njobs, st);
    else throw Error(SAP_ERR_UNSUPPORTED, "half-bandwidth too large for the banded factorization kernel");
}

template <class T>
void launch_zero_pivot_check(const T* B, int k, int64_t row0, int64_t rows, int* d_flag, hipStream_t st)
{
    if (rows <= 0) return;
    int64_t g = (rows + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(zero_pivot_kernel<T>, dim3((unsigned)g), dim3(256), 0, st, B, k, row0, rows, d_flag);
    SAP_LAUNCH_CHECK();
}

template <class T>
void launch_extract_coupling(const T* B, int k, PartInfo pi, T* offB, T* offC, hipStream_t st)
{
    if (pi.P < 2) return;
    hipLaunchKernelGGL(extract_coupling_kernel<T>, dim3(pi.P - 1, std::min(k, 32)), dim3(256), 0, st, B, k, pi, offB,
                       offC);
    SAP_LAUNCH_CHECK();
}

template <class T>
void launch_spike_tips(const T* B, const T* Bul, int k, PartInfo pi, int fact_lu_ul, const T* offB, const T* offC,
                       T* WV, hipStream_t st)
{
    if (pi.P < 2) return;
    if (fact_lu_ul && k <= 32 * STK_IT && StkLayout<T>(k).bytes() <= LDS_LIMIT) {
        const size_t lds = StkLayout<T>(k).bytes();
        SAP_HIP_CHECK(hipFuncSetAttribute((const void*)spike_tips_mfma_kernel<T>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        constexpr int NR = StkCfg<T>::NR;
        hipLaunchKernelGGL(spike_tips_mfma_kernel<T>, dim3(pi.P - 1, 2, (k + NR - 1) / NR), dim3(STK_THREADS), lds,
                           st, B, Bul, k, pi, offB, offC, WV);
        SAP_LAUNCH_CHECK();
        return;
    }
    const size_t lds = sizeof(T) * (size_t)k * (RC + 1);
    if (lds > LDS_LIMIT) throw Error(SAP_ERR_UNSUPPORTED, "half-bandwidth too large for the spike kernel");
    SAP_HIP_CHECK(hipFuncSetAttribute((const void*)spike_tips_kernel<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds));
    const unsigned chunks = (unsigned)((k + RC - 1) / RC);
    if (fact_lu_ul) {
        hipLaunchKernelGGL(spike_tips_kernel<T>, dim3(pi.P - 1, 2, chunks), dim3(256), lds, st, B, Bul, k, pi, offB,
                           offC, WV);
        SAP_LAUNCH_CHECK();
    } else {
        hipLaunchKernelGGL(spike_tips_kernel<T>, dim3(pi.P - 1, 1, chunks), dim3(256), lds, st, B, Bul, k, pi, offB,
                           offC, WV);
        SAP_LAUNCH_CHECK();
        const int64_t maxlen = pi.n / pi.P + 1;
        const int64_t stride = maxlen * RC;
        DevBuf<T> scratch((size_t)stride * (pi.P - 1) * chunks);
        hipLaunchKernelGGL(spike_w_full_kernel<T>, dim3(pi.P - 1, chunks), dim3(256), 0, st, B, k, pi, offC, WV,
                           scratch.get(), stride);
        SAP_LAUNCH_CHECK();
        SAP_HIP_CHECK(hipStreamSynchronize(st));   // scratch is released on return
    }
}

template <class T>
void launch_reduced_factor(const T* WV, int k, int nb, T* S, T* Sinv, int safe, hipStream_t st)
{
    if (nb <= 0) return;
    const unsigned tiles = (unsigned)((k + GT - 1) / GT);
    hipLaunchKernelGGL(schur_gemm_kernel<T>, dim3(nb, tiles, tiles), dim3(256), 0, st, WV, k, S);
    SAP_LAUNCH_CHECK();
    const size_t lds_lu = sizeof(T) * (size_t)k * k;
    if (lds_lu <= LDS_LIMIT && k <= 256) {
        // LU in LDS; the inverse by the blocked matrix-core sweeps when they fit
        const bool mfma_inv = k <= 32 * STK_IT && StkLayout<T>(k).bytes() <= LDS_LIMIT;
        if (!mfma_inv || !launch_schur_lu_reg<T, 1>(S, k, nb, safe, st)) {
            SAP_HIP_CHECK(hipFuncSetAttribute((const void*)schur_lu_inv_kernel<T>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_lu));
            hipLaunchKernelGGL(schur_lu_inv_kernel<T>, dim3(nb), dim3(1024), lds_lu, st, S, Sinv, k, safe,
                               mfma_inv ? 0 : 1);
            SAP_LAUNCH_CHECK();
        }
        if (mfma_inv) {
            const size_t lds = StkLayout<T>(k).bytes();
            constexpr int NR = StkCfg<T>::NR;
            SAP_HIP_CHECK(hipFuncSetAttribute((const void*)schur_inverse_mfma_kernel<T>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            hipLaunchKernelGGL(schur_inverse_mfma_kernel<T>, dim3(nb, (k + NR - 1) / NR), dim3(STK_THREADS), lds, st,
                               S, k, Sinv);
            SAP_LAUNCH_CHECK();
        }
        return;
    }
    hipLaunchKernelGGL(schur_lu_kernel<T>, dim3(nb), dim3(256), 0, st, S, k, safe);
    SAP_LAUNCH_CHECK();
    const size_t lds = sizeof(T) * (size_t)k * (RC + 1);
    SAP_HIP_CHECK(hipFuncSetAttribute((const void*)schur_inverse_kernel<T>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(schur_inverse_kernel<T>, dim3(nb, (k + RC - 1) / RC), dim3(256), lds, st, S, k, Sinv);
    SAP_LAUNCH_CHECK();
}

template <class T>
void count_band_zeros_async(const T* A, int64_t na, const T* B, int64_t nb, unsigned long long* d_cnt,
                            unsigned long long* h_out, hipStream_t st, int max_groups)
{
    SAP_HIP_CHECK(hipMemsetAsync(d_cnt, 0, sizeof(unsigned long long), st));
    int64_t g = (na + nb + 255) / 256;
    if (g > max_groups) g = max_groups;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(count_zero_kernel<T>, dim3((unsigned)g), dim3(256), 0, st, A, na, B, nb, d_cnt);
    SAP_LAUNCH_CHECK();
    SAP_HIP_CHECK(hipMemcpyAsync(h_out, d_cnt, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
}

template <class T>
int64_t count_band_zeros(const T* A, int64_t na, const T* B, int64_t nb, unsigned long long* d_cnt, hipStream_t st)
{
    unsigned long long h = 0;
    count_band_zeros_async<T>(A, na, B, nb, d_cnt, &h, st, 8192);
    SAP_HIP_CHECK(hipStreamSynchronize(st));
    return (int64_t)h;
}

#define SAP_INST(T)                                                                                          \
    template int64_t count_band_zeros<T>(const T*, int64_t, const T*, int64_t, unsigned long long*, hipStream_t); \
    template void count_band_zeros_async<T>(const T*, int64_t, const T*, int64_t, unsigned long long*,        \
                                            unsigned long long*, hipStream_t, int);                           \
    template void launch_band_factor<T>(T*, T*, const T*, int, const FactorJob*, int, int64_t, hipStream_t);                     \
    template int band_factor_kernel_class<T>(int, int64_t);                                                   \
    template void launch_zero_pivot_check<T>(const T*, int, int64_t, int64_t, int*, hipStream_t);             \
    template void launch_extract_coupling<T>(const T*, int, PartInfo, T*, T*, hipStream_t);                   \
    template void launch_spike_tips<T>(const T*, const T*, int, PartInfo, int, const T*, const T*, T*,        \
                                       hipStream_t);                                                          \
    template void launch_reduced_factor<T>(const T*, int, int, T*, T*, int, hipStream_t);
SAP_INST(float)
SAP_INST(double)
#undef SAP_INST

}  // namespace sap_amd
