// Register-resident MFMA banded LU/UL of the partitions (the factorization
// used for half-bandwidths up to 224 in FP32 and 128 in FP64).
//
// Band layout (column-major, as the reference's m_B, sap/device/
// data_transfer.cuh:440-447): A(r, c) at B[c*(2k+1) + k + r - c]; a UL
// factorization is the LU factorization of the same storage read backwards.
#include "sap_internal.hpp"

namespace sap_amd {

namespace {

// ---------------------------------------------------------------------------
// Register-resident blocked banded LU, 512 threads (8 wavefronts) per
// partition, 32-column panels, 16 x 16 MFMA tiles.
// The active region of panel p, rows and columns [32p, 32p + 16 KT) with
// KT = ceil((k + 32) / 16), stays in the MFMA accumulators of the eight
// wavefronts for the whole factorization: tile (R, C) of the partition (tile
// row R, tile column C) lives in slot (R mod KT, C mod KT), and slot s
// belongs to wavefront s mod 8.  When the elimination moves down the diagonal
// by 32 rows, the tiles that leave the region are exactly the next panel's
// columns and U rows, and the slots they free take the tiles that enter, so
// every band entry is read from HBM once and every factor entry written once.
// Per panel:
//   A.  slots holding the panel columns and the U rows are copied to LDS and
//       start loading the entering tiles (used one panel later);
//   B1. wavefront 0 factors the 32 x 32 diagonal block in registers, one row
//       per lane, pivot rows broadcast with readlane, boosting at pivot time;
//   B2. 16(KT-2) threads solve L21 = A21 U11^-1 row by row, another 16(KT-2)
//       threads U12 = L11^-1 A12 column by column (unit L11);
//   C.  trailing tiles += (-L21) U12 on the matrix cores
//       (v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64, the 32 panel columns
//       split over the four lane groups), U12 / pivot written out.
// Outside the band every stored value is an exact zero, so the padded tiles
// never create fill-in.  Factors are stored as L (unit), D, U' like the other
// factorization kernels.
// ---------------------------------------------------------------------------
template <class T>
struct Mfma16;

template <>
struct Mfma16<float> {
    typedef float acc_t __attribute__((ext_vector_type(4)));
    __device__ static inline acc_t mma(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // accumulator register q of lane l holds D[row(l, q)][l % 16]
    __device__ static inline int row(int l, int q) { return 4 * (l >> 4) + q; }
};

template <>
struct Mfma16<double> {
    typedef double acc_t __attribute__((ext_vector_type(4)));
    __device__ static inline acc_t mma(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static inline int row(int l, int q) { return (l >> 4) + 4 * q; }
};

__device__ inline float lane_bcast(float v, int src)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

__device__ inline double lane_bcast(double v, int src)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <class T, int KT>
struct ResLds {
    static constexpr int PS = 36;                         // row stride: 16-byte aligned rows, few bank conflicts
    static constexpr int NTR = 16 * (KT - 2);             // trailing rows / columns
    static constexpr int pn_off = 0;                      // Pn [16 KT][PS]: panel rows, then -L21
    static constexpr int ub_off = pn_off + 16 * KT * PS;  // UbT [NTR][PS]: U block, one column per row
    static constexpr int lt_off = ub_off + NTR * PS;      // L11 transposed [32][33]
    static constexpr int pv_off = lt_off + 32 * 33;       // pivots [32]
    static constexpr size_t bytes = sizeof(T) * (size_t)(pv_off + 32);
};

template <class T, int KT>
__global__ __launch_bounds__(512) void band_factor_resident_kernel(T* __restrict__ B0, T* __restrict__ B1, int k,
                                                                   const FactorJob* __restrict__ jobs)
{
    using MF = Mfma16<T>;
    using acc_t = typename MF::acc_t;
    using LY = ResLds<T, KT>;
    constexpr int PS = LY::PS, NTR = LY::NTR, NS = KT * KT, SPW = (NS + 7) / 8;
    const FactorJob job = jobs[blockIdx.x];
    T* __restrict__ base = job.buf ? B1 : B0;
    const int64_t w = 2 * k + 1;
    const int64_t len = job.len;
    const int dir = job.dir;
    const int64_t origin = dir > 0 ? job.start * w : (job.start + len) * w - 1;
    const int wm1 = 2 * k;   // next column, same row

    extern __shared__ unsigned char smem_raw[];
    T* const sm = reinterpret_cast<T*>(smem_raw);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // uniform: slot logic stays scalar

    // Entry (j0 + dr, j0 + dc) of the partition is P[dir * (dc * 2k + dr)]
    // with P = base + origin + dir * (j0 * (2k + 1) + k); rem = len - j0.
    auto inb = [&](int dr, int dc, int rem) { return dr < rem && dc < rem && dr - dc <= k && dc - dr <= k; };
    auto load_tile = [&](acc_t& t, const T* P, int rem, int dr0, int dc0, int lr, int lane) {
        const int dc = dc0 + lr;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int dr = dr0 + MF::row(lane, q);
            t[q] = inb(dr, dc, rem) ? P[dir * (dc * wm1 + dr)] : (T)0;
        }
    };

    acc_t acc[SPW];
    {
        const int lane = threadIdx.x & 63;
        const T* P = base + origin + dir * (int64_t)k;
        const int rem = (int)min(len, (int64_t)(1 << 30));
#pragma unroll
        for (int j = 0; j < SPW; ++j) {
            const int s = wave + 8 * j;
            if (s < NS) load_tile(acc[j], P, rem, 16 * (s / KT), 16 * (s % KT), lane & 15, lane);
        }
    }

    const int64_t npan = (len + 31) / 32;
    for (int64_t p = 0; p < npan; ++p) {
        // lane-derived values are re-materialised every panel (keeps the
        // per-lane predicates of the unrolled loops out of the registers)
        int tid = threadIdx.x, zoff = 0;
        asm volatile("" : "+v"(tid), "+v"(zoff));
        // LDS addresses from an opaque zero fold into the ds_* immediate
        // offsets instead of living in scalar registers across the loop
        T* Pn = sm + zoff + LY::pn_off;
        T* UbT = sm + zoff + LY::ub_off;
        T* L11T = sm + zoff + LY::lt_off;
        T* pivs = sm + zoff + LY::pv_off;
        const int lane = tid & 63, lr = lane & 15, lg = lane >> 4;
        const int64_t j0 = 32 * p;
        T* __restrict__ P = base + origin + dir * (j0 * w + k);
        const int rem = (int)min(len - j0, (int64_t)(1 << 30));
        const int nb = min(32, rem);
        const int sh = (int)((2 * p) % KT);

        // A. leaving tiles to LDS; their slots load the entering tiles
#pragma unroll
        for (int j = 0; j < SPW; ++j) {
            const int s = wave + 8 * j;
            if (s < NS) {
                const int rho = (s / KT - sh + KT) % KT, gam = (s % KT - sh + KT) % KT;
                if (gam < 2) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) Pn[(16 * rho + MF::row(lane, q)) * PS + 16 * gam + lr] = acc[j][q];
                } else if (rho < 2) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) UbT[(16 * (gam - 2) + lr) * PS + 16 * rho + MF::row(lane, q)] = acc[j][q];
                }
                if (gam < 2 || rho < 2) {
                    const int rn = rho < 2 ? rho + KT - 2 : rho - 2;
                    const int cn = gam < 2 ? gam + KT - 2 : gam - 2;
                    load_tile(acc[j], P, rem, 32 + 16 * rn, 32 + 16 * cn, lr, lane);
                }
            }
        }
        __syncthreads();

        // B1. LU of the diagonal block, row r in lane r (lanes 32..63 mirror)
        if (wave == 0) {
            const int r = lane & 31;
            T row[32];
            T pr = (T)0;   // this lane's pivot
#pragma unroll
            for (int c = 0; c < 32; ++c) row[c] = Pn[r * PS + c];
#pragma unroll
            for (int jj = 0; jj < 32; ++jj) {
                if (jj < nb) {
                    T piv = lane_bcast(row[jj], jj);
                    if (job.boost && (jj + 1 < rem || job.boost_last)) piv = boost_pivot(piv);
                    const bool below = r > jj;
                    const T l = row[jj] / piv;
                    pr = r == jj ? piv : pr;
                    row[jj] = below ? l : (r == jj ? piv : row[jj]);
#pragma unroll
                    for (int c = jj + 1; c < 32; ++c) {
                        const T u = lane_bcast(row[c], jj);   // pivot row, read before any lane changes it
                        row[c] = below ? row[c] - l * u : row[c];
                    }
                }
            }
            if (lane < 32) {
                pivs[r] = pr;
#pragma unroll
                for (int c = 0; c < 32; ++c) {
                    Pn[r * PS + c] = row[c];
                    if (c < r) L11T[c * 33 + r] = row[c];
                    if (inb(r, c, nb)) P[dir * (c * wm1 + r)] = c > r ? row[c] / pr : row[c];
                }
            }
        }
        __syncthreads();

        // B2. L21 rows and U12 columns
        if (tid < NTR) {
            const int t = tid;
            T a[32];
#pragma unroll
            for (int c = 0; c < 32; ++c) a[c] = Pn[(32 + t) * PS + c];
#pragma unroll
            for (int jj = 0; jj < 32; ++jj) {
                if (jj < nb) {
                    const T l = a[jj] / pivs[jj];
                    a[jj] = l;
#pragma unroll
                    for (int c = jj + 1; c < 32; ++c) a[c] -= l * Pn[jj * PS + c];
                }
            }
            const bool rowin = 32 + t < rem;
#pragma unroll
            for (int c = 0; c < 32; ++c) {
                Pn[(32 + t) * PS + c] = -a[c];
                if (rowin && c < nb && 32 + t - c <= k) P[dir * (c * wm1 + 32 + t)] = a[c];
            }
        } else if (tid >= 256 && tid < 256 + NTR) {
            const int cc = tid - 256;
            T u[32];
#pragma unroll
            for (int m = 0; m < 32; ++m) u[m] = UbT[cc * PS + m];
#pragma unroll
            for (int jj = 0; jj < 31; ++jj) {
                if (jj < nb) {
                    const T uj = u[jj];
#pragma unroll
                    for (int m = jj + 1; m < 32; ++m) u[m] -= L11T[jj * 33 + m] * uj;
                }
            }
#pragma unroll
            for (int m = 0; m < 32; ++m) UbT[cc * PS + m] = u[m];
        }
        __syncthreads();

        // C. trailing update on the matrix cores; U' rows out
#pragma unroll
        for (int j = 0; j < SPW; ++j) {
            const int s = wave + 8 * j;
            if (s < NS) {
                const int rho = (s / KT - sh + KT) % KT, gam = (s % KT - sh + KT) % KT;
                if (rho >= 2 && gam >= 2) {
                    const T* ap = Pn + (16 * rho + lr) * PS + 8 * lg;
                    const T* bp = UbT + (16 * (gam - 2) + lr) * PS + 8 * lg;
                    T av[8], bv[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q) { av[q] = ap[q]; bv[q] = bp[q]; }
#pragma unroll
                    for (int q = 0; q < 8; ++q) acc[j] = MF::mma(av[q], bv[q], acc[j]);
                }
            }
        }
        for (int e = tid; e < 32 * NTR; e += 512) {
            const int cc = e >> 5, m = e & 31;
            if (m < nb && 32 + cc < rem && 32 + cc - m <= k) P[dir * ((32 + cc) * wm1 + m)] = UbT[cc * PS + m] / pivs[m];
        }
        __syncthreads();
    }
}

template <class T, int KT>
void launch_factor_resident(T* B0, T* B1, int k, const FactorJob* jobs, int njobs, hipStream_t st)
{
    const size_t lds = ResLds<T, KT>::bytes;
    SAP_HIP_CHECK(hipFuncSetAttribute((const void*)band_factor_resident_kernel<T, KT>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((band_factor_resident_kernel<T, KT>), dim3(njobs), dim3(512), lds, st, B0, B1, k, jobs);
    SAP_LAUNCH_CHECK();
}

// Largest window (in 16-row tiles) kept in registers: FP32 up to k = 224,
// FP64 (twice the registers per tile) up to k = 128.
template <class T>
constexpr int resident_max_kt() { return sizeof(T) == 4 ? 16 : 10; }

template <class T, int KT>
bool dispatch_resident(int kt, T* B0, T* B1, int k, const FactorJob* jobs, int njobs, hipStream_t st)
{
    if constexpr (KT > resident_max_kt<T>()) {
        return false;
    } else {
        if (kt == KT) { launch_factor_resident<T, KT>(B0, B1, k, jobs, njobs, st); return true; }
        return dispatch_resident<T, KT + 1>(kt, B0, B1, k, jobs, njobs, st);
    }
}

}  // namespace

template <class T>
bool launch_band_factor_resident(T* B0, T* B1, int k, const FactorJob* d_jobs, int njobs, hipStream_t st)
{
    return dispatch_resident<T, 3>((k + 32 + 15) / 16, B0, B1, k, d_jobs, njobs, st);
}

template bool launch_band_factor_resident<float>(float*, float*, int, const FactorJob*, int, hipStream_t);
template bool launch_band_factor_resident<double>(double*, double*, int, const FactorJob*, int, hipStream_t);

}  // namespace sap_amd
