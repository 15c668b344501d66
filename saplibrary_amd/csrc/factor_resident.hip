// Register-resident MFMA banded LU/UL of the partitions (the factorization
// used for half-bandwidths up to 224 in FP32 and 112 in FP64).
//
// Band layout (column-major, as the reference's m_B, sap/device/
// data_transfer.cuh:440-447): A(r, c) at B[c*(2k+1) + k + r - c]; a UL
// factorization is the LU factorization of the same storage read backwards.
#include "mfma.hpp"
#include "sap_internal.hpp"

#include <climits>

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
//   A0. the slots of the 32 x 32 diagonal block are copied to LDS;
//   B1. wavefront 0 factors that block in registers, one row per lane, pivot
//       rows broadcast with readlane, boosting at pivot time, while
//   A.  the other wavefronts (and wavefront 0 after B1) copy the other panel
//       column and U row slots to LDS, start loading the entering tiles
//       (first used one panel later) and store the previous panel's factors
//       from the other LDS buffer;
//   B2. 16(KT-2) threads solve L21 = A21 U11^-1 row by row, another 16(KT-2)
//       threads U12 = L11^-1 A12 column by column (unit L11);
//   C.  trailing tiles += (-L21) U12 on the matrix cores
//       (v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64, the 32 panel columns
//       split over the four lane groups).  Tiles are kept transposed in the
//       accumulators (the update is acc^T += U12^T (-L21)^T), which makes each
//       tile load read whole 16-row column segments.
// Outside the band every stored value is an exact zero, so the padded tiles
// never create fill-in.  The wavefronts exchange data only through LDS (each
// tile load is consumed by the wavefront that issued it; factor stores are
// not read back), so the barriers do not wait for global memory, and all
// global accesses of a panel are issued in phase A, where they have B1 and
// B2 to complete.  Factors are stored as L (unit), D, U' like the other
// factorization kernels; the multipliers use the pivot's reciprocal.
// ---------------------------------------------------------------------------
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
    // two panel buffers, used by alternate panels (a panel's factors are
    // stored while the next panel is being factored):
    static constexpr int pn_off = 0;                      // Pn [16 KT][PS]: panel rows; after B2 rows >= 32 hold -L21
    static constexpr int ub_off = pn_off + 16 * KT * PS;  // UbT [NTR][PS]: U block, one column per row
    static constexpr int pv_off = ub_off + NTR * PS;      // pivots [32]
    static constexpr int rp_off = pv_off + 32;            // their reciprocals [32]
    static constexpr int buf_elems = rp_off + 32;
    static constexpr int LTS = 36;                        // L11T row stride (16-byte aligned rows)
    static constexpr int lt_off = 2 * buf_elems;          // L11 transposed [32][LTS]
    static constexpr size_t bytes = sizeof(T) * (size_t)(lt_off + 32 * LTS);
};

template <class T>
__device__ inline T buffer_load(__amdgpu_buffer_rsrc_t r, int off);
template <>
__device__ inline float buffer_load<float>(__amdgpu_buffer_rsrc_t r, int off)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <>
__device__ inline double buffer_load<double>(__amdgpu_buffer_rsrc_t r, int off)
{
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
    return __longlong_as_double((long long)(((unsigned long long)v[1] << 32) | v[0]));
}

template <class T>
__device__ inline void buffer_store(T v, __amdgpu_buffer_rsrc_t r, int off);
template <>
__device__ inline void buffer_store<float>(float v, __amdgpu_buffer_rsrc_t r, int off)
{
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, off, 0, 0);
}
template <>
__device__ inline void buffer_store<double>(double v, __amdgpu_buffer_rsrc_t r, int off)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    typedef unsigned int u2 __attribute__((ext_vector_type(2)));
    const u2 x = {(unsigned)b, (unsigned)(b >> 32)};
    __builtin_amdgcn_raw_buffer_store_b64(x, r, off, 0, 0);
}

// N values of a 16-byte aligned LDS row (float: 16-byte loads), from
// element 4 * (FROM / 4) on; the rest of v is left as it is
template <class T, int N>
__device__ inline void lds_row_from(const T* p, T (&v)[N], int from)
{
    if constexpr (sizeof(T) == 4) {
        const float4* q = reinterpret_cast<const float4*>(p);
#pragma unroll
        for (int i = 0; i < N / 4; ++i) {
            if (4 * i + 3 >= from) {
                const float4 x = q[i];
                v[4 * i] = x.x; v[4 * i + 1] = x.y; v[4 * i + 2] = x.z; v[4 * i + 3] = x.w;
            }
        }
    } else {
#pragma unroll
        for (int c = 0; c < N; ++c)
            if (c >= from) v[c] = p[c];
    }
}
template <int FROM, class T, int N>
__device__ inline void lds_row(const T* p, T (&v)[N])
{
    lds_row_from<T, N>(p, v, FROM);
}
template <int N>
__device__ inline void lds_store_row(float* p, const float (&v)[N])
{
    float4* q = reinterpret_cast<float4*>(p);
#pragma unroll
    for (int i = 0; i < N / 4; ++i) q[i] = make_float4(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]);
}
template <int N>
__device__ inline void lds_store_neg_row(float* p, const float (&v)[N])
{
    float4* q = reinterpret_cast<float4*>(p);
#pragma unroll
    for (int i = 0; i < N / 4; ++i) q[i] = make_float4(-v[4 * i], -v[4 * i + 1], -v[4 * i + 2], -v[4 * i + 3]);
}

__device__ inline float pivot_rcp(float v) { return __builtin_amdgcn_rcpf(v); }
__device__ inline double pivot_rcp(double v) { return 1.0 / v; }

// All ones when x < 0, else 0: one field extract of the sign bit.  Range
// predicates are combined as sign masks of differences (ANDed), so no
// per-element compare writes a scalar mask register.
__device__ inline int neg_mask(int x) { return __builtin_amdgcn_sbfe(x, 31, 1); }
// the sign of v flipped where m is all ones
__device__ inline float flip_sign(float v, int m) { return __int_as_float(__float_as_int(v) ^ (m & (int)0x80000000)); }
__device__ inline double flip_sign(double v, int m)
{
    const unsigned long long sb = (unsigned long long)((unsigned)m & 0x80000000u) << 32;
    return __longlong_as_double((long long)((unsigned long long)__double_as_longlong(v) ^ sb));
}
// a where m is all ones, else b
__device__ inline float pick(float a, float b, int m)
{
    return __int_as_float((__float_as_int(a) & m) | (__float_as_int(b) & ~m));
}
__device__ inline double pick(double a, double b, int m)
{
    const long long x = __double_as_longlong(a), y = __double_as_longlong(b);
    const unsigned lo = ((unsigned)x & (unsigned)m) | ((unsigned)y & ~(unsigned)m);
    const unsigned hi = ((unsigned)(x >> 32) & (unsigned)m) | ((unsigned)(y >> 32) & ~(unsigned)m);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <class T, int KT, int NW>
__global__ __launch_bounds__(64 * NW) void band_factor_resident_kernel(T* __restrict__ B0, T* __restrict__ B1,
                                                                   const T* __restrict__ Bin, int k_all,
                                                                   const FactorJob* __restrict__ jobs)
{
    using MF = Mfma16<T>;
    using acc_t = typename MF::acc_t;
    using LY = ResLds<T, KT>;
    constexpr int PS = LY::PS, NTR = LY::NTR, NS = KT * KT, SPW = (NS + NW - 1) / NW, HR = 32 + NTR;
    constexpr int NT = 64 * NW;         // threads
    constexpr int UOFF = 64 * (NW / 2); // first thread of the U12 solves in B2
    static_assert(NTR <= UOFF && UOFF + NTR <= NT, "B2 thread ranges");
    static_assert(SPW <= 32, "slot class masks are 32-bit");
    const FactorJob job = jobs[blockIdx.x];
    const int k = job.kb > 0 ? job.kb : k_all;
    T* __restrict__ base = (job.buf ? B1 : B0) + job.boff;
    const int64_t w = 2 * k + 1;
    const int64_t len = job.len;
    const int dir = job.dir;
    const int wm1 = 2 * k;   // next column, same row

    extern __shared__ unsigned char smem_raw[];
    T* const sm = reinterpret_cast<T*>(smem_raw);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // uniform: slot logic stays scalar

    // rem = len - j0 is the number of partition rows left from panel row j0;
    // entry (dr, dc) of the panel's window is stored iff this mask is set
    auto inb = [&](int dr, int dc, int rem) {
        return neg_mask((dr - rem) & (dc - rem) & (dr - dc - k - 1) & (dc - dr - k - 1));
    };
    // Tile loads are buffer loads over the partition's storage: every element
    // is one unconditional load, and entries outside the band or the
    // partition use an offset past the end of the buffer, which reads zero.
    // (A register only ever written by loads needs no wait for an earlier
    // load into it.)  Element (j0 + dr, j0 + dc) is at element offset
    // ob + dir * (dc * 2k + dr) with ob = rel0 + dir * (j0 * (2k + 1) + k).
    // The band is read from Bin when given (the LU and UL factorizations of
    // one partition then share one unfactored copy) and the factors are
    // written to the job's band; without Bin the factorization is in place.
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(base + job.start * w), 0, (int)(len * w * (int64_t)sizeof(T)), 0x00020000);
    const __amdgpu_buffer_rsrc_t rsrc_in = __builtin_amdgcn_make_buffer_rsrc(
        (void*)((Bin ? Bin + job.boff : base) + job.start * w), 0, (int)(len * w * (int64_t)sizeof(T)), 0x00020000);
    const int rel0 = dir > 0 ? 0 : (int)(len * w - 1);
    // Tiles are held transposed: accumulator register q of lane l holds
    // A(dr0 + l % 16, dc0 + row(l, q)), so for each register the 16 lanes of
    // a group read 16 consecutive rows of one column (one contiguous segment
    // per lane group).
    auto load_tile = [&](acc_t& t, int ob, int rem, int dr0, int dc0, int lr, int lane) {
        const int dr = dr0 + lr;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int dc = dc0 + MF::row(lane, q);
            const int m = inb(dr, dc, rem);
            const int off = (((ob + dir * (dc * wm1 + dr)) * (int)sizeof(T)) & m) | ~m;
            t[q] = buffer_load<T>(rsrc_in, off);
        }
    };
    // factors of panel pp from panel buffer bq: its 32 columns (rows 0 .. HR-1:
    // L, D, U' of the diagonal block and L21) and its 32 U' rows, coalesced
    // (values are gathered from LDS first and stored afterwards, so the LDS
    // reads overlap instead of each store waiting for its own read).  The
    // stores are unconditional buffer stores: entries outside the band get
    // an offset past the end of the buffer and are dropped, so every panel
    // issues the same number of stores (pp < 0: none land).
    // Stored by wavefronts 1 .. NW-1 (st = thread index - 64): wavefront 0
    // factors the next diagonal block meanwhile.
    constexpr int NTS = NT - 64;
    auto store_panel = [&](int64_t pp, const T* bq, int st) {
        constexpr int N1 = (32 * HR + NTS - 1) / NTS, N2 = (32 * NTR + NTS - 1) / NTS;
        const int64_t pj0 = 32 * max(pp, (int64_t)0);
        const int obq = rel0 + dir * (int)(pj0 * w + k);
        const int prem = pp < 0 ? 0 : (int)min(len - pj0, (int64_t)(1 << 30));
        const int pnb = min(32, prem);
        const T* qPn = bq + LY::pn_off;
        const T* qUb = bq + LY::ub_off;
        const T* qrp = bq + LY::rp_off;   // pivot reciprocals (0 past the partition's end)
        T v1[N1], v2[N2];
        int o1[N1], o2[N2];
#pragma unroll
        for (int i = 0; i < N1; ++i) {
            const int e = st + NTS * i;
            const int c = e / HR, r = e - c * HR;
            const int ok = neg_mask((e - 32 * HR) & (c - pnb) & (r - prem) & (r - c - k - 1) & (c - r - k - 1));
            const int lrow = ~neg_mask(r - 32);   // L21 rows, kept negated in LDS
            const int upper = neg_mask(r - c);    // U' of the diagonal block: scaled by 1 / pivot
            const T v = qPn[(r * PS + c) & ok];
            const T f = qrp[r & ~lrow & 31];
            v1[i] = flip_sign(v * pick(f, (T)1, upper), lrow);
            o1[i] = (((obq + dir * (c * wm1 + r)) * (int)sizeof(T)) & ok) | ~ok;
        }
#pragma unroll
        for (int i = 0; i < N2; ++i) {
            const int e = st + NTS * i;
            const int cc = e >> 5, m = e & 31;
            const int ok = neg_mask((e - 32 * NTR) & (m - pnb) & (32 + cc - prem) & (32 + cc - m - k - 1));
            v2[i] = qUb[(cc * PS + m) & ok] * qrp[m];
            o2[i] = (((obq + dir * ((32 + cc) * wm1 + m)) * (int)sizeof(T)) & ok) | ~ok;
        }
#pragma unroll
        for (int i = 0; i < N1; ++i) buffer_store<T>(v1[i], rsrc, o1[i]);
#pragma unroll
        for (int i = 0; i < N2; ++i) buffer_store<T>(v2[i], rsrc, o2[i]);
    };

    acc_t acc[SPW];
    {
        const int lane = threadIdx.x & 63;
        const int ob = rel0 + dir * k;
        const int rem = (int)min(len, (int64_t)(1 << 30));
#pragma unroll
        for (int j = 0; j < SPW; ++j) {
            const int s = wave + NW * j;
            if (s < NS) load_tile(acc[j], ob, rem, 16 * (s / KT), 16 * (s % KT), lane & 15, lane);
        }
    }

    const int64_t npan = (len + 31) / 32;
    for (int64_t p = 0; p < npan; ++p) {
        // lane-derived values are re-materialised every panel (keeps the
        // per-lane predicates of the unrolled loops out of the registers)
        int tid = threadIdx.x, zoff = 0;
        asm volatile("" : "+v"(tid), "+v"(zoff));
        const int lane = tid & 63, lr = lane & 15, lg = lane >> 4;
        const int cur = (int)(p & 1);
        // LDS addresses from an opaque zero fold into the ds_* immediate
        // offsets instead of living in scalar registers across the loop
        T* bcur = sm + zoff + cur * LY::buf_elems;
        T* Pn = bcur + LY::pn_off;
        T* UbT = bcur + LY::ub_off;
        T* pivs = bcur + LY::pv_off;
        T* rpiv = bcur + LY::rp_off;
        T* L11T = sm + zoff + LY::lt_off;
        const int64_t j0 = 32 * p;
        const int rem = (int)min(len - j0, (int64_t)(1 << 30));
        const int nb = min(32, rem);
        const int sh = (int)((2 * p) % KT);
        const int obp = rel0 + dir * (int)(j0 * w + k);
        // Slot bookkeeping, lane-parallel: lane j holds the rotated tile
        // coordinates (rho, gam) of this wavefront's slot j (s = wave + NW j,
        // clamped to NS - 1 past the last slot), and three ballots give the
        // slot classes: panel column (gam < 2), U row (rho < 2 <= gam) and
        // trailing (both >= 2).  Each unrolled slot step then costs a mask bit
        // test and one readlane instead of its own scalar index arithmetic.
        unsigned pm, um, tm, dm;
        int rg;
        {
            const int sl = min(wave + NW * lane, NS - 1);
            int rl = sl / KT - sh, gl = sl % KT - sh;
            rl += KT & neg_mask(rl);
            gl += KT & neg_mask(gl);
            const bool okl = lane < SPW && wave + NW * lane < NS;
            pm = (unsigned)__builtin_amdgcn_ballot_w64(okl && gl < 2);
            um = (unsigned)__builtin_amdgcn_ballot_w64(okl && gl >= 2 && rl < 2);
            tm = (unsigned)__builtin_amdgcn_ballot_w64(okl && gl >= 2 && rl >= 2);
            dm = (unsigned)__builtin_amdgcn_ballot_w64(okl && gl < 2 && rl < 2);   // diagonal block
            rg = rl | (gl << 8);
        }
        // rho and gam of slot j (j a compile-time constant)
        auto slot_rho = [&](int j) { return __builtin_amdgcn_readlane(rg, j) & 0xff; };
        auto slot_gam = [&](int j) { return __builtin_amdgcn_readlane(rg, j) >> 8; };

        // Every load and store of the previous panel has long completed (B2
        // and C ran since); saying so with an explicit wait clears the
        // compiler's record of loads pending across the loop back-edge, which
        // would otherwise make each slot below wait for the loads just issued
        // by the slots before it (vmcnt bits [3:0] and [15:14] = 0, expcnt and
        // lgkmcnt at their maxima).
        __builtin_amdgcn_s_waitcnt((0x7 << 4) | (0xF << 8));
        // A0. the diagonal block's four tiles to LDS: B1 needs nothing else,
        //     so wavefront 0 factors it while the others finish phase A.
#pragma unroll
        for (int j = 0; j < SPW; ++j) {
            if ((dm >> j) & 1) {
                const int rho = slot_rho(j), gam = slot_gam(j);
#pragma unroll
                for (int q = 0; q < 4; ++q) Pn[(16 * rho + lr) * PS + 16 * gam + MF::row(lane, q)] = acc[j][q];
            }
        }
        lds_barrier();

        // B1. LU of the diagonal block (wavefront 0), row r in lane r (lanes
        //     32..63 mirror).  It writes rows 0..31 of Pn, the pivots and L11^T
        //     only, which the rest of phase A does not touch.
        //     Straight-line code (no branch per column), so the scheduler can
        //     run ahead: a lane that must not change gets multiplier 0, and
        //     columns past the partition's end (last panel) get reciprocal 0,
        //     which makes their steps exact no-ops.
        if (wave == 0) {
            const int r = lane & 31;
            T row[32];
            T pr = (T)0;   // this lane's pivot
#pragma unroll
            for (int c = 0; c < 32; ++c) row[c] = Pn[r * PS + c];
#pragma unroll
            for (int jj = 0; jj < 32; ++jj) {
                T piv = lane_bcast(row[jj], jj);
                if (job.boost && (jj + 1 < rem || job.boost_last)) piv = boost_pivot(piv);
                const T rp = jj < nb ? pivot_rcp(piv) : (T)0;
                const T l = r > jj ? row[jj] * rp : (T)0;
                pr = r == jj ? piv : pr;
#pragma unroll
                for (int c = jj + 1; c < 32; ++c) row[c] -= l * lane_bcast(row[c], jj);   // pivot row, before any update
                row[jj] = r > jj ? l : (r == jj ? piv : row[jj]);
            }
            if (lane < 32) {
                pivs[r] = pr;
                rpiv[r] = r < nb ? pivot_rcp(pr) : (T)0;
#pragma unroll
                for (int c = 0; c < 32; ++c) {
                    Pn[r * PS + c] = row[c];
                    if (c < r) L11T[c * LY::LTS + r] = row[c];
                }
            }
        }
        // A. the other leaving tiles to LDS, then their slots load the entering
        //    tiles, then the previous panel's factors go out: every global
        //    access of the panel is issued here, so loads and stores drain
        //    during B2 and C before the next panel needs them.  (All copies to
        //    LDS come first: reading a slot that may still be loading waits for
        //    every load issued before it.)
#pragma unroll
        for (int j = 0; j < SPW; ++j) {
            if (((pm & ~dm) >> j) & 1) {
                const int rho = slot_rho(j), gam = slot_gam(j);
#pragma unroll
                for (int q = 0; q < 4; ++q) Pn[(16 * rho + lr) * PS + 16 * gam + MF::row(lane, q)] = acc[j][q];
            } else if ((um >> j) & 1) {
                const int rho = slot_rho(j), gam = slot_gam(j);
#pragma unroll
                for (int q = 0; q < 4; ++q) UbT[(16 * (gam - 2) + MF::row(lane, q)) * PS + 16 * rho + lr] = acc[j][q];
            }
        }
#pragma unroll
        for (int j = 0; j < SPW; ++j) {
            if (((pm | um) >> j) & 1) {
                const int rho = slot_rho(j), gam = slot_gam(j);
                const int rn = rho < 2 ? rho + KT - 2 : rho - 2;
                const int cn = gam < 2 ? gam + KT - 2 : gam - 2;
                load_tile(acc[j], obp, rem, 32 + 16 * rn, 32 + 16 * cn, lr, lane);
            }
        }
        if (wave > 0) store_panel(p - 1, sm + zoff + (1 - cur) * LY::buf_elems, tid - 64);
        lds_barrier();

        // B2. L21 rows (stored negated: the trailing update adds) and U12 columns
        if constexpr (sizeof(T) == 4) {
            // FP32, by halves: columns 0..15 of L21 = A21 U11^-1 (rows 0..15 of
            // U12 = L11^-1 A12) by substitution; their contribution to the
            // other half (A21_b -= L21_a U11_ab, A12_b -= L11_ba U12_a) on the
            // matrix cores; then the other half by substitution.  The U11
            // (L11^T) rows come in 16-byte loads one step ahead of their use.
            if (tid < NTR) {
                const int t = tid;
                T a[16], cur[16], nxt[16];
                lds_row<0>(Pn + (32 + t) * PS, a);
                lds_row<1>(Pn, cur);
#pragma unroll
                for (int jj = 0; jj < 16; ++jj) {   // rpiv = 0 past the partition's end
                    if (jj < 15) lds_row_from<T>(Pn + (jj + 1) * PS, nxt, jj + 2);
                    const T l = a[jj] * rpiv[jj];
                    a[jj] = l;
#pragma unroll
                    for (int c = jj + 1; c < 16; ++c) a[c] -= l * cur[c];
#pragma unroll
                    for (int c = 0; c < 16; ++c) cur[c] = nxt[c];
                }
                lds_store_neg_row(Pn + (32 + t) * PS, a);
            } else if (tid >= UOFF && tid < UOFF + NTR) {
                const int cc = tid - UOFF;
                T u[16], cur[16], nxt[16];
                lds_row<0>(UbT + cc * PS, u);
                lds_row<1>(L11T, cur);
#pragma unroll
                for (int jj = 0; jj < 15; ++jj) {   // u[jj] = 0 past the partition's end
                    if (jj < 14) lds_row_from<T>(L11T + (jj + 1) * LY::LTS, nxt, jj + 2);
                    const T uj = u[jj];
#pragma unroll
                    for (int m = jj + 1; m < 16; ++m) u[m] -= cur[m] * uj;
#pragma unroll
                    for (int m = 0; m < 16; ++m) cur[m] = nxt[m];
                }
                lds_store_row(UbT + cc * PS, u);
            }
            lds_barrier();
            {
                // 16 x 16 tiles: KT-2 of L21's rows (A = the stored -L21_a,
                // B = U11 rows 0..15, columns 16..31) and KT-2 of U12's
                // columns (A = -U12_a^T, B = L11^T rows 0..15, columns 16..31)
                const int g = lane >> 4, c16 = lane & 15;
                for (int f = wave; f < 2 * (KT - 2); f += NW) {
                    const bool lside = f < KT - 2;
                    const int r0 = lside ? 32 + 16 * f : 16 * (f - (KT - 2));
                    T* const rows = lside ? Pn : UbT;
                    const T* const bm = lside ? Pn : L11T;
                    const int bs = lside ? PS : LY::LTS;
                    acc_t acc;
#pragma unroll
                    for (int q = 0; q < 4; ++q) acc[q] = rows[(r0 + MF::row(lane, q)) * PS + 16 + c16];
#pragma unroll
                    for (int st = 0; st < 4; ++st) {
                        const T av = rows[(r0 + c16) * PS + 4 * st + g];
                        const T bv = bm[(4 * st + g) * bs + 16 + c16];
                        acc = MF::mma(lside ? av : -av, bv, acc);
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q) rows[(r0 + MF::row(lane, q)) * PS + 16 + c16] = acc[q];
                }
            }
            lds_barrier();
            if (tid < NTR) {
                const int t = tid;
                T a[16], cur[16], nxt[16];   // columns 16..31
                lds_row<0>(Pn + (32 + t) * PS + 16, a);
                lds_row<1>(Pn + 16 * PS + 16, cur);
#pragma unroll
                for (int jj = 0; jj < 16; ++jj) {
                    if (jj < 15) lds_row_from<T>(Pn + (17 + jj) * PS + 16, nxt, jj + 2);
                    const T l = a[jj] * rpiv[16 + jj];
                    a[jj] = l;
#pragma unroll
                    for (int c = jj + 1; c < 16; ++c) a[c] -= l * cur[c];
#pragma unroll
                    for (int c = 0; c < 16; ++c) cur[c] = nxt[c];
                }
                lds_store_neg_row(Pn + (32 + t) * PS + 16, a);
            } else if (tid >= UOFF && tid < UOFF + NTR) {
                const int cc = tid - UOFF;
                T u[16], cur[16], nxt[16];   // rows 16..31
                lds_row<0>(UbT + cc * PS + 16, u);
                lds_row<1>(L11T + 16 * LY::LTS + 16, cur);
#pragma unroll
                for (int jj = 0; jj < 15; ++jj) {
                    if (jj < 14) lds_row_from<T>(L11T + (17 + jj) * LY::LTS + 16, nxt, jj + 2);
                    const T uj = u[jj];
#pragma unroll
                    for (int m = jj + 1; m < 16; ++m) u[m] -= cur[m] * uj;
#pragma unroll
                    for (int m = 0; m < 16; ++m) cur[m] = nxt[m];
                }
                lds_store_row(UbT + cc * PS + 16, u);
            }
        } else if (tid < NTR) {
            const int t = tid;
            T a[32];
#pragma unroll
            for (int c = 0; c < 32; ++c) a[c] = Pn[(32 + t) * PS + c];
#pragma unroll
            for (int jj = 0; jj < 32; ++jj) {   // rpiv = 0 past the partition's end
                const T l = a[jj] * rpiv[jj];
                a[jj] = l;
#pragma unroll
                for (int c = jj + 1; c < 32; ++c) a[c] -= l * Pn[jj * PS + c];
            }
#pragma unroll
            for (int c = 0; c < 32; ++c) Pn[(32 + t) * PS + c] = -a[c];
        } else if (tid >= UOFF && tid < UOFF + NTR) {
            const int cc = tid - UOFF;
            T u[32];
#pragma unroll
            for (int m = 0; m < 32; ++m) u[m] = UbT[cc * PS + m];
#pragma unroll
            for (int jj = 0; jj < 31; ++jj) {   // u[jj] = 0 past the partition's end
                const T uj = u[jj];
#pragma unroll
                for (int m = jj + 1; m < 32; ++m) u[m] -= L11T[jj * LY::LTS + m] * uj;
            }
#pragma unroll
            for (int m = 0; m < 32; ++m) UbT[cc * PS + m] = u[m];
        }
        lds_barrier();

        // C. trailing update on the matrix cores.  The next panel writes the
        //    other buffer, so no barrier is needed after this phase.
        if constexpr (sizeof(T) == 4) {
            // FP32: each slot's operands are two 16-byte rows (lane group lg
            // takes k = 8 lg .. 8 lg + 7), fetched one slot ahead of the slot
            // being multiplied.  The fetches are unconditional: for a slot
            // that is not updated (panel rows or columns; s >= NS reads slot
            // NS - 1) the addresses still fall inside this panel buffer (rows
            // 16 rho + lr < 16 KT of Pn, or rows of UbT at most 32 before it,
            // which are rows of Pn), and the values go unused.
            auto fetch = [&](int j, float4 (&o)[4]) {
                const int rho = slot_rho(j), gam = slot_gam(j);
                const float4* ap = reinterpret_cast<const float4*>(Pn + (16 * rho + lr) * PS + 8 * lg);
                const float4* bp = reinterpret_cast<const float4*>(UbT + (16 * (gam - 2) + lr) * PS + 8 * lg);
                o[0] = ap[0]; o[1] = ap[1]; o[2] = bp[0]; o[3] = bp[1];
            };
            float4 op[4], nx[4];
            fetch(0, op);
#pragma unroll
            for (int j = 0; j < SPW; ++j) {
                if (j + 1 < SPW) fetch(j + 1, nx);
                if ((tm >> j) & 1) {
                    const float av[8] = {op[0].x, op[0].y, op[0].z, op[0].w, op[1].x, op[1].y, op[1].z, op[1].w};
                    const float bv[8] = {op[2].x, op[2].y, op[2].z, op[2].w, op[3].x, op[3].y, op[3].z, op[3].w};
#pragma unroll
                    // transposed tiles: acc^T += U12^T (-L21)^T
                    for (int q = 0; q < 8; ++q) acc[j] = MF::mma(bv[q], av[q], acc[j]);
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) op[i] = nx[i];
            }
        } else {
#pragma unroll
            for (int j = 0; j < SPW; ++j) {
                if ((tm >> j) & 1) {
                    const int rho = slot_rho(j), gam = slot_gam(j);
                    const T* ap = Pn + (16 * rho + lr) * PS + 8 * lg;
                    const T* bp = UbT + (16 * (gam - 2) + lr) * PS + 8 * lg;
                    T av[8], bv[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q) { av[q] = ap[q]; bv[q] = bp[q]; }
#pragma unroll
                    for (int q = 0; q < 8; ++q) acc[j] = MF::mma(bv[q], av[q], acc[j]);
                }
            }
        }
    }
    if (npan > 0 && wave > 0) store_panel(npan - 1, sm + ((npan - 1) & 1) * LY::buf_elems, (int)threadIdx.x - 64);
}

// Wavefronts per workgroup: the window's tiles are spread over them, so more
// wavefronts leave each more registers for the panel phases (and more
// latency hiding), as long as the accumulators still fit.
template <class T, int KT>
constexpr int resident_waves() { return 8; }

template <class T, int KT>
void launch_factor_resident(T* B0, T* B1, const T* Bin, int k, const FactorJob* jobs, int njobs, hipStream_t st)
{
    constexpr int NW = resident_waves<T, KT>();
    const size_t lds = ResLds<T, KT>::bytes;
    SAP_HIP_CHECK(hipFuncSetAttribute((const void*)band_factor_resident_kernel<T, KT, NW>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((band_factor_resident_kernel<T, KT, NW>), dim3(njobs), dim3(64 * NW), lds, st, B0, B1, Bin,
                       k, jobs);
    SAP_LAUNCH_CHECK();
}

// Largest window (in 16-row tiles) kept in registers: FP32 up to k = 224,
// FP64 (twice the registers and LDS per tile) up to k = 112.
template <class T>
constexpr int resident_max_kt() { return sizeof(T) == 4 ? 16 : 9; }

template <class T, int KT>
bool dispatch_resident(int kt, T* B0, T* B1, const T* Bin, int k, const FactorJob* jobs, int njobs,
                       hipStream_t st)
{
    if constexpr (KT > resident_max_kt<T>()) {
        return false;
    } else {
        if (kt == KT) { launch_factor_resident<T, KT>(B0, B1, Bin, k, jobs, njobs, st); return true; }
        return dispatch_resident<T, KT + 1>(kt, B0, B1, Bin, k, jobs, njobs, st);
    }
}

}  // namespace

template <class T>
bool band_factor_resident_fits(int k, int64_t max_len)
{
    // buffer offsets are 32-bit: a partition's storage must stay below 2 GiB
    if (max_len * (2 * (int64_t)k + 1) * (int64_t)sizeof(T) >= ((int64_t)1 << 31)) return false;
    return (k + 32 + 15) / 16 <= resident_max_kt<T>();
}

template <class T>
bool launch_band_factor_resident(T* B0, T* B1, const T* Bin, int k, const FactorJob* d_jobs, int njobs,
                                 int64_t max_len, hipStream_t st)
{
    if (!band_factor_resident_fits<T>(k, max_len)) return false;
    return dispatch_resident<T, 3>((k + 32 + 15) / 16, B0, B1, Bin, k, d_jobs, njobs, st);
}

template bool band_factor_resident_fits<float>(int, int64_t);
template bool band_factor_resident_fits<double>(int, int64_t);
template bool launch_band_factor_resident<float>(float*, float*, const float*, int, const FactorJob*, int, int64_t,
                                                 hipStream_t);
template bool launch_band_factor_resident<double>(double*, double*, const double*, int, const FactorJob*, int,
                                                  int64_t, hipStream_t);

}  // namespace sap_amd
