// Preconditioner apply: z = M^{-1} v (Precond::solve / getSRev,
// sap/precond.h:1392-1528).
//
// SaP-C runs three launches:
//   pass 1  interface values of g = A_D^{-1} v.  The reduced system needs
//           only the bottom k rows of A_i^{-1} v_i and the top k rows of
//           A_{i+1}^{-1} v_{i+1}: the first come from a full L sweep of the
//           LU factors plus a k-row U' sweep, the second from a full U' sweep
//           of the UL factors plus a k-row L' sweep.  The two are
//           independent, so every partition boundary gets two workgroups.
//   iface   reduced solve per boundary with S_i^{-1} and purification
//           (partFullFwdSweep/partFullBckSweep + purifyRHS,
//           sap/precond.h:5418-5530), producing corrections of the rhs rows
//           next to each boundary.
//   pass 2  full L, D^{-1}, U' sweep of (v - corrections) for every
//           partition (partBandedFwdSweep/partBandedBckSweep); the last
//           partition uses its UL factors when factMethod = LU_UL.
// SaP-D (or P = 1) runs pass 2 alone.
//
// Triangular sweeps: one workgroup per (partition, direction) walks the rows
// in blocks; wavefront 0 solves each block's unit-lower diagonal block with
// a lane-broadcast chain, the other wavefronts apply the block to the next k
// rows, and the pending values live in an LDS ring.  Three variants:
//   MODE 2 (k <= 384 + BS while the staging fits, the default; above
//          192 + BS each far-row lane takes two rows): multipliers and
//          right-hand side
//          staged HBM -> LDS by producer wavefronts several blocks ahead;
//   MODE 1 (same k range): multipliers prefetched into registers one block
//          ahead;
//   MODE 0 (any k): the plain blocked sweep with blocks of 64.
#include "sap_internal.hpp"

#include <algorithm>
#include <mutex>
#include <unordered_map>

namespace sap_amd {

namespace {

constexpr int NBS = 64;

template <class T>
struct BandView {
    const T* base;
    int64_t origin;
    int64_t w;
    int k;
    int dir;
    __device__ inline T at(int64_t r, int64_t c) const { return base[origin + dir * (c * w + k + r - c)]; }
};

template <class T>
__device__ inline BandView<T> make_view(const T* B, int64_t s, int64_t len, int k, int dir)
{
    BandView<T> v;
    v.base = B;
    v.w = 2 * k + 1;
    v.k = k;
    v.dir = dir;
    v.origin = dir > 0 ? s * v.w : (s + len) * v.w - 1;
    return v;
}

// v if bit j of m is set, else +0 (bitwise: a signed one-bit field extract
// and an AND, so no compare and no scalar mask register per element)
__device__ inline float keep_if(float v, unsigned m, int j)
{
    return __int_as_float(__float_as_int(v) & __builtin_amdgcn_sbfe((int)m, j, 1));
}
__device__ inline double keep_if(double v, unsigned m, int j)
{
    // the 32-bit field mask applied to each half (hipcc drops the sign
    // extension of the field extract when it is widened to 64 bits first)
    const unsigned sm = (unsigned)__builtin_amdgcn_sbfe((int)m, j, 1);
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)b & sm, hi = (unsigned)(b >> 32) & sm;
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// bits 0 .. n-1 set (n clamped to [0, 32])
__device__ inline unsigned bits_below(int n)
{
    return n <= 0 ? 0u : n >= 32 ? ~0u : (1u << n) - 1u;
}

__device__ inline float bcast(float v, int j)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}
__device__ inline double bcast(double v, int j)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), j);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), j);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// y = L^{-1} b over local rows [0, len) of a view (L unit lower, band k).
// in(r) gives b_r; out(r, y_r) receives each final value.
template <class T, class In, class Out>
__device__ void lower_sweep(const BandView<T>& V, int64_t len, In in, Out out, T* pend, int ring_mask, T* xblk)
{
    const int k = V.k;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int64_t init = min(len, (int64_t)NBS + k);
    for (int64_t r = tid; r < init; r += blockDim.x) pend[r & ring_mask] = in(r);
    __syncthreads();
    for (int64_t c0 = 0; c0 < len; c0 += NBS) {
        const int nb = (int)min((int64_t)NBS, len - c0);
        if (wave == 0) {
            T Ld[NBS];
#pragma unroll
            for (int j = 0; j < NBS; ++j)
                Ld[j] = (j < lane && lane < nb && lane - j <= k) ? V.at(c0 + lane, c0 + j) : (T)0;
            T x = lane < nb ? pend[(c0 + lane) & ring_mask] : (T)0;
#pragma unroll
            for (int j = 0; j < NBS - 1; ++j) {
                const T xj = bcast(x, j);
                x -= Ld[j] * xj;
            }
            if (lane < nb) {
                xblk[lane] = x;
                out(c0 + lane, x);
            }
        }
        __syncthreads();
        for (int t = tid; t < k; t += blockDim.x) {
            const int64_t r = c0 + nb + t;
            if (r >= len) break;
            const int jlo = max(0, nb + t - k);
            T acc = (T)0;
            for (int j = jlo; j < nb; ++j) acc += V.at(r, c0 + j) * xblk[j];
            pend[r & ring_mask] -= acc;
        }
        {
            const int64_t lo = c0 + nb + k;
            const int64_t hi = min(len, c0 + nb + k + NBS);
            for (int64_t r = lo + tid; r < hi; r += blockDim.x) pend[r & ring_mask] = in(r);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Pipelined variant for k <= 192 + BS (BS = 32 rows per block for FP32, 16
// for FP64, so two blocks of multipliers fit in registers).  Per block b:
//   wave 0    solves the BS x BS diagonal block with the lane-broadcast chain
//             and immediately applies the block to the next BS rows (the
//             sub-diagonal BS x BS block), so the chain never waits for the
//             other waves;
//   waves 1-3 apply the previous block b-1 to its remaining k - BS rows
//             (rows of block b+1 go to a separate accumulator, later ones to
//             the LDS ring) and stream the next input rows into the ring.
// Every wave loads its multipliers for the next block into registers one
// block ahead, so the loads overlap the chain.  One barrier per block.
// ---------------------------------------------------------------------------
template <class T>
struct SweepRegs {
    static constexpr int BS = sizeof(T) == 4 ? 32 : 16;
    T a[BS];     // wave 0: diagonal-block multipliers of its row; waves 1-3: its far row
    T b[BS];     // wave 0: sub-diagonal block row
    T d;         // wave 0: pivot of its row (for the D^{-1} scaling)
};

template <class T>
constexpr int sweep_bs() { return SweepRegs<T>::BS; }

template <class T, class In, class Out>
__device__ __forceinline__ void lower_sweep_fast(const BandView<T>& V, int64_t len, In in, Out out, bool scale,
                                                 T* pend, int ring_mask, T* xblk2, T* acc2)
{
    constexpr int BS = sweep_bs<T>();
    const int k = V.k;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int t = tid - 64;                   // far-row index of waves 1-3
    const int nfar = k - BS;                  // far rows per block

    // rows below c0 + H can be touched while block b is processed; the input
    // rows [c0 + H, c0 + H + BS) are streamed in during step b.
    const int64_t H = BS + max(BS, k);
    const int64_t init = min(len, H);
    for (int64_t r = tid; r < init; r += blockDim.x) pend[r & ring_mask] = in(r);
    if (tid < BS) { acc2[tid] = (T)0; acc2[BS + tid] = (T)0; }

    // Multiplier (r, c0 + j) of the view sits at colp + j * cstep + (r - c0) * dir
    // with colp = V.base + V.origin + dir * (c0 * w + k) and cstep = dir * (w - 1):
    // one 64-bit lane pointer per row, uniform steps per column.
    const int64_t cstep = (int64_t)V.dir * (V.w - 1);
    auto load = [&](SweepRegs<T>& R, int64_t c0) {
        const int nb = (int)min((int64_t)BS, len - c0);
        const T* colp = V.base + V.origin + (int64_t)V.dir * (c0 * V.w + k);
        if (wave == 0) {
            const int i = lane;
            const int64_t r = c0 + i;
            const bool rv = i < BS && r < len;
            const T* p = colp + (int64_t)V.dir * i;
#pragma unroll
            for (int j = 0; j < BS; ++j) R.a[j] = (rv && j < i && i - j <= k) ? p[j * cstep] : (T)0;
            const bool sv = i < BS && r + BS < len;
            const T* q = p + (int64_t)V.dir * BS;
#pragma unroll
            for (int j = 0; j < BS; ++j) R.b[j] = (sv && j < nb && BS + i - j <= k) ? q[j * cstep] : (T)0;
            R.d = (rv && scale) ? p[i * cstep] : (T)1;
        } else {
            const int64_t r = c0 + 2 * BS + t;
            const bool fv = t < nfar && r < len;
            const T* p = colp + (int64_t)V.dir * (2 * BS + t);
#pragma unroll
            for (int j = 0; j < BS; ++j) R.a[j] = (fv && j < nb && 2 * BS + t - j <= k) ? p[j * cstep] : (T)0;
        }
    };

    // one pipeline step: block b (wave 0) and block b-1 (waves 1-3)
    auto step = [&](int64_t b, SweepRegs<T>& cur, SweepRegs<T>& nxt) {
        const int64_t c0 = b * BS;
        const int nb = (int)min((int64_t)BS, len - c0);
        const int par = (int)(b & 1);
        if (wave == 0) {
            if (c0 + BS < len) load(nxt, c0 + BS);
            T x = (lane < nb) ? pend[(c0 + lane) & ring_mask] - acc2[par * BS + lane] : (T)0;
#pragma unroll
            for (int j = 0; j < BS - 1; ++j) {
                const T xj = bcast(x, j);
                x -= cur.a[j] * xj;
            }
            if (lane < nb) {
                xblk2[par * BS + lane] = x;
                out(c0 + lane, scale ? x / cur.d : x);
            }
            T s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
            for (int j = 0; j < BS; j += 4) {
                s0 += cur.b[j] * bcast(x, j);
                s1 += cur.b[j + 1] * bcast(x, j + 1);
                s2 += cur.b[j + 2] * bcast(x, j + 2);
                s3 += cur.b[j + 3] * bcast(x, j + 3);
            }
            const int64_t rn = c0 + BS + lane;
            if (lane < BS && rn < len) pend[rn & ring_mask] -= (s0 + s1) + (s2 + s3);
        } else {
            // far rows of block b-1 (multipliers loaded one step earlier)
            if (b >= 1) {
                const int64_t cp = c0 - BS;
                const T* xp = xblk2 + (1 - par) * BS;
                T s0 = 0, s1 = 0;
#pragma unroll
                for (int j = 0; j < BS; j += 2) {
                    s0 += cur.a[j] * xp[j];
                    s1 += cur.a[j + 1] * xp[j + 1];
                }
                const T acc = s0 + s1;
                const int64_t r = cp + 2 * BS + t;
                if (t < BS) acc2[(1 - par) * BS + t] = (t < nfar && r < len) ? acc : (T)0;
                else if (t < nfar && r < len) pend[r & ring_mask] -= acc;
            } else if (t < BS) {
                acc2[BS + t] = (T)0;
            }
            if (c0 < len) load(nxt, c0);
            // stream the input rows first touched in a later step
            const int64_t rr = c0 + H + t;
            if (t < BS && rr < len) pend[rr & ring_mask] = in(rr);
        }
        __syncthreads();
    };

    SweepRegs<T> R0, R1;
    if (wave == 0) load(R0, 0);
    __syncthreads();
    const int64_t nblk = (len + BS - 1) / BS;
    // waves 1-3 lag one block behind wave 0: one extra step drains them
    for (int64_t b = 0; b <= nblk; b += 2) {
        step(b, R0, R1);
        if (b + 1 <= nblk) step(b + 1, R1, R0);
    }
}

// ---------------------------------------------------------------------------
// LDS-prefetched variant (k <= 192 + BS, 512 threads).  The entries a block
// of BS columns needs (column c: its diagonal and the k entries below it,
// contiguous in the band storage) are copied HBM -> LDS by four producer
// wavefronts with global_load_lds_dwordx4, as 16-byte chunks covering each
// column segment, up to NBUF - 2 blocks ahead of the solve.  The right-hand
// side rows stream into an LDS staging ring the same way, SW_DIN steps
// ahead.  No step waits for HBM; the sweep runs at the pace of its
// dependency chain.  In step b:
//   wave 0     diagonal-block chain and sub-diagonal block of block b (as in
//              lower_sweep_fast), multipliers from buffer b;
//   waves 1-3  far rows of block b-1 from buffer b-1;
//   wave 4     moves the staged right-hand side rows into the pending ring
//              and streams the next ones;
//   waves 5-7  producers, loading block b + NBUF - 2 into the buffer of
//              block b-2.
// Each producer (and wave 4 for the staging ring) waits with an explicit
// vmcnt count for its own copies before the barrier that publishes them.
// ---------------------------------------------------------------------------
// Wavefront roles: 0 the chain, 1-3 far rows, SW_STREAMER the right-hand
// side stream, SW_PROD0 .. SW_PROD0 + SW_PROD - 1 producers.  Wavefronts w
// and w + 4 share a SIMD, so the chain's partner is the light streamer.
constexpr int SW_PROD = 3;           // producer wavefronts
constexpr int SW_STREAMER = 4;
constexpr int SW_PROD0 = 5;
// load instructions per producer per block: up to 12 in FP32 (k <= 284),
// 16 in FP64 (k <= 382, the wide FP64 bands; FP32 keeps its per-lane
// offset tables small)
template <class T>
constexpr int sw_ni_max() { return sizeof(T) == 4 ? 12 : 16; }
constexpr int SW_FAR = 192;          // far rows of waves 1-3 per pass (one row per lane)
constexpr int SW_THREADS = 512;
constexpr int SW_DIN = 8;            // right-hand side steps in flight
constexpr int SW_NIN = SW_DIN + 1;   // staging slots of 64 dwords

// s_waitcnt takes an immediate: dispatch on a wave-uniform count (<= 63).
// gfx9 encoding: vmcnt bits [3:0] and [15:14]; expcnt [6:4] and lgkmcnt
// [11:8] left at their maxima (no wait).
template <int N>
__device__ __forceinline__ void wait_vmcnt_imm()
{
    __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

template <int N>
__device__ __forceinline__ void wait_vmcnt_le(int n)
{
    if constexpr (N == 0) {
        wait_vmcnt_imm<0>();
    } else {
        if (n >= N) wait_vmcnt_imm<N>();
        else wait_vmcnt_le<N - 1>(n);
    }
}

__device__ __forceinline__ void wait_vmcnt(int n) { wait_vmcnt_le<63>(n); }

// Chunk geometry of one view for the LDS-prefetched sweep.
template <class T>
struct SweepChunks {
    static constexpr int E = 16 / sizeof(T);   // elements per 16-byte chunk
    int nch;      // chunks per column (covers any alignment of a (k+1)-segment)
    int segl;     // elements per column in LDS = nch * E
    int ncb;      // chunks per block, padded to a whole number of producer rounds
    int ni;       // load instructions per producer wavefront per block
    __host__ __device__ SweepChunks(int k, int bs)
    {
        nch = (k + E - 1) / E + 1;
        segl = nch * E;
        const int round = 64 * SW_PROD;
        ncb = ((bs * nch + round - 1) / round) * round;
        ni = ncb / round;
    }
    __host__ __device__ size_t buf_elems() const { return (size_t)ncb * E; }
};

// Right-hand side of a sweep: view row r reads src[s0 + sdir * r] (RawT is
// double for the input vector, T for the intermediate), and fin(r, raw)
// turns it into the preconditioner type.
template <class T, class RawT, class Fin, class Out>
__device__ __forceinline__ void lower_sweep_lds(const BandView<T>& V, int64_t len, const RawT* __restrict__ src,
                                                int64_t s0, int sdir, Fin fin, Out out, bool scale, T* pend,
                                                int ring_mask, T* xblk2, T* acc2, T* dyn, int nbuf)
{
    constexpr int BS = sweep_bs<T>();
    constexpr int E = SweepChunks<T>::E;
    constexpr int WPR = sizeof(RawT) / 4;     // dwords per right-hand side row
    const SweepChunks<T> G(V.k, BS);
    const int k = V.k;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nfar = k - BS;                  // far rows per block
    const int64_t nblk = (len + BS - 1) / BS;
    const size_t bufe = G.buf_elems();
    unsigned* instage = reinterpret_cast<unsigned*>(dyn);
    T* bufs = dyn + SW_NIN * 64 * 4 / sizeof(T);

    const int64_t H = BS + max(BS, k);
    const int64_t init = min(len, H);
    for (int64_t r = threadIdx.x; r < init; r += blockDim.x) pend[r & ring_mask] = fin(r, src[s0 + sdir * r]);
    if (threadIdx.x < BS) { acc2[threadIdx.x] = (T)0; acc2[BS + threadIdx.x] = (T)0; }

    // Segment of view column c: elements q = 0..k at V.origin + dir * (c*w + k + q);
    // lo = its lowest address, lo & ~(E-1) the first chunk.  In LDS, element q of
    // block column j is at j*segl + off_j + dir*q with off_j = (lo_j & (E-1)) +
    // (dir < 0 ? k : 0); lo_j = lo_0 + j * dir * w.
    auto seg_lo = [&](int64_t c) -> int64_t {
        const int64_t a = V.origin + (int64_t)V.dir * (c * V.w + k);
        return V.dir > 0 ? a : a - k;
    };
    const int dwm = (int)(V.dir * V.w);
    const int kadj = V.dir < 0 ? k : 0;

    // producer: chunks of block b into buffer slot = b % nbuf (columns past
    // the end and padding lanes re-read the block's first chunk: in bounds,
    // never used)
    // The chunk a producer lane copies in load i of a block is the same for
    // every block up to a uniform shift: seg_lo(c0 + j) = seg_lo(j) +
    // dir * w * c0, and w * c0 is a multiple of E.  So the offsets (for block
    // 0) and columns are tabulated once per lane.
    constexpr int NIM = sw_ni_max<T>();
    int64_t poff[NIM];
    int pcol[NIM];
    const int64_t first0 = seg_lo(0) & ~(int64_t)(E - 1);
    if (wave >= SW_PROD0) {
        const int lane = threadIdx.x & 63;
#pragma unroll
        for (int i = 0; i < NIM; ++i) {
            const int rnd = i * SW_PROD + (wave - SW_PROD0);   // 64-chunk round of this block
            const int g = rnd * 64 + lane;
            const int j = g / G.nch, ch = g - j * G.nch;
            const bool real = i < G.ni && g < BS * G.nch;
            poff[i] = real ? (seg_lo(j) & ~(int64_t)(E - 1)) + (int64_t)ch * E : first0;
            pcol[i] = real ? j : 0;
        }
    }
    const int64_t blk_step = (int64_t)V.dir * V.w * BS;
    auto produce = [&](int64_t b, int slot, int lane) {
        T* dst = bufs + (size_t)slot * bufe;
        const int64_t shift = b * blk_step;
        const int64_t rem = len - b * BS;                // columns of this block inside the view
#pragma unroll
        for (int i = 0; i < NIM; ++i) {
            if (i < G.ni) {
                const int rnd = i * SW_PROD + (wave - SW_PROD0);
                // columns past the end re-read the block's first chunk
                const int64_t off = (pcol[i] < rem ? poff[i] : first0) + shift;
                __builtin_amdgcn_global_load_lds(V.base + off, dst + (size_t)rnd * 64 * E, 16, 0, 0);
            }
        }
        (void)lane;
    };
    auto produced_after = [&](int64_t b) -> int {   // this wave's block loads issued for blocks > b
        const int64_t last = min(nblk - 1, b + nbuf - 3);
        return (int)min(max((int64_t)0, last - b) * G.ni, (int64_t)63);
    };
    // right-hand side stream (wave 4): step s carries rows [s*BS + H, s*BS + H + BS)
    const int64_t nstream = len > H ? (len - H + BS - 1) / BS : 0;
    auto stream = [&](int64_t st, int lane) {
        const int tr = (lane / WPR) % BS, word = lane % WPR;
        int64_t r = st * BS + H + tr;
        if (r >= len) r = len - 1;
        const unsigned* sp = reinterpret_cast<const unsigned*>(src + (s0 + sdir * r)) + word;
        __builtin_amdgcn_global_load_lds(sp, instage + (st % SW_NIN) * 64, 4, 0, 0);
    };

    if (wave >= SW_PROD0) {
        for (int64_t b = 0; b < min(nblk, (int64_t)nbuf - 2); ++b) produce(b, (int)b, threadIdx.x & 63);
        wait_vmcnt(produced_after(0));
    } else if (wave == SW_STREAMER) {
        for (int64_t st = 0; st < min(nstream, (int64_t)SW_DIN); ++st) stream(st, threadIdx.x & 63);
    }
    lds_barrier();

    // Wave 0's band masks, bit j for block column j (lanes 0..BS-1):
    //   ma0: diagonal block, row i, column j < i with i - j <= k;
    //   mb0: sub-diagonal block, row BS + i, column j with BS + i - j <= k.
    //   mf0: far row t of waves 1-3 (row 2 BS + t of the previous block),
    //        column j with 2 BS + t - j <= k.
    unsigned ma0 = 0u, mb0 = 0u;
    {
        const int i = threadIdx.x & 63;
        if (i < BS) {
            ma0 = bits_below(i) & ~bits_below(i - k);
            mb0 = ~bits_below(BS + i - k);
        }
    }
    const unsigned mf0 = ~bits_below(2 * BS + ((int)threadIdx.x - 64) - k);
    // k - BS > 192: each lane of waves 1-3 takes a second far row, t + 192
    const unsigned mf1 = ~bits_below(2 * BS + SW_FAR + ((int)threadIdx.x - 64) - k);
    // slot = b % nbuf, kept as a 32-bit counter (a 64-bit modulo per step
    // costs more than the step's arithmetic)
    int slot = 0;
    for (int64_t b = 0; b <= nblk; ++b, slot = slot + 1 == nbuf ? 0 : slot + 1) {
        // lane-derived values re-materialised per step (no hoisted predicates)
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int lane = tid & 63;
        const int64_t c0 = b * BS;
        const int nb = (int)min((int64_t)BS, len - c0);
        const int par = (int)(b & 1);
        if (wave == 0) {
            if (c0 < len) {
                // element (row c0 + i, column c0 + j) of the block sits at
                // colb(j) + dir * i with colb(j) = j * segl + off_j - dir * j
                // (wave-uniform); rows outside the band read a neighbouring
                // LDS word (inside the buffers or their guard) and are masked
                const T* buf = bufs + (size_t)slot * bufe;
                const int lo0 = (int)seg_lo(c0);
                const int i = lane;
                const int di = V.dir * i;
                const bool rv = i < nb;
                const bool sv = i < BS && c0 + BS + i < len;
                // masks of this step (rows past the end, columns past nb)
                const unsigned ma = rv ? ma0 : 0u;
                const unsigned mb = sv ? (mb0 & (nb >= 32 ? ~0u : (1u << nb) - 1u)) : 0u;
                T a[BS], sb[BS];
#pragma unroll
                for (int j = 0; j < BS; ++j) {
                    const int colb = j * G.segl + ((lo0 + j * dwm) & (E - 1)) + kadj - V.dir * j;
                    a[j] = keep_if(buf[colb + di], ma, j);
                    sb[j] = keep_if(buf[colb + V.dir * BS + di], mb, j);
                }
                const int ii = i < BS ? i : 0;
                const T dv = buf[ii * G.segl + ((lo0 + ii * dwm) & (E - 1)) + kadj];
                const T d = (rv && scale) ? dv : (T)1;
                T x = rv ? pend[(c0 + lane) & ring_mask] - acc2[par * BS + lane] : (T)0;
#pragma unroll
                for (int j = 0; j < BS - 1; ++j) {
                    const T xj = bcast(x, j);
                    x -= a[j] * xj;
                }
                if (rv) {
                    xblk2[par * BS + lane] = x;
                    out(c0 + lane, scale ? x / d : x);
                }
                T s0_ = 0, s1_ = 0, s2_ = 0, s3_ = 0;
#pragma unroll
                for (int j = 0; j < BS; j += 4) {
                    s0_ += sb[j] * bcast(x, j);
                    s1_ += sb[j + 1] * bcast(x, j + 1);
                    s2_ += sb[j + 2] * bcast(x, j + 2);
                    s3_ += sb[j + 3] * bcast(x, j + 3);
                }
                if (sv) pend[(c0 + BS + lane) & ring_mask] -= (s0_ + s1_) + (s2_ + s3_);
            }
        } else if (wave < 4) {
            const int t = tid - 64;   // far-row index
            if (b >= 1) {
                const int64_t cp = c0 - BS;
                const T* buf = bufs + (size_t)(slot == 0 ? nbuf - 1 : slot - 1) * bufe;
                const int lo0 = (int)seg_lo(cp);
                const T* xp = xblk2 + (1 - par) * BS;
                const int64_t r = cp + 2 * BS + t;
                const bool fv = t < nfar && r < len;   // r < len: block b-1 is a full block
                const int dt = V.dir * (2 * BS + t);
                T s0_ = 0, s1_ = 0;
#pragma unroll
                for (int j = 0; j < BS; j += 2) {
                    const int colb0 = j * G.segl + ((lo0 + j * dwm) & (E - 1)) + kadj - V.dir * j;
                    const int colb1 = (j + 1) * G.segl + ((lo0 + (j + 1) * dwm) & (E - 1)) + kadj - V.dir * (j + 1);
                    s0_ += keep_if(buf[colb0 + dt], mf0, j) * xp[j];
                    s1_ += keep_if(buf[colb1 + dt], mf0, j + 1) * xp[j + 1];
                }
                const T acc = s0_ + s1_;
                if (t < BS) acc2[(1 - par) * BS + t] = fv ? acc : (T)0;
                else if (fv) pend[r & ring_mask] -= acc;
                if (nfar > SW_FAR) {   // wave-uniform
                    const int t2 = t + SW_FAR;
                    const int64_t r2 = cp + 2 * BS + t2;
                    const int dt2 = V.dir * (2 * BS + t2);
                    T u0 = 0, u1 = 0;
#pragma unroll
                    for (int j = 0; j < BS; j += 2) {
                        const int colb0 = j * G.segl + ((lo0 + j * dwm) & (E - 1)) + kadj - V.dir * j;
                        const int colb1 = (j + 1) * G.segl + ((lo0 + (j + 1) * dwm) & (E - 1)) + kadj - V.dir * (j + 1);
                        u0 += keep_if(buf[colb0 + dt2], mf1, j) * xp[j];
                        u1 += keep_if(buf[colb1 + dt2], mf1, j + 1) * xp[j + 1];
                    }
                    if (t2 < nfar && r2 < len) pend[r2 & ring_mask] -= u0 + u1;
                }
            } else if (t < BS) {
                acc2[BS + t] = (T)0;
            }
        } else if (wave >= SW_PROD0) {
            if (b + nbuf - 2 < nblk) produce(b + nbuf - 2, slot >= 2 ? slot - 2 : slot + nbuf - 2, lane);
            wait_vmcnt(produced_after(b + 1));
        } else if (b < nstream) {
            // right-hand side rows c0 + H .. c0 + H + BS - 1 into the pending
            // ring (rows no other wavefront touches in this step)
            if (b + SW_DIN < nstream) stream(b + SW_DIN, lane);
            wait_vmcnt((int)(min(b + SW_DIN, nstream - 1) - b));
            const int64_t rr = c0 + H + lane;
            if (lane < BS && rr < len) {
                const RawT raw = reinterpret_cast<const RawT*>(instage + (b % SW_NIN) * 64)[lane];
                pend[rr & ring_mask] = fin(rr, raw);
            }
        }
        lds_barrier();
    }
}

inline int ring_size(int k)
{
    int r = 1;
    while (r < 2 * NBS + k) r <<= 1;
    return r;
}

template <class T>
__device__ inline T* lds_ring()
{
    extern __shared__ unsigned char smem_raw[];
    return reinterpret_cast<T*>(smem_raw);
}

// Several right-hand sides: workgroup -> (job, right-hand side) for nj jobs
// and nr right-hand sides.  The nr workgroups of one job get ids that are
// equal mod 8, and workgroup ids are dealt to the 8 XCDs round-robin, so the
// sweeps of one partition for all right-hand sides run on one XCD and read
// its factors through that XCD's L2.  With nr = 1 the job is blockIdx.x.
// Grid: ceil(nj / 8) * 8 * nr workgroups; false for the padding ones.
__device__ inline bool job_and_rhs(int nj, int nr, int& job, int& r)
{
    const int b = blockIdx.x;
    const int q = b >> 3;
    r = q % nr;
    job = (q / nr) * 8 + (b & 7);
    return job < nj;
}

inline unsigned job_rhs_grid(int nj, int nr) { return (unsigned)(((nj + 7) / 8) * 8 * nr); }

// ---------------------------------------------------------------------------
// pass 1 (LU_UL): blockIdx.x = 2*i + side, i = boundary.
//   side 0: partition i, LU factors (B), bottom k rows of A_i^{-1} v_i
//   side 1: partition i+1, UL factors (Bul), top k rows of A_{i+1}^{-1} v_{i+1}
// g receives those 2k values per boundary at their row positions;
// scr is a length-n scratch vector per side.
// ---------------------------------------------------------------------------
// MODE: 0 simple blocked sweep, 1 register-pipelined, 2 LDS-prefetched
// (512 threads; nbuf block buffers after the ring in dynamic LDS).
template <class T, int MODE>
__global__ __launch_bounds__(MODE == 2 ? SW_THREADS : 256) void sweep_interface_kernel(
    const T* __restrict__ B, const T* __restrict__ Bul, int k, PartInfo pi, const double* __restrict__ v,
    T* __restrict__ g, T* __restrict__ scr0, T* __restrict__ scr1, int ring_mask, int nbuf, int nr, int64_t ldv,
    const int* __restrict__ gate)
{
    T* pend = lds_ring<T>();
    __shared__ T xblk[NBS];
    __shared__ T xblk2[2 * NBS];
    __shared__ T acc2[2 * NBS];
    int job, rh;
    if (!job_and_rhs(2 * (pi.P - 1), nr, job, rh)) return;
    if (gate && !gate[rh]) return;
    v += rh * ldv;
    g += rh * pi.n;
    scr0 += rh * pi.n;
    scr1 += rh * pi.n;
    const int i = job >> 1;
    const int side = job & 1;
    const int p = i + side;
    int64_t s, len;
    pi.bounds(p, s, len);
    const T* F = side == 0 ? B : Bul;
    T* scr = side == 0 ? scr0 : scr1;
    const int d1 = side == 0 ? 1 : -1;       // stage-1 direction
    const BandView<T> V1 = make_view(F, s, len, k, d1);
    const BandView<T> V2 = make_view(F, s, len, k, -d1);
    // local row r of a view with direction d -> global row
    auto grow = [&](int d, int64_t r) -> int64_t { return d > 0 ? s + r : s + len - 1 - r; };

    auto in1 = [&](int64_t r) -> T { return (T)v[grow(d1, r)]; };
    auto in2 = [&](int64_t r) -> T { return scr[grow(-d1, r)]; };
    auto out2 = [&](int64_t r, T y) { g[grow(-d1, r)] = y; };
    const int64_t kl = min((int64_t)k, len);
    if constexpr (MODE == 2) {
        T* dyn = pend + ring_mask + 1;
        auto cvt = [](int64_t, double raw) -> T { return (T)raw; };
        auto same = [](int64_t, T raw) -> T { return raw; };
        lower_sweep_lds<T>(V1, len, v, grow(d1, 0), d1, cvt, [&](int64_t r, T y) { scr[grow(d1, r)] = y; }, true,
                           pend, ring_mask, xblk2, acc2, dyn, nbuf);
        __syncthreads();
        lower_sweep_lds<T>(V2, kl, (const T*)scr, grow(-d1, 0), -d1, same, out2, false, pend, ring_mask, xblk2, acc2,
                           dyn, nbuf);
    } else if constexpr (MODE == 1) {
        lower_sweep_fast<T>(V1, len, in1, [&](int64_t r, T y) { scr[grow(d1, r)] = y; }, true, pend, ring_mask,
                            xblk2, acc2);
        __syncthreads();
        lower_sweep_fast<T>(V2, kl, in2, out2, false, pend, ring_mask, xblk2, acc2);
    } else {
        lower_sweep<T>(V1, len, in1, [&](int64_t r, T y) { scr[grow(d1, r)] = y / V1.at(r, r); }, pend, ring_mask,
                       xblk);
        __syncthreads();
        lower_sweep<T>(V2, kl, in2, out2, pend, ring_mask, xblk);
    }
}

// ---------------------------------------------------------------------------
// pass 2: blockIdx.x = partition.  out = A_p^{-1} (v - corr) with the LU
// factors (stage 1 forward L with D^{-1}, stage 2 backward U'), or for the
// UL-factored last partition the mirror order.  corr (2k per boundary) may
// be null.  Output FP64 (outd) or preconditioner precision (outt).
// ---------------------------------------------------------------------------
// The full sweep of partition p for one right-hand side (pointers already
// at that column), shared by the per-system and the job-table kernels.
template <class T, int MODE>
__device__ __forceinline__ void sweep_full_body(const T* __restrict__ B, const T* __restrict__ Bul, int k, PartInfo pi,
                                                int last_is_ul, int p, const double* __restrict__ v,
                                                const T* __restrict__ corr, double* __restrict__ outd,
                                                T* __restrict__ outt, T* __restrict__ scr, int ring_mask, int nbuf,
                                                T* pend, T* xblk, T* xblk2, T* acc2, int half = -1,
                                                int64_t s_ovr = 0, int64_t len_ovr = 0)
{
    // len_ovr > 0: the rows [s_ovr, s_ovr + len_ovr) of the band instead of
    // partition p (a block of a partition)
    int64_t s, len;
    if (len_ovr > 0) { s = s_ovr; len = len_ovr; }
    else pi.bounds(p, s, len);
    // half 0: rows [s + len/2, s + len) from the LU factors (stage 2 from
    // the bottom up to the middle); half 1: rows [s, s + len/2) from the UL
    // factors (stage 2 from the top down to the middle); -1: all rows
    const bool ul = half >= 0 ? half == 1 : (last_is_ul && p == pi.P - 1);
    const int64_t mid = len / 2;
    const int64_t len2 = half < 0 ? len : (half == 1 ? mid : len - mid);
    const T* F = ul ? Bul : B;
    const int d1 = ul ? -1 : 1;
    const BandView<T> V1 = make_view(F, s, len, k, d1);
    const BandView<T> V2 = make_view(F, s, len, k, -d1);
    auto grow = [&](int d, int64_t r) -> int64_t { return d > 0 ? s + r : s + len - 1 - r; };
    const bool has_top = corr && p > 0;
    const bool has_bot = corr && p < pi.P - 1;
    auto rhs = [&](int64_t r) -> T {
        const int64_t gr = grow(d1, r);
        T val = (T)v[gr];
        const int64_t lr = gr - s;
        if (has_top && lr < k) val -= corr[(size_t)(p - 1) * 2 * k + k + lr];
        if (has_bot && lr >= len - k) val -= corr[(size_t)p * 2 * k + (lr - (len - k))];
        return val;
    };
    auto in2 = [&](int64_t r) -> T { return scr[grow(-d1, r)]; };
    if constexpr (MODE == 2) {
        T* dyn = pend + ring_mask + 1;
        auto rhs_fin = [&](int64_t r, double raw) -> T {
            T val = (T)raw;
            const int64_t lr = grow(d1, r) - s;
            if (has_top && lr < k) val -= corr[(size_t)(p - 1) * 2 * k + k + lr];
            if (has_bot && lr >= len - k) val -= corr[(size_t)p * 2 * k + (lr - (len - k))];
            return val;
        };
        auto same = [](int64_t, T raw) -> T { return raw; };
        lower_sweep_lds<T>(V1, len, v, grow(d1, 0), d1, rhs_fin, [&](int64_t r, T y) { scr[grow(d1, r)] = y; }, true,
                           pend, ring_mask, xblk2, acc2, dyn, nbuf);
        __syncthreads();
        if (outd)
            lower_sweep_lds<T>(V2, len2, (const T*)scr, grow(-d1, 0), -d1, same,
                               [&](int64_t r, T y) { outd[grow(-d1, r)] = (double)y; }, false, pend, ring_mask, xblk2,
                               acc2, dyn, nbuf);
        else
            lower_sweep_lds<T>(V2, len2, (const T*)scr, grow(-d1, 0), -d1, same,
                               [&](int64_t r, T y) { outt[grow(-d1, r)] = y; }, false, pend, ring_mask, xblk2, acc2,
                               dyn, nbuf);
    } else if constexpr (MODE == 1) {
        lower_sweep_fast<T>(V1, len, rhs, [&](int64_t r, T y) { scr[grow(d1, r)] = y; }, true, pend, ring_mask, xblk2,
                            acc2);
        __syncthreads();
        if (outd)
            lower_sweep_fast<T>(V2, len2, in2, [&](int64_t r, T y) { outd[grow(-d1, r)] = (double)y; }, false, pend,
                                ring_mask, xblk2, acc2);
        else
            lower_sweep_fast<T>(V2, len2, in2, [&](int64_t r, T y) { outt[grow(-d1, r)] = y; }, false, pend, ring_mask,
                                xblk2, acc2);
    } else {
        lower_sweep<T>(V1, len, rhs, [&](int64_t r, T y) { scr[grow(d1, r)] = y / V1.at(r, r); }, pend, ring_mask,
                       xblk);
        __syncthreads();
        if (outd)
            lower_sweep<T>(V2, len2, in2, [&](int64_t r, T y) { outd[grow(-d1, r)] = (double)y; }, pend, ring_mask,
                           xblk);
        else
            lower_sweep<T>(V2, len2, in2, [&](int64_t r, T y) { outt[grow(-d1, r)] = y; }, pend, ring_mask, xblk);
    }
}

template <class T, int MODE>
__global__ __launch_bounds__(MODE == 2 ? SW_THREADS : 256) void sweep_full_kernel(
    const T* __restrict__ B, const T* __restrict__ Bul, int k, PartInfo pi, int last_is_ul,
    const double* __restrict__ v, const T* __restrict__ corr, double* __restrict__ outd, T* __restrict__ outt,
    T* __restrict__ scr, int ring_mask, int nbuf, int nr, int64_t ldv, int64_t ldo, const int* __restrict__ gate,
    int split, T* __restrict__ scr_alt)
{
    T* pend = lds_ring<T>();
    __shared__ T xblk[NBS];
    __shared__ T xblk2[2 * NBS];
    __shared__ T acc2[2 * NBS];
    // split: two workgroups per partition, the bottom half of its rows from
    // the LU factors and the top half from the UL factors (scratch scr_alt)
    int job, rh;
    if (!job_and_rhs(split ? 2 * pi.P : pi.P, nr, job, rh)) return;
    if (gate && !gate[rh]) return;
    const int p = split ? job >> 1 : job;
    const int half = split ? (job & 1) : -1;
    v += rh * ldv;
    if (corr) corr += rh * (int64_t)2 * k * (pi.P - 1);
    if (outd) outd += rh * ldo;
    if (outt) outt += rh * pi.n;
    T* sc = (half == 1 ? scr_alt : scr) + rh * pi.n;
    sweep_full_body<T, MODE>(B, Bul, k, pi, last_is_ul, p, v, corr, outd, outt, sc, ring_mask, nbuf, pend, xblk,
                             xblk2, acc2, half);
}

// Full sweeps of several systems in one launch: workgroup b takes job b of
// a table (one partition of one system for one right-hand side; null B:
// padding), with that system's band, width, staging depth and LDS ring;
// the input column is V + voff and the output Z + ooff.
template <class T, int MODE>
__global__ __launch_bounds__(MODE == 2 ? SW_THREADS : 256) void sweep_full_jobs_kernel(
    const SweepJob* __restrict__ jobs, const double* __restrict__ V, double* __restrict__ Z,
    const int* __restrict__ gate)
{
    T* pend = lds_ring<T>();
    __shared__ T xblk[NBS];
    __shared__ T xblk2[2 * NBS];
    __shared__ T acc2[2 * NBS];
    const SweepJob J = jobs[blockIdx.x];
    if (!J.B) return;
    if (gate && !gate[J.gate]) return;
    const PartInfo pi{J.n, J.P};
    sweep_full_body<T, MODE>((const T*)J.B, (const T*)J.Bul, J.k, pi, J.last_is_ul, J.part, V + J.voff,
                          (const T*)J.corr, Z + J.ooff, nullptr, (T*)J.scr, J.ring_mask, J.nbuf, pend, xblk, xblk2,
                          acc2, -1, J.s0, J.len0);
}

// ---------------------------------------------------------------------------
// Reduced solve and purification per boundary i (one workgroup):
//   zt' = g_t - W g_b ;  x_t = S^{-1} zt' ;  x_b = g_b - V x_t
//   corr_b = B_i x_t (rows e0-k..e0-1),  corr_t = C_{i+1} x_b (rows s1..s1+k-1)
// All k x k matrices column-major.
// ---------------------------------------------------------------------------
template <class T>
__device__ void gemv_cm(const T* __restrict__ M, int k, const T* x, T* y, T beta_src_sign, const T* src)
{
    // y[a] = src[a] + sign * sum_b M(a, b) x[b]   (src may be null -> 0)
    for (int a = threadIdx.x; a < k; a += blockDim.x) {
        T acc = (T)0;
        for (int b = 0; b < k; ++b) acc += M[(size_t)b * k + a] * x[b];
        y[a] = (src ? src[a] : (T)0) + beta_src_sign * acc;
    }
}

// The same with 1024 threads: thread (a mod 256, q) sums columns
// [q k / 4, (q + 1) k / 4) in order, and the four partial sums are added as
// (p0 + p1) + (p2 + p3) through part[4][256] in LDS, so each output row
// has four times the loads in flight.  Ends with a barrier (y is complete).
template <class T>
__device__ void gemv_cm4(const T* __restrict__ M, int k, const T* x, T* y, T beta_src_sign, const T* src, T* part)
{
    const int t = threadIdx.x & 255, q = threadIdx.x >> 8;
    const int b0 = (k * q) / 4, b1 = (k * (q + 1)) / 4;
    for (int a0 = 0; a0 < k; a0 += 256) {
        const int a = a0 + t;
        T acc = (T)0;
        if (a < k) {
            // eight loads issued together, then added in column order
            int b = b0;
            for (; b + 8 <= b1; b += 8) {
                T m[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) m[u] = M[(size_t)(b + u) * k + a];
#pragma unroll
                for (int u = 0; u < 8; ++u) acc += m[u] * x[b + u];
            }
            for (; b < b1; ++b) acc += M[(size_t)b * k + a] * x[b];
        }
        part[q * 256 + t] = acc;
        __syncthreads();
        if (q == 0 && a < k)
            y[a] = (src ? src[a] : (T)0) + beta_src_sign * ((part[t] + part[256 + t]) + (part[512 + t] + part[768 + t]));
        __syncthreads();
    }
}

template <class T>
__global__ __launch_bounds__(256) void interface_kernel(const T* __restrict__ WV, const T* __restrict__ Sinv,
                                                        const T* __restrict__ offB, const T* __restrict__ offC, int k,
                                                        PartInfo pi, const T* __restrict__ g, T* __restrict__ corr,
                                                        const int* __restrict__ gate)
{
    extern __shared__ unsigned char smem_raw[];
    T* gb = reinterpret_cast<T*>(smem_raw);
    T* gt = gb + k;
    T* zt = gt + k;
    T* xt = zt + k;
    T* xb = xt + k;
    const int i = blockIdx.x;
    if (gate && !gate[blockIdx.y]) return;
    g += blockIdx.y * pi.n;
    corr += blockIdx.y * (int64_t)2 * k * (pi.P - 1);
    int64_t s1, l1;
    pi.bounds(i + 1, s1, l1);
    const size_t kk = (size_t)k * k;
    for (int a = threadIdx.x; a < k; a += blockDim.x) {
        gb[a] = g[s1 - k + a];
        gt[a] = g[s1 + a];
    }
    __syncthreads();
    gemv_cm<T>(WV + (2 * i + 1) * kk, k, gb, zt, (T)-1, gt);   // zt = gt - W gb
    __syncthreads();
    gemv_cm<T>(Sinv + i * kk, k, zt, xt, (T)1, nullptr);        // xt = S^{-1} zt
    __syncthreads();
    gemv_cm<T>(WV + 2 * i * kk, k, xt, xb, (T)-1, gb);          // xb = gb - V xt
    __syncthreads();
    T* c = corr + (size_t)i * 2 * k;
    gemv_cm<T>(offB + i * kk, k, xt, c, (T)1, nullptr);         // B_i xt
    gemv_cm<T>(offC + i * kk, k, xb, c + k, (T)1, nullptr);     // C_{i+1} xb
}

// interface_kernel with 1024 threads and gemv_cm4
template <class T>
__global__ __launch_bounds__(1024) void interface4_kernel(const T* __restrict__ WV, const T* __restrict__ Sinv,
                                                          const T* __restrict__ offB, const T* __restrict__ offC,
                                                          int k, PartInfo pi, const T* __restrict__ g,
                                                          T* __restrict__ corr, const int* __restrict__ gate)
{
    extern __shared__ unsigned char smem_raw[];
    T* part = reinterpret_cast<T*>(smem_raw);
    T* gb = part + 1024;
    T* gt = gb + k;
    T* zt = gt + k;
    T* xt = zt + k;
    T* xb = xt + k;
    const int i = blockIdx.x;
    if (gate && !gate[blockIdx.y]) return;
    g += blockIdx.y * pi.n;
    corr += blockIdx.y * (int64_t)2 * k * (pi.P - 1);
    int64_t s1, l1;
    pi.bounds(i + 1, s1, l1);
    const size_t kk = (size_t)k * k;
    for (int a = threadIdx.x; a < k; a += blockDim.x) {
        gb[a] = g[s1 - k + a];
        gt[a] = g[s1 + a];
    }
    __syncthreads();
    gemv_cm4<T>(WV + (2 * i + 1) * kk, k, gb, zt, (T)-1, gt, part);   // zt = gt - W gb
    gemv_cm4<T>(Sinv + i * kk, k, zt, xt, (T)1, nullptr, part);        // xt = S^{-1} zt
    gemv_cm4<T>(WV + 2 * i * kk, k, xt, xb, (T)-1, gb, part);          // xb = gb - V xt
    T* c = corr + (size_t)i * 2 * k;
    gemv_cm4<T>(offB + i * kk, k, xt, c, (T)1, nullptr, part);         // B_i xt
    gemv_cm4<T>(offC + i * kk, k, xb, c + k, (T)1, nullptr, part);     // C_{i+1} xb
}

template <class T>
__global__ void diag_divide_kernel(const T* __restrict__ B, int64_t n, const double* __restrict__ v,
                                   double* __restrict__ out, int64_t ldv, int64_t ldo, const int* __restrict__ gate)
{
    if (gate && !gate[blockIdx.y]) return;
    v += blockIdx.y * ldv;
    out += blockIdx.y * ldo;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = (double)((T)v[i] / B[i]);
}

template <class T>
size_t sweep_lds(int k)
{
    return sizeof(T) * (size_t)ring_size(k);
}

}  // namespace

template <class T>
bool sweep_fast(int k)
{
    return k <= 192 + sweep_bs<T>();
}

// the LDS-staged sweep with two far rows per lane (MODE 2 only)
template <class T>
bool sweep_staged_wide(int k)
{
    return k > 192 + sweep_bs<T>() && k <= 2 * SW_FAR + sweep_bs<T>();
}

// Sweep configuration: mode and (mode 2) number of LDS block buffers, as
// many as fit next to the ring (at most 8) within the 63 outstanding loads a
// producer wavefront can track.
struct SweepCfg { int mode, nbuf; size_t lds; int threads; };

template <class T>
SweepCfg sweep_cfg(int k, int nbuf_req = 0)
{
    constexpr size_t LDS_MAX = 160 * 1024 - 4096;   // static LDS of the kernels + margin
    const size_t ring = sweep_lds<T>(k);
    const bool wide = sweep_staged_wide<T>(k);
    if (!sweep_fast<T>(k) && !wide) return {0, 0, ring, 256};
    const SweepChunks<T> G(k, sweep_bs<T>());
    const size_t bb = G.buf_elems() * sizeof(T);
    const size_t stage = SW_NIN * 64 * 4;
    const size_t guard = 256 * sizeof(T);   // masked reads past the last buffer stay inside the allocation
    int nbuf = (int)std::min<size_t>(8, (LDS_MAX - ring - stage - guard) / bb);
    nbuf = std::min(nbuf, 63 / G.ni + 3);
    // at least four block buffers (loads two blocks ahead); wide FP64 bands,
    // whose blocks of 16 columns no longer fit four times, run with three
    // (the next block's loads then complete within the step that issues
    // them, still far faster than the plain blocked sweep)
    const int min_nbuf = (sizeof(T) == 8 && wide) ? 3 : 4;
    if (nbuf < min_nbuf || G.ni > sw_ni_max<T>()) return {wide ? 0 : 1, 0, ring, 256};
    if (nbuf_req >= 4 && nbuf_req < nbuf) nbuf = nbuf_req;
    return {2, nbuf, ring + stage + nbuf * bb + guard, SW_THREADS};
}

// Raises a kernel's dynamic LDS limit once per size (not a stream
// operation, but kept out of the launch path so that launches can be
// captured into a graph with nothing else in between).
void set_lds_limit(const void* kernel, size_t lds)
{
    static std::mutex mu;
    static std::unordered_map<const void*, size_t> done;
    std::lock_guard<std::mutex> lock(mu);
    size_t& have = done[kernel];
    if (lds <= have) return;
    SAP_HIP_CHECK(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    have = lds;
}

template <class T, int MODE>
void launch_sweep_interface_t(const T* B, const T* Bul, int k, PartInfo pi, const double* v, T* g, T* scr0, T* scr1,
                              const SweepCfg& c, const Rhs& rh, hipStream_t st)
{
    set_lds_limit((const void*)sweep_interface_kernel<T, MODE>, c.lds);
    hipLaunchKernelGGL((sweep_interface_kernel<T, MODE>), dim3(job_rhs_grid(2 * (pi.P - 1), rh.nr)), dim3(c.threads),
                       c.lds, st, B, Bul, k, pi, v, g, scr0, scr1, ring_size(k) - 1, c.nbuf, rh.nr, rh.ldv, rh.gate);
    SAP_LAUNCH_CHECK();
}

template <class T>
void launch_sweep_interface(const T* B, const T* Bul, int k, PartInfo pi, const double* v, T* g, T* scr0, T* scr1,
                            hipStream_t st, const Rhs& rh)
{
    if (pi.P < 2) return;
    // several columns: the fewest staging buffers, so more workgroups fit on a CU
    const SweepCfg c = sweep_cfg<T>(k, rh.nr > 1 ? 4 : rh.nbuf);
    if (c.mode == 2)      launch_sweep_interface_t<T, 2>(B, Bul, k, pi, v, g, scr0, scr1, c, rh, st);
    else if (c.mode == 1) launch_sweep_interface_t<T, 1>(B, Bul, k, pi, v, g, scr0, scr1, c, rh, st);
    else                  launch_sweep_interface_t<T, 0>(B, Bul, k, pi, v, g, scr0, scr1, c, rh, st);
}

template <class T>
void launch_interface_solve(const T* WV, const T* Sinv, const T* offB, const T* offC, int k, PartInfo pi, const T* g,
                            T* corr, hipStream_t st, const Rhs& rh)
{
    if (pi.P < 2) return;
    const size_t lds4 = sizeof(T) * (1024 + 5 * (size_t)k);
    if (lds4 <= 64 * 1024) {
        hipLaunchKernelGGL(interface4_kernel<T>, dim3(pi.P - 1, rh.nr), dim3(1024), lds4, st, WV, Sinv, offB, offC, k,
                           pi, g, corr, rh.gate);
    } else {
        const size_t lds = sizeof(T) * 5 * (size_t)k;
        hipLaunchKernelGGL(interface_kernel<T>, dim3(pi.P - 1, rh.nr), dim3(256), lds, st, WV, Sinv, offB, offC, k,
                           pi, g, corr, rh.gate);
    }
    SAP_LAUNCH_CHECK();
}

template <class T, int MODE>
void launch_sweep_full_t(const T* B, const T* Bul, int k, PartInfo pi, int last_is_ul, const double* v, const T* corr,
                         double* outd, T* outt, T* scr, const SweepCfg& c, const Rhs& rh, hipStream_t st)
{
    set_lds_limit((const void*)sweep_full_kernel<T, MODE>, c.lds);
    const int split = rh.split_scr != nullptr;
    hipLaunchKernelGGL((sweep_full_kernel<T, MODE>), dim3(job_rhs_grid(split ? 2 * pi.P : pi.P, rh.nr)),
                       dim3(c.threads), c.lds, st, B, Bul, k, pi, last_is_ul, v, corr, outd, outt, scr,
                       ring_size(k) - 1, c.nbuf, rh.nr, rh.ldv, rh.ldo, rh.gate, split, (T*)rh.split_scr);
    SAP_LAUNCH_CHECK();
}

template <class T>
void launch_sweep_full(const T* B, const T* Bul, int k, PartInfo pi, int last_is_ul, const double* v, const T* corr,
                       double* outd, T* outt, T* scr, hipStream_t st, const Rhs& rh)
{
    const SweepCfg c = sweep_cfg<T>(k, rh.nr > 1 ? 4 : rh.nbuf);
    if (c.mode == 2)      launch_sweep_full_t<T, 2>(B, Bul, k, pi, last_is_ul, v, corr, outd, outt, scr, c, rh, st);
    else if (c.mode == 1) launch_sweep_full_t<T, 1>(B, Bul, k, pi, last_is_ul, v, corr, outd, outt, scr, c, rh, st);
    else                  launch_sweep_full_t<T, 0>(B, Bul, k, pi, last_is_ul, v, corr, outd, outt, scr, c, rh, st);
}

template <class T>
void launch_diag_divide(const T* B, int64_t n, const double* v, double* out, hipStream_t st, const Rhs& rh)
{
    int64_t g = (n + 255) / 256;
    if (g > 8192) g = 8192;
    hipLaunchKernelGGL(diag_divide_kernel<T>, dim3((unsigned)g, rh.nr), dim3(256), 0, st, B, n, v, out, rh.ldv,
                       rh.ldo, rh.gate);
    SAP_LAUNCH_CHECK();
}

template <class T>
bool sweep_jobs_append(std::vector<SweepJob>& jobs, size_t& lds, const T* B, const T* Bul, int k, PartInfo pi,
                       int last_is_ul, const T* corr, T* scr, int64_t voff, int64_t ldv, int64_t ooff, int64_t ldo,
                       int nr, int gate0, int nbuf_req, int* mode_out)
{
    // many workgroups: the fewest staging buffers, so more fit on a CU
    (void)nbuf_req;
    const SweepCfg c = sweep_cfg<T>(k, 4);
    if (k == 0) return false;
    if (mode_out) *mode_out = c.mode;
    else if (c.mode != 2) return false;
    lds = std::max(lds, c.lds);
    // the kernel's (job, right-hand side) order per system: the nr
    // workgroups of a partition 8 ids apart (one XCD), padded to whole
    // groups of 8 partitions
    const int groups = (pi.P + 7) / 8;
    for (int q = 0; q < groups * 8 * nr; ++q) {
        const int r = (q >> 3) % nr;
        const int part = ((q >> 3) / nr) * 8 + (q & 7);
        SweepJob J{};
        if (part < pi.P) {
            J.B = B;
            J.Bul = Bul;
            J.corr = corr ? corr + (int64_t)r * 2 * k * (pi.P - 1) : nullptr;
            J.scr = scr + (int64_t)r * pi.n;
            J.voff = voff + r * ldv;
            J.ooff = ooff + r * ldo;
            J.n = pi.n;
            J.P = pi.P;
            J.k = k;
            J.part = part;
            J.gate = gate0 + r;
            J.last_is_ul = last_is_ul;
            J.ring_mask = ring_size(k) - 1;
            J.nbuf = c.nbuf;
        }
        jobs.push_back(J);
    }
    return true;
}

template <class T>
int sweep_range_job_append(std::vector<SweepJob>& jobs, size_t& lds, const T* B, int k, int64_t s0, int64_t len0,
                           T* scr, int64_t n, int64_t ldv, int64_t ldo, int nr)
{
    const SweepCfg c = sweep_cfg<T>(k, 4);
    lds = std::max(lds, c.lds);
    for (int r = 0; r < nr; ++r) {
        SweepJob J{};
        J.B = B;
        J.Bul = B;
        J.corr = nullptr;
        J.scr = scr + (int64_t)r * n;
        J.voff = r * ldv;
        J.ooff = r * ldo;
        J.n = n;
        J.P = 1;
        J.k = k;
        J.part = 0;
        J.gate = r;
        J.last_is_ul = 0;
        J.ring_mask = ring_size(k) - 1;
        J.nbuf = c.nbuf;
        J.s0 = s0;
        J.len0 = len0;
        jobs.push_back(J);
    }
    return c.mode;
}

template <class T, int MODE>
void launch_sweep_full_jobs_t(const SweepJob* d_jobs, int njobs, size_t lds, const double* V, double* Z,
                              const int* gate, hipStream_t st)
{
    set_lds_limit((const void*)sweep_full_jobs_kernel<T, MODE>, lds);
    hipLaunchKernelGGL((sweep_full_jobs_kernel<T, MODE>), dim3(njobs), dim3(MODE == 2 ? SW_THREADS : 256), lds, st,
                       d_jobs, V, Z, gate);
    SAP_LAUNCH_CHECK();
}

template <class T>
void launch_sweep_full_jobs(const SweepJob* d_jobs, int njobs, size_t lds, const double* V, double* Z,
                            const int* gate, hipStream_t st, int mode)
{
    if (njobs <= 0) return;
    if (mode == 2)      launch_sweep_full_jobs_t<T, 2>(d_jobs, njobs, lds, V, Z, gate, st);
    else if (mode == 1) launch_sweep_full_jobs_t<T, 1>(d_jobs, njobs, lds, V, Z, gate, st);
    else                launch_sweep_full_jobs_t<T, 0>(d_jobs, njobs, lds, V, Z, gate, st);
}

template <class T>
int sweep_max_nbuf(int k)
{
    const SweepCfg c = sweep_cfg<T>(k);
    return c.mode == 2 ? c.nbuf : 0;
}

#define SAP_INST(T)                                                                                               \
    template int sweep_max_nbuf<T>(int);                                                                           \
    template bool sweep_jobs_append<T>(std::vector<SweepJob>&, size_t&, const T*, const T*, int, PartInfo, int,    \
                                       const T*, T*, int64_t, int64_t, int64_t, int64_t, int, int, int, int*);      \
    template void launch_sweep_full_jobs<T>(const SweepJob*, int, size_t, const double*, double*, const int*,      \
                                            hipStream_t, int);                                                     \
    template int sweep_range_job_append<T>(std::vector<SweepJob>&, size_t&, const T*, int, int64_t, int64_t, T*,  \
                                           int64_t, int64_t, int64_t, int);                                        \
    template void launch_sweep_interface<T>(const T*, const T*, int, PartInfo, const double*, T*, T*, T*,          \
                                            hipStream_t, const Rhs&);                                              \
    template void launch_interface_solve<T>(const T*, const T*, const T*, const T*, int, PartInfo, const T*, T*,   \
                                            hipStream_t, const Rhs&);                                              \
    template void launch_sweep_full<T>(const T*, const T*, int, PartInfo, int, const double*, const T*, double*,   \
                                       T*, T*, hipStream_t, const Rhs&);                                           \
    template void launch_diag_divide<T>(const T*, int64_t, const double*, double*, hipStream_t, const Rhs&);
SAP_INST(float)
SAP_INST(double)
#undef SAP_INST

}  // namespace sap_amd
