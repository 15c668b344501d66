// Band-format kernels: workload generator, FP64 banded SpMV, and the
// row-major FP64 band -> column-major preconditioner band conversion.
#include "sap_internal.hpp"

#include <atomic>

#include <algorithm>

namespace sap_amd {

namespace {

// ---------------------------------------------------------------------------
// xoshiro256** (Blackman & Vigna), one stream per row seeded by splitmix64.
// Same definition as the CPU generator used by the tests.
// ---------------------------------------------------------------------------
__device__ inline uint64_t splitmix64(uint64_t& x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

struct Xoshiro {
    uint64_t s0, s1, s2, s3;
    __device__ Xoshiro(uint64_t seed, uint64_t stream)
    {
        uint64_t x = seed ^ (stream * 0xD1B54A32D192ED03ULL);
        s0 = splitmix64(x); s1 = splitmix64(x); s2 = splitmix64(x); s3 = splitmix64(x);
    }
    __device__ inline uint64_t next()
    {
        const uint64_t result = rotl64(s1 * 5, 7) * 9;
        const uint64_t t = s1 << 17;
        s2 ^= s0; s3 ^= s1; s1 ^= s2; s0 ^= s3; s2 ^= t;
        s3 = rotl64(s3, 45);
        return result;
    }
    __device__ inline double uniform_pm1() { return (double)(next() >> 11) * 0x1.0p-52 - 1.0; }
};

// One thread per row; the row's random values are drawn in column order.
__global__ void gen_banded_kernel(int64_t n, int k, double dom, uint64_t seed, double* __restrict__ band)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int w = 2 * k + 1;
    Xoshiro rng(seed, (uint64_t)i + 1);
    double* row = band + i * w;
    double s = 0.0;
    for (int d = 0; d < w; ++d) {
        if (d == k) continue;
        const int64_t j = i + d - k;
        if (j < 0 || j >= n) { row[d] = 0.0; continue; }
        const double v = rng.uniform_pm1();
        row[d] = v;
        s += fabs(v);
    }
    row[k] = dom * s;
}

__global__ void gen_vector_kernel(int64_t n, uint64_t seed, double* __restrict__ x)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Xoshiro rng(seed, (uint64_t)i + 1);
    x[i] = rng.uniform_pm1();
}

// ---------------------------------------------------------------------------
// y = A x on the row-major FP64 band (the MVBanded functor, sap/spmv.h:27-74).
// G lanes cooperate on one row: lane l sums diagonals l, l+G, ... in order,
// then a fixed shuffle tree combines the G partial sums.  Each lane issues
// the loads of NI of its diagonals before the multiply-adds (NI * G entries
// of the row per round); the band is read with non-temporal loads since a
// product streams it exactly once, leaving the caches to x.  Each group of G
// lanes takes RPW consecutive rows and issues the loads of all of them
// before their multiply-adds (more bytes in flight per wavefront, fewer
// wavefronts to launch); every row keeps its lanes, order and shuffle tree,
// so the result does not depend on RPW.
// ---------------------------------------------------------------------------
template <int G, int NI, int RPW>
__global__ __launch_bounds__(256) void band_spmv_kernel(int64_t n, int k, const double* __restrict__ band,
                                                        const double* __restrict__ x, double* __restrict__ y)
{
    const int w = 2 * k + 1;
    const int lane = threadIdx.x & (G - 1);
    const int64_t row0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G * RPW;
    double s[RPW];
    int dlo[RPW], dhi[RPW];
    int span = -1;   // the longest row's last offset from its first diagonal
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
        const int64_t row = row0 + r;
        s[r] = 0.0;
        dlo[r] = (int)(row < k ? k - row : 0);
        const int64_t dhi64 = n - 1 - row + k;
        dhi[r] = row < n ? (int)(dhi64 < w - 1 ? dhi64 : w - 1) : -1;   // -1: past the last row, nothing read
        span = max(span, dhi[r] - dlo[r]);
    }
    for (int o = lane; o <= span; o += NI * G) {
        double av[RPW][NI], xv[RPW][NI];
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            const double* rb = band + (row0 + r) * w;
            const double* xr = x + (row0 + r) - k;
#pragma unroll
            for (int m = 0; m < NI; ++m) {
                const int d = dlo[r] + o + m * G;
                av[r][m] = d <= dhi[r] ? __builtin_nontemporal_load(rb + d) : 0.0;
                xv[r][m] = d <= dhi[r] ? xr[d] : 0.0;
            }
        }
#pragma unroll
        for (int r = 0; r < RPW; ++r)
#pragma unroll
            for (int m = 0; m < NI; ++m)
                if (dlo[r] + o + m * G <= dhi[r]) s[r] = fma(av[r][m], xv[r][m], s[r]);
    }
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) s[r] += __shfl_down(s[r], off, G);
        if (lane == 0 && row0 + r < n) y[row0 + r] = s[r];
    }
}

// Y(:, c) = A X(:, c) for RB columns per workgroup row (grid.y = column
// chunks): each band entry is loaded once for the RB columns, and every
// column gets band_spmv_kernel's lane assignment (G, NI), multiply-add order
// and shuffle tree, so the same bits as a product of that column alone.
// Columns whose gate flag is 0 are not computed.
template <int G, int NI, int RB>
__global__ __launch_bounds__(256) void band_spmm_kernel(int64_t n, int k, const double* __restrict__ band,
                                                        const double* __restrict__ X, int64_t ldx,
                                                        double* __restrict__ Y, int64_t ldy, int nr,
                                                        const int* __restrict__ gate)
{
    const int w = 2 * k + 1;
    const int lane = threadIdx.x & (G - 1);
    const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
    const int c0 = blockIdx.y * RB;
    bool on[RB];
#pragma unroll
    for (int c = 0; c < RB; ++c) on[c] = c0 + c < nr && (!gate || gate[c0 + c]);
    double s[RB];
#pragma unroll
    for (int c = 0; c < RB; ++c) s[c] = 0.0;
    if (row < n) {
        const double* r = band + row * w;
        const int dlo = (int)(row < k ? k - row : 0);
        const int64_t dhi64 = n - 1 - row + k;
        const int dhi = (int)(dhi64 < w - 1 ? dhi64 : w - 1);
        const double* xr = X + (int64_t)c0 * ldx + row - k;
        for (int d0 = dlo + lane; d0 <= dhi; d0 += NI * G) {
            double av[NI];
#pragma unroll
            for (int m = 0; m < NI; ++m) {
                const int d = d0 + m * G;
                av[m] = d <= dhi ? __builtin_nontemporal_load(r + d) : 0.0;
            }
#pragma unroll
            for (int c = 0; c < RB; ++c) {
                if (!on[c]) continue;
                double xv[NI];
#pragma unroll
                for (int m = 0; m < NI; ++m) {
                    const int d = d0 + m * G;
                    xv[m] = d <= dhi ? xr[c * ldx + d] : 0.0;
                }
#pragma unroll
                for (int m = 0; m < NI; ++m)
                    if (d0 + m * G <= dhi) s[c] = fma(av[m], xv[m], s[c]);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < RB; ++c) {
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) s[c] += __shfl_down(s[c], off, G);
        if (lane == 0 && row < n && on[c]) Y[(int64_t)(c0 + c) * ldy + row] = s[c];
    }
}

// ---------------------------------------------------------------------------
// Column-major band from the row-major input: B[c*(2k+1) + d] = A(c+d-k, c),
// cast to the preconditioner type (bandedMatrixTranspose,
// sap/device/data_transfer.cuh:732-770).  The tile of TC columns x TD
// diagonals is staged through LDS so both the reads (along input rows) and
// the writes (along output columns) are contiguous.
// ---------------------------------------------------------------------------
constexpr int TC = 64;
constexpr int TD = 32;

template <class T>
__global__ __launch_bounds__(256) void band_to_colmajor_kernel(int64_t n, int k, const double* __restrict__ in,
                                                               T* __restrict__ out, T* __restrict__ out2)
{
    __shared__ double tile[TC + TD][TD + 1];
    const int w = 2 * k + 1;
    const int64_t c0 = (int64_t)blockIdx.x * TC;
    const int d0 = blockIdx.y * TD;
    // input rows r = c + d - k for c in [c0, c0+TC), d in [d0, d0+TD):
    // r in [c0 + d0 - k, c0 + d0 - k + TC + TD - 1), input offset o = 2k - d.
    const int64_t rbase = c0 + d0 - k;
    for (int e = threadIdx.x; e < (TC + TD - 1) * TD; e += blockDim.x) {
        const int rr = e / TD;           // row within the box
        const int oo = e % TD;           // offset index: o = 2k - d0 - (TD - 1) + oo
        const int64_t r = rbase + rr;
        const int o = 2 * k - d0 - (TD - 1) + oo;
        double v = 0.0;
        if (r >= 0 && r < n && o >= 0 && o < w) v = in[r * w + o];
        tile[rr][oo] = v;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < TC * TD; e += blockDim.x) {
        const int cc = e / TD;
        const int dd = e % TD;
        const int64_t c = c0 + cc;
        const int d = d0 + dd;
        if (c >= n || d >= w) continue;
        const int64_t r = c + d - k;
        double v = 0.0;
        if (r >= 0 && r < n) v = tile[cc + dd][(TD - 1) - dd];
        out[c * w + d] = (T)v;
        if (out2) out2[c * w + d] = (T)v;
    }
}

// Column tiles: one workgroup builds tc whole output columns (all 2k+1
// diagonals) in LDS and writes them as one contiguous run.  Input row r
// holds A(r, c) at offset o = c - r + k, so the tile's columns take one
// contiguous segment of at most tc values from each of the tc + 2k rows
// r = c0 - k + rr, o in [max(0, 2k - rr), min(2k+1, 2k + tc - rr)): every
// input element is read once, in whole-segment loads.  In LDS, column cc
// has stride 2k+2 (even), so the scatter of one row's segment (consecutive
// o: stride 2k+1 words, odd) hits distinct banks.  out2, when given,
// receives a second copy (the band the UL factorizations overwrite).
// ---------------------------------------------------------------------------
constexpr int B2C_THREADS = 1024;
constexpr int B2C_ROWS = 8;   // rows per wavefront whose loads are in flight together

template <class T>
__global__ __launch_bounds__(B2C_THREADS) void band_to_colmajor_cols_kernel(int64_t n, int k, int tc,
                                                                        const double* __restrict__ in,
                                                                        T* __restrict__ out, T* __restrict__ out2)
{
    extern __shared__ unsigned char smem_raw[];
    T* const lds = reinterpret_cast<T*>(smem_raw);
    const int w = 2 * k + 1, ls = w + 1;
    const int64_t c0 = (int64_t)blockIdx.x * tc;
    const int ncol = (int)min((int64_t)tc, n - c0);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int NWV = B2C_THREADS / 64;
    const int nrow = ncol + 2 * k;
    for (int rb = wave; rb < nrow; rb += NWV * B2C_ROWS) {
        double v[B2C_ROWS];
#pragma unroll
        for (int u = 0; u < B2C_ROWS; ++u) {
            const int rr = rb + u * NWV;
            const int64_t r = c0 - k + rr;
            const int o = max(0, 2 * k - rr) + lane;
            const bool ok = rr < nrow && o < min(w, 2 * k + ncol - rr) && r >= 0 && r < n;
            v[u] = ok ? in[r * w + o] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < B2C_ROWS; ++u) {
            const int rr = rb + u * NWV;
            const int o = max(0, 2 * k - rr) + lane;
            if (rr < nrow && o < min(w, 2 * k + ncol - rr)) lds[(rr - 2 * k + o) * ls + 2 * k - o] = (T)v[u];
        }
    }
    __syncthreads();
    const int64_t tot = (int64_t)ncol * w;
    T* const dst = out + c0 * w;
    T* const dst2 = out2 ? out2 + c0 * w : nullptr;
    // element e of the run is column e / w, diagonal e % w (stepped, no division per element)
    int cc = (int)threadIdx.x / w, d = (int)threadIdx.x - cc * w;
    const int qs = B2C_THREADS / w, rs = B2C_THREADS - qs * w;
    for (int64_t e = threadIdx.x; e < tot; e += B2C_THREADS) {
        const T v = lds[cc * ls + d];
        dst[e] = v;
        if (dst2) dst2[e] = v;
        cc += qs;
        d += rs;
        if (d >= w) { d -= w; ++cc; }
    }
}

// Persistent, software-pipelined form of the column-tile conversion (tiles
// of tc <= 64 columns, k small enough that a tile's tc + 2k input rows fit
// RPW rows per wavefront): each workgroup converts a run of tiles, and the
// row segments of its next tile are loaded into registers while the current
// tile goes from LDS to memory, so the loads of one tile and the stores of
// the previous one overlap (the non-persistent kernel's workgroups, one per
// CU for its LDS, alternate between the two).  Tiles are dealt to the
// workgroups XCD by XCD (workgroup b runs on XCD b mod 8): each XCD converts
// one contiguous eighth of the band, so the cache lines two neighbouring
// tiles share at their segment ends come through one L2.  The stores are
// 16-byte vectors of the contiguous output run.  Same values as
// band_to_colmajor_cols_kernel, element for element.
// ---------------------------------------------------------------------------
template <class T, int RPW>
__global__ __launch_bounds__(B2C_THREADS) void band_to_colmajor_pipe_kernel(int64_t n, int k, int tc, int64_t nchunks,
                                                                        const double* __restrict__ in,
                                                                        T* __restrict__ out, T* __restrict__ out2)
{
    extern __shared__ unsigned char smem_raw[];
    T* const lds = reinterpret_cast<T*>(smem_raw);
    constexpr int NWV = B2C_THREADS / 64;
    constexpr int V = 16 / (int)sizeof(T);   // elements per 16-byte store
    const int w = 2 * k + 1, ls = w + 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // tiles of this workgroup: XCD x = b mod 8 owns tiles [lo, hi), its
    // per = G / 8 workgroups take every per-th of them
    const int G = gridDim.x, b = blockIdx.x;
    const int x = b & 7, slot = b >> 3, per = G >> 3;
    const int64_t lo = nchunks * x / 8, hi = nchunks * (x + 1) / 8;
    int64_t ch = lo + slot;
    if (ch >= hi) return;   // uniform over the workgroup

    // Row rr of a tile (input row c0 - k + rr) gives the tile the segment
    // o in [ob, ob + len) with ob = max(0, 2k - rr): lane l loads offset
    // ob + l, so every per-row quantity but the lane is uniform, and the
    // loads are buffer loads relative to the tile's first input row (lanes
    // past the segment, and rows outside the matrix, read zero through an
    // offset past the buffer's end).
    double v[RPW];
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    auto load = [&](int64_t c) {
        const int64_t c0 = c * tc;
        const int ncol = (int)min((int64_t)tc, n - c0);
        const int nrow = ncol + 2 * k;
        const int64_t rlo = max((int64_t)0, c0 - k), rhi = min(n, c0 + ncol + k);
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(in + rlo * w), 0, (int)((rhi - rlo) * w * (int64_t)sizeof(double)), 0x00020000);
#pragma unroll
        for (int u = 0; u < RPW; ++u) {
            const int rr = wv + NWV * u;
            const int64_t r = c0 - k + rr;
            const int ob = max(0, 2 * k - rr);
            const int len = min(w, 2 * k + ncol - rr) - ob;
            const bool row_ok = rr < nrow && r >= 0 && r < n;
            const int base = row_ok ? (int)((r - rlo) * w + ob) : 0;
            const int off = (row_ok && lane < len) ? (base + lane) * (int)sizeof(double) : (int)0x80000000;
            const auto q = __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0);
            v[u] = __longlong_as_double((long long)(((unsigned long long)q[1] << 32) | q[0]));
        }
    };
    load(ch);
    for (;;) {
        const int64_t c0 = ch * tc;
        const int ncol = (int)min((int64_t)tc, n - c0);
        const int nrow = ncol + 2 * k;
        // segment element ob + l of row rr is column rr - 2k + ob + l,
        // diagonal 2k - ob - l: LDS index (rr - 2k + ob) ls + 2k - ob + l w
#pragma unroll
        for (int u = 0; u < RPW; ++u) {
            const int rr = wv + NWV * u;
            const int ob = max(0, 2 * k - rr);
            const int len = min(w, 2 * k + ncol - rr) - ob;
            if (rr < nrow && lane < len) lds[(rr - 2 * k + ob) * ls + 2 * k - ob + lane * w] = (T)v[u];
        }
        __syncthreads();
        const int64_t nx = ch + per;
        if (nx < hi) load(nx);
        // the run of ncol * w elements: element e is column e / w, diagonal
        // e % w, at LDS index e + e / w (column stride w + 1)
        const int tot = ncol * w;
        T* const dst = out + c0 * w;
        T* const dst2 = out2 ? out2 + c0 * w : nullptr;
        const int nvec = tot / V;
        constexpr int STEP = B2C_THREADS * V;
        const int qs = STEP / w, rs = STEP - qs * w;
        int e = (int)threadIdx.x * V;
        int cc = e / w, d = e - cc * w;
        for (int q = threadIdx.x; q < nvec; q += B2C_THREADS) {
            T val[V];
#pragma unroll
            for (int i = 0; i < V; ++i) {
                const int di = d + i;
                val[i] = lds[cc * ls + di + (di >= w ? 1 : 0)];
            }
            if constexpr (V == 4) {
                const float4 f = make_float4(val[0], val[1], val[2], val[3]);
                *reinterpret_cast<float4*>(dst + e) = f;
                if (dst2) *reinterpret_cast<float4*>(dst2 + e) = f;
            } else {
                const double2 f = make_double2(val[0], val[1]);
                *reinterpret_cast<double2*>(dst + e) = f;
                if (dst2) *reinterpret_cast<double2*>(dst2 + e) = f;
            }
            e += STEP;
            cc += qs;
            d += rs;
            if (d >= w) { d -= w; ++cc; }
        }
        for (int t = nvec * V + (int)threadIdx.x; t < tot; t += B2C_THREADS) {
            const int tc_ = t / w;
            const T val = lds[t + tc_];
            dst[t] = val;
            if (dst2) dst2[t] = val;
        }
        __syncthreads();   // the LDS tile is free for the next scatter
        if (nx >= hi) break;
        ch = nx;
    }
}

// ---------------------------------------------------------------------------
// CSR entries into the zero-filled column-major band, cast to the band type
// (the COO scatter copyFromCOOMatrixToBandedMatrix,
// sap/device/data_transfer.cuh:421-448, driven from the CSR rows: each
// entry lands on its own band slot, so the order does not matter).
// ---------------------------------------------------------------------------
template <class Tin, class T>
__global__ void csr_to_band_kernel(int64_t n, int k, const int* __restrict__ rp, const int* __restrict__ ci,
                                   const Tin* __restrict__ vals, T* __restrict__ B)
{
    const int64_t w = 2 * k + 1;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
        const int e1 = rp[r + 1];
        for (int e = rp[r]; e < e1; ++e) {
            const int64_t c = ci[e];
            B[c * w + k + r - c] = (T)vals[e];
        }
    }
}

// max |row - col| over the stored entries (the half-bandwidth of
// convertToBandedMatrix(CSR), sap/precond.h:2253-2257)
__global__ void csr_bandwidth_kernel(int64_t n, const int* __restrict__ rp, const int* __restrict__ ci,
                                     int* __restrict__ out)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int m = 0;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride)
        for (int e = rp[r]; e < rp[r + 1]; ++e) m = max(m, abs((int)r - ci[e]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// ---------------------------------------------------------------------------
// New values of a tracked matrix into the zero-filled band: entry e goes to
// band slot map[e] (-1: not in the band) times its DB scale factor, in the
// band type (Precond::update, sap/precond.h:751-775: scatter_if of
// entries * scaleMap through bandedMatMap).
// ---------------------------------------------------------------------------
template <class T>
__global__ void update_band_kernel(int64_t nnz, const double* __restrict__ entries,
                                   const int64_t* __restrict__ map, const T* __restrict__ scale, T* __restrict__ B)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += stride) {
        const int64_t slot = map[e];
        if (slot < 0) continue;
        T v = (T)entries[e];
        if (scale) v = v * scale[e];
        B[slot] = v;
    }
}

// ---------------------------------------------------------------------------
// y = A x for FP64 CSR (the SpmvCusp functor, sap/spmv.h:83-129): G lanes per
// row, each summing its entries in order, then a fixed shuffle tree.
// ---------------------------------------------------------------------------
template <int G>
__global__ __launch_bounds__(256) void csr_spmv_kernel(int64_t n, const int* __restrict__ rp,
                                                       const int* __restrict__ ci, const double* __restrict__ val,
                                                       const double* __restrict__ x, double* __restrict__ y)
{
    const int lane = threadIdx.x & (G - 1);
    const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
    double s = 0.0;
    if (row < n)
        for (int j = rp[row] + lane; j < rp[row + 1]; j += G) s = fma(val[j], x[ci[j]], s);
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) s += __shfl_down(s, off, G);
    if (lane == 0 && row < n) y[row] = s;
}

// csr_spmv_kernel for column blockIdx.y of X / Y (same arithmetic per
// column); columns whose gate flag is 0 are skipped
template <int G>
__global__ __launch_bounds__(256) void csr_spmm_kernel(int64_t n, const int* __restrict__ rp,
                                                       const int* __restrict__ ci, const double* __restrict__ val,
                                                       const double* __restrict__ X, int64_t ldx,
                                                       double* __restrict__ Y, int64_t ldy,
                                                       const int* __restrict__ gate)
{
    if (gate && !gate[blockIdx.y]) return;
    const double* x = X + blockIdx.y * ldx;
    double* y = Y + blockIdx.y * ldy;
    const int lane = threadIdx.x & (G - 1);
    const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
    double s = 0.0;
    if (row < n)
        for (int j = rp[row] + lane; j < rp[row + 1]; j += G) s = fma(val[j], x[ci[j]], s);
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) s += __shfl_down(s, off, G);
    if (lane == 0 && row < n) y[row] = s;
}

// ---------------------------------------------------------------------------
// Vector permutations of the reordered preconditioner, in the
// preconditioner type as the reference computes them on PrecVectors:
//   leftTrans  (scaleAndPermute, sap/precond.h:1582-1598): w[perm[i]] = v[i] * s[i]
//   rightTrans (permuteAndScale, sap/precond.h:1560-1576): w[perm[i]] = v[i] * s[perm[i]]
// (scale may be null: permute only).
// ---------------------------------------------------------------------------
template <class T, bool SCALE_AT_TARGET>
__global__ void permute_scale_kernel(int64_t n, const double* __restrict__ v, const int* __restrict__ perm,
                                     const T* __restrict__ scale, double* __restrict__ w, int64_t ldv, int64_t ldw,
                                     const int* __restrict__ gate)
{
    if (gate && !gate[blockIdx.y]) return;   // column blockIdx.y
    v += blockIdx.y * ldv;
    w += blockIdx.y * ldw;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int p = perm[i];
        T x = (T)v[i];
        if (scale) x = x * scale[SCALE_AT_TARGET ? p : i];
        w[p] = (double)x;
    }
}

}  // namespace

void launch_gen_banded(int64_t n, int k, double dom, uint64_t seed, double* band, hipStream_t st)
{
    const int bs = 256;
    hipLaunchKernelGGL(gen_banded_kernel, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, st, n, k, dom, seed,
                       band);
    SAP_LAUNCH_CHECK();
}

void launch_gen_vector(int64_t n, uint64_t seed, double* x, hipStream_t st)
{
    const int bs = 256;
    hipLaunchKernelGGL(gen_vector_kernel, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, st, n, seed, x);
    SAP_LAUNCH_CHECK();
}

void launch_band_spmv(int64_t n, int k, const double* band, const double* x, double* y, hipStream_t st)
{
    const int w = 2 * k + 1;
    const int bs = 256;
    if (w >= 192) {
        constexpr int RPW = 4;   // 4 rows per wavefront: 0.661 against 0.696-0.721 ms at n = 1M, k = 200
        const int64_t groups = (n + RPW - 1) / RPW;
        const int64_t blocks = (groups * 64 + bs - 1) / bs;
        hipLaunchKernelGGL((band_spmv_kernel<64, 8, RPW>), dim3((unsigned)blocks), dim3(bs), 0, st, n, k, band, x, y);
    } else if (w >= 48) {
        const int64_t blocks = (n * 16 + bs - 1) / bs;
        hipLaunchKernelGGL((band_spmv_kernel<16, 12, 1>), dim3((unsigned)blocks), dim3(bs), 0, st, n, k, band, x, y);
    } else {
        const int64_t blocks = (n * 4 + bs - 1) / bs;
        hipLaunchKernelGGL((band_spmv_kernel<4, 12, 1>), dim3((unsigned)blocks), dim3(bs), 0, st, n, k, band, x, y);
    }
    SAP_LAUNCH_CHECK();
}

template <int RB>
void launch_band_spmm_rb(int64_t n, int k, const double* band, const double* X, int64_t ldx, double* Y, int64_t ldy,
                         int nr, const int* gate, hipStream_t st)
{
    const int w = 2 * k + 1;
    const int bs = 256;
    const unsigned chunks = (unsigned)((nr + RB - 1) / RB);
    if (w >= 192) {
        const int64_t blocks = (n * 64 + bs - 1) / bs;
        hipLaunchKernelGGL((band_spmm_kernel<64, 8, RB>), dim3((unsigned)blocks, chunks), dim3(bs), 0, st, n, k, band,
                           X, ldx, Y, ldy, nr, gate);
    } else if (w >= 48) {
        const int64_t blocks = (n * 16 + bs - 1) / bs;
        hipLaunchKernelGGL((band_spmm_kernel<16, 12, RB>), dim3((unsigned)blocks, chunks), dim3(bs), 0, st, n, k,
                           band, X, ldx, Y, ldy, nr, gate);
    } else {
        const int64_t blocks = (n * 4 + bs - 1) / bs;
        hipLaunchKernelGGL((band_spmm_kernel<4, 12, RB>), dim3((unsigned)blocks, chunks), dim3(bs), 0, st, n, k, band,
                           X, ldx, Y, ldy, nr, gate);
    }
    SAP_LAUNCH_CHECK();
}

void launch_band_spmm(int64_t n, int k, const double* band, const double* X, int64_t ldx, double* Y, int64_t ldy,
                      int nr, const int* gate, hipStream_t st)
{
    // columns per workgroup row: the band is read once per RB columns
    if (nr > 8) launch_band_spmm_rb<16>(n, k, band, X, ldx, Y, ldy, nr, gate, st);
    else if (nr > 4) launch_band_spmm_rb<8>(n, k, band, X, ldx, Y, ldy, nr, gate, st);
    else launch_band_spmm_rb<4>(n, k, band, X, ldx, Y, ldy, nr, gate, st);
}

void launch_csr_spmm(int64_t n, int64_t nnz, const int* rp, const int* ci, const double* val, const double* X,
                     int64_t ldx, double* Y, int64_t ldy, int nr, const int* gate, hipStream_t st)
{
    const int bs = 256;
    const double avg = n > 0 ? (double)nnz / (double)n : 0.0;
    if (avg >= 24) {
        const int64_t blocks = (n * 32 + bs - 1) / bs;
        hipLaunchKernelGGL(csr_spmm_kernel<32>, dim3((unsigned)blocks, nr), dim3(bs), 0, st, n, rp, ci, val, X, ldx,
                           Y, ldy, gate);
    } else if (avg >= 6) {
        const int64_t blocks = (n * 8 + bs - 1) / bs;
        hipLaunchKernelGGL(csr_spmm_kernel<8>, dim3((unsigned)blocks, nr), dim3(bs), 0, st, n, rp, ci, val, X, ldx, Y,
                           ldy, gate);
    } else {
        const int64_t blocks = (n * 2 + bs - 1) / bs;
        hipLaunchKernelGGL(csr_spmm_kernel<2>, dim3((unsigned)blocks, nr), dim3(bs), 0, st, n, rp, ci, val, X, ldx, Y,
                           ldy, gate);
    }
    SAP_LAUNCH_CHECK();
}

// compute units of the current device (cached per device)
int device_cus()
{
    static std::atomic<int> cus[64] = {};
    int dev = 0;
    SAP_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) dev = 0;
    int v = cus[dev].load(std::memory_order_relaxed);
    if (v == 0) {
        SAP_HIP_CHECK(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev));
        if (v <= 0) v = 256;
        cus[dev].store(v, std::memory_order_relaxed);
    }
    return v;
}

template <class T>
int band_to_colmajor_max_tile(int k)
{
    // column tiles of up to 64 columns within 150 KiB of LDS
    constexpr size_t budget = 150 * 1024;
    return (int)std::min<size_t>(64, budget / ((size_t)(2 * k + 2) * sizeof(T)));
}

template <class T>
void launch_band_to_colmajor(int64_t n, int k, const double* band_rm, T* B, hipStream_t st, T* B2, int tile)
{
    const int w = 2 * k + 1;
    // bands too wide for 8 columns of LDS use the diagonal-box kernel
    int tc = band_to_colmajor_max_tile<T>(k);
    if (tile >= 8 && tile < tc) tc = tile;
    const int rows = tc + 2 * k;   // input rows of one tile
    constexpr int V = 16 / (int)sizeof(T);
    const bool aligned = ((uintptr_t)B % 16 == 0) && ((uintptr_t)B2 % 16 == 0);
    if (tc >= 8 && rows <= 32 * (B2C_THREADS / 64) && tc % V == 0 && w >= V && aligned) {
        const size_t lds = (size_t)tc * (w + 1) * sizeof(T);
        const int64_t nchunks = (n + tc - 1) / tc;
        const int G = 8 * std::max(1, std::min(device_cus() / 8, (int)((nchunks + 7) / 8)));
        const int rpw = (rows + B2C_THREADS / 64 - 1) / (B2C_THREADS / 64);
        auto go = [&](auto kern) {
            SAP_HIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            hipLaunchKernelGGL(kern, dim3((unsigned)G), dim3(B2C_THREADS), lds, st, n, k, tc, nchunks, band_rm, B, B2);
        };
        if (rpw <= 8)       go(band_to_colmajor_pipe_kernel<T, 8>);
        else if (rpw <= 16) go(band_to_colmajor_pipe_kernel<T, 16>);
        else                go(band_to_colmajor_pipe_kernel<T, 32>);
    } else if (tc >= 8) {
        const size_t lds = (size_t)tc * (w + 1) * sizeof(T);
        SAP_HIP_CHECK(hipFuncSetAttribute((const void*)band_to_colmajor_cols_kernel<T>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(band_to_colmajor_cols_kernel<T>, dim3((unsigned)((n + tc - 1) / tc)), dim3(B2C_THREADS),
                           lds, st, n, k, tc, band_rm, B, B2);
    } else {
        dim3 grid((unsigned)((n + TC - 1) / TC), (unsigned)((w + TD - 1) / TD));
        hipLaunchKernelGGL(band_to_colmajor_kernel<T>, grid, dim3(256), 0, st, n, k, band_rm, B, B2);
    }
    SAP_LAUNCH_CHECK();
}

template void launch_band_to_colmajor<float>(int64_t, int, const double*, float*, hipStream_t, float*, int);
template void launch_band_to_colmajor<double>(int64_t, int, const double*, double*, hipStream_t, double*, int);
template int band_to_colmajor_max_tile<float>(int);
template int band_to_colmajor_max_tile<double>(int);

// The first k and last k columns of partitions p0 .. p1-1 (each column one
// contiguous run of 2k+1 band entries), copied from in to out.
// (workgroup (p - p0, y) copies columns y, y + gridDim.y, ... of the 2k;
// each column's 2k+1 entries are contiguous)
template <class T>
__global__ void copy_partition_edges_kernel(const T* __restrict__ in, T* __restrict__ out, int k, PartInfo pi, int p0)
{
    const int64_t w = 2 * (int64_t)k + 1;
    int64_t s, len;
    pi.bounds(p0 + (int)blockIdx.x, s, len);
    for (int c = blockIdx.y; c < 2 * k; c += gridDim.y) {
        const int64_t col = c < k ? s + c : s + len - 2 * k + c;
        const T* src = in + col * w;
        T* dst = out + col * w;
        for (int e = threadIdx.x; e < w; e += blockDim.x) dst[e] = src[e];
    }
}

template <class T>
void launch_copy_partition_edges(const T* in, T* out, int k, PartInfo pi, int p0, int p1, hipStream_t st)
{
    if (k <= 0 || p1 <= p0) return;
    hipLaunchKernelGGL(copy_partition_edges_kernel<T>, dim3((unsigned)(p1 - p0), (unsigned)std::min(2 * k, 64)),
                       dim3(256), 0, st, in, out, k, pi, p0);
    SAP_LAUNCH_CHECK();
}
template void launch_copy_partition_edges<float>(const float*, float*, int, PartInfo, int, int, hipStream_t);
template void launch_copy_partition_edges<double>(const double*, double*, int, PartInfo, int, int, hipStream_t);

template <class Tin, class T>
void launch_csr_to_band(int64_t n, int k, const int* rp, const int* ci, const Tin* vals, T* B, hipStream_t st)
{
    int64_t g = (n + 255) / 256;
    if (g > 8192) g = 8192;
    hipLaunchKernelGGL((csr_to_band_kernel<Tin, T>), dim3((unsigned)g), dim3(256), 0, st, n, k, rp, ci, vals, B);
    SAP_LAUNCH_CHECK();
}
template void launch_csr_to_band<double, float>(int64_t, int, const int*, const int*, const double*, float*,
                                                hipStream_t);
template void launch_csr_to_band<double, double>(int64_t, int, const int*, const int*, const double*, double*,
                                                 hipStream_t);
template void launch_csr_to_band<float, float>(int64_t, int, const int*, const int*, const float*, float*,
                                               hipStream_t);

template <class T>
void launch_update_band(int64_t nnz, const double* entries, const int64_t* map, const T* scale, T* B, hipStream_t st)
{
    if (nnz <= 0) return;
    int64_t g = (nnz + 255) / 256;
    if (g > 8192) g = 8192;
    hipLaunchKernelGGL(update_band_kernel<T>, dim3((unsigned)g), dim3(256), 0, st, nnz, entries, map, scale, B);
    SAP_LAUNCH_CHECK();
}
template void launch_update_band<float>(int64_t, const double*, const int64_t*, const float*, float*, hipStream_t);
template void launch_update_band<double>(int64_t, const double*, const int64_t*, const double*, double*,
                                         hipStream_t);

void launch_csr_bandwidth(int64_t n, const int* rp, const int* ci, int* out, hipStream_t st)
{
    int64_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(csr_bandwidth_kernel, dim3((unsigned)g), dim3(256), 0, st, n, rp, ci, out);
    SAP_LAUNCH_CHECK();
}

void launch_csr_spmv(int64_t n, int64_t nnz, const int* rp, const int* ci, const double* val, const double* x,
                     double* y, hipStream_t st)
{
    const int bs = 256;
    const double avg = n > 0 ? (double)nnz / (double)n : 0.0;
    if (avg >= 24) {
        const int64_t blocks = (n * 32 + bs - 1) / bs;
        hipLaunchKernelGGL(csr_spmv_kernel<32>, dim3((unsigned)blocks), dim3(bs), 0, st, n, rp, ci, val, x, y);
    } else if (avg >= 6) {
        const int64_t blocks = (n * 8 + bs - 1) / bs;
        hipLaunchKernelGGL(csr_spmv_kernel<8>, dim3((unsigned)blocks), dim3(bs), 0, st, n, rp, ci, val, x, y);
    } else {
        const int64_t blocks = (n * 2 + bs - 1) / bs;
        hipLaunchKernelGGL(csr_spmv_kernel<2>, dim3((unsigned)blocks), dim3(bs), 0, st, n, rp, ci, val, x, y);
    }
    SAP_LAUNCH_CHECK();
}

template <class T>
void launch_permute_scale(int64_t n, const double* v, const int* perm, const T* scale, bool scale_at_target,
                          double* w, hipStream_t st, int nr, int64_t ldv, int64_t ldw, const int* gate)
{
    int64_t g = (n + 255) / 256;
    if (g > 8192) g = 8192;
    if (scale_at_target)
        hipLaunchKernelGGL((permute_scale_kernel<T, true>), dim3((unsigned)g, nr), dim3(256), 0, st, n, v, perm, scale,
                           w, ldv, ldw, gate);
    else
        hipLaunchKernelGGL((permute_scale_kernel<T, false>), dim3((unsigned)g, nr), dim3(256), 0, st, n, v, perm,
                           scale, w, ldv, ldw, gate);
    SAP_LAUNCH_CHECK();
}
template void launch_permute_scale<float>(int64_t, const double*, const int*, const float*, bool, double*, hipStream_t,
                                          int, int64_t, int64_t, const int*);
template void launch_permute_scale<double>(int64_t, const double*, const int*, const double*, bool, double*,
                                           hipStream_t, int, int64_t, int64_t, const int*);

}  // namespace sap_amd
