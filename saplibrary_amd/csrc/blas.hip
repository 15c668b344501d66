// FP64 vector kernels for the Krylov loops (the cusp::blas calls of
// sap/bicgstab2.h, sap/minres.h and sap/monitor.h).  Reductions are
// deterministic: a fixed grid writes per-block partial sums, and a single
// block combines them in a fixed order.
#include "sap_internal.hpp"

#include <algorithm>

namespace sap_amd {

namespace {

constexpr int RED_BLOCKS = 512;
constexpr int RED_THREADS = 256;

__device__ inline double block_reduce_sum(double v)
{
    __shared__ double warp_sums[RED_THREADS / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) warp_sums[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += warp_sums[i];
    }
    return s;
}

// Up to three dot products in one pass (q = 0 .. nd-1: a_q . b_q); each
// keeps the summation order of a single one, so the values do not depend on
// what it is computed with.
struct DotArgs {
    const double* a[3];
    const double* b[3];
};

template <int ND>
__global__ __launch_bounds__(RED_THREADS) void dot_partial_kernel(int64_t n, DotArgs d, double* __restrict__ partials)
{
    double s[ND];
#pragma unroll
    for (int q = 0; q < ND; ++q) s[q] = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
#pragma unroll
        for (int q = 0; q < ND; ++q) s[q] = fma(d.a[q][i], d.b[q][i], s[q]);
    }
#pragma unroll
    for (int q = 0; q < ND; ++q) {
        const double t = block_reduce_sum(s[q]);
        if (threadIdx.x == 0) partials[q * RED_BLOCKS + blockIdx.x] = t;
        __syncthreads();   // the reduction's LDS is reused by the next q
    }
}

// out[q] = sum of the partials of dot q; out may be host memory mapped into
// the device's address space (read by the host after a stream synchronize)
__global__ __launch_bounds__(RED_THREADS) void reduce_final_kernel(int m, int nd, const double* __restrict__ partials,
                                                                   double* __restrict__ out)
{
    for (int q = 0; q < nd; ++q) {
        double s = 0.0;
        for (int i = threadIdx.x; i < m; i += blockDim.x) s += partials[q * m + i];
        s = block_reduce_sum(s);
        if (threadIdx.x == 0) out[q] = s;
        __syncthreads();
    }
}

struct SpreadArgs {
    const double* src;
    double* dst[4];
    double* zero[4];
    int nd, nz;
};

// one element of z = alpha x + beta y (shared by the single and batched
// kernels, so both round the same way)
__device__ __forceinline__ double axpby_elem(double alpha, double x, double beta, double y)
{
    return alpha * x + beta * y;
}

__global__ void axpby_kernel(int64_t n, double alpha, const double* __restrict__ x, double beta,
                             const double* __restrict__ y, double* __restrict__ z)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        z[i] = axpby_elem(alpha, x[i], beta, y[i]);
}

// z = alpha x + beta y with alpha = sign * (num / *den) read on the device
// (the same division the host would make, so the same alpha)
__global__ void axpby_ratio_kernel(int64_t n, double num, const double* __restrict__ den, double sign,
                                   const double* __restrict__ x, double beta, const double* __restrict__ y,
                                   double* __restrict__ z)
{
    // a zero gamma stops the solve (code -11) right after this update; the
    // updated vectors are then dead, and alpha = 0 keeps them finite
    const double d = *den;
    const double alpha = d == 0.0 ? 0.0 : sign * (num / d);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        z[i] = axpby_elem(alpha, x[i], beta, y[i]);
}

// ------------------------------------------------------ batched (columns)
// Column c of a batch vector spans [off[c], off[c] + len[c]); grid.y is the
// column.  Each column gets the arithmetic of the single-vector kernel above
// on that column alone (same grid-stride partition, same reduction tree), so
// the same bits.  A column whose gate flag is 0 is skipped.
struct BDotPtrs {
    const double* a[3];
    const double* b[3];
};

template <int ND>
__global__ __launch_bounds__(RED_THREADS) void bdot_partial_kernel(const int64_t* __restrict__ off,
                                                                   const int* __restrict__ len, BDotPtrs d,
                                                                   const int* __restrict__ gate,
                                                                   double* __restrict__ partials)
{
    const int c = blockIdx.y;
    if (gate && !gate[c]) return;
    const int64_t n = len[c];
    const int64_t o = off[c];
    double s[ND];
#pragma unroll
    for (int q = 0; q < ND; ++q) s[q] = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
#pragma unroll
        for (int q = 0; q < ND; ++q) s[q] = fma(d.a[q][o + i], d.b[q][o + i], s[q]);
    }
#pragma unroll
    for (int q = 0; q < ND; ++q) {
        const double t = block_reduce_sum(s[q]);
        if (threadIdx.x == 0) partials[((size_t)c * ND + q) * RED_BLOCKS + blockIdx.x] = t;
        __syncthreads();
    }
}

// out(q, c) at out[c * out_stride + q]
__global__ __launch_bounds__(RED_THREADS) void bdot_final_kernel(int nd, const double* __restrict__ partials,
                                                                 const int* __restrict__ gate, double* __restrict__ out,
                                                                 int out_stride)
{
    const int c = blockIdx.x;
    if (gate && !gate[c]) return;
    const double* pc = partials + (size_t)c * nd * RED_BLOCKS;
    for (int q = 0; q < nd; ++q) {
        double s = 0.0;
        for (int i = threadIdx.x; i < RED_BLOCKS; i += blockDim.x) s += pc[q * RED_BLOCKS + i];
        s = block_reduce_sum(s);
        if (threadIdx.x == 0) out[(size_t)c * out_stride + q] = s;
        __syncthreads();
    }
}

// coefficient of column c: sign * S[c * s_stride + field] (field >= 0) or cst
__device__ __forceinline__ double bcoef(const BCoef& k, const double* S, int s_stride, int c)
{
    return k.field >= 0 ? k.sign * S[(size_t)c * s_stride + k.field] : k.cst;
}

__global__ void baxpby_kernel(const int64_t* __restrict__ off, const int* __restrict__ len, BCoef ka,
                              const double* __restrict__ x, BCoef kb, const double* __restrict__ y,
                              double* __restrict__ z, const double* __restrict__ S, int s_stride,
                              const int* __restrict__ gate)
{
    const int c = blockIdx.y;
    if (gate && !gate[c]) return;
    const double alpha = bcoef(ka, S, s_stride, c);
    const double beta = bcoef(kb, S, s_stride, c);
    const int64_t n = len[c];
    const int64_t o = off[c];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        z[o + i] = axpby_elem(alpha, x[o + i], beta, y[o + i]);
}

__global__ void bcopy_kernel(const int64_t* __restrict__ off, const int* __restrict__ len,
                             const double* __restrict__ x, double* __restrict__ y, const int* __restrict__ gate)
{
    const int c = blockIdx.y;
    if (gate && !gate[c]) return;
    const int64_t n = len[c];
    const int64_t o = off[c];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) y[o + i] = x[o + i];
}

// dst[q] = src (q < nd), zero[q] = 0 (q < nz), per gated column
__global__ void bspread_kernel(const int64_t* __restrict__ off, const int* __restrict__ len, SpreadArgs a,
                               const int* __restrict__ gate)
{
    const int c = blockIdx.y;
    if (gate && !gate[c]) return;
    const int64_t n = len[c];
    const int64_t o = off[c];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (a.nd > 0) {
            const double v = a.src[o + i];
            for (int q = 0; q < a.nd; ++q) a.dst[q][o + i] = v;
        }
        for (int q = 0; q < a.nz; ++q) a.zero[q][o + i] = 0.0;
    }
}

__global__ void copy_kernel(int64_t n, const double* __restrict__ x, double* __restrict__ y)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) y[i] = x[i];
}

__global__ void fill_kernel(int64_t n, double v, double* __restrict__ x)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) x[i] = v;
}


// dst[q] = src for q < nd and zero[q] = 0 for q < nz, in one pass
__global__ void spread_kernel(int64_t n, SpreadArgs a)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (a.nd > 0) {
            const double v = a.src[i];
            for (int q = 0; q < a.nd; ++q) a.dst[q][i] = v;
        }
        for (int q = 0; q < a.nz; ++q) a.zero[q][i] = 0.0;
    }
}

// Streaming read of n doubles (16-byte non-temporal loads, a per-block sum
// kept so the loads are not dead): the device's practical read bandwidth
// for the kernels that stream a band once.
typedef double dvec2 __attribute__((ext_vector_type(2)));

__global__ __launch_bounds__(256) void stream_read_kernel(int64_t n2, const dvec2* __restrict__ x,
                                                          double* __restrict__ partials)
{
    double s = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride) {
        const dvec2 v = __builtin_nontemporal_load(x + i);
        s += v.x + v.y;
    }
    s = block_reduce_sum(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

inline unsigned grid_for(int64_t n)
{
    int64_t g = (n + 255) / 256;
    if (g > 8192) g = 8192;
    if (g < 1) g = 1;
    return (unsigned)g;
}

}  // namespace

int dot_partials_size() { return 3 * RED_BLOCKS; }

int bdot_partials_size(int cols) { return 3 * RED_BLOCKS * cols; }

double stream_read_best_ms(const double* x, int64_t n, int reps, hipStream_t st)
{
    const int blocks = 256 * 8;
    DevBuf<double> part(blocks);
    hipEvent_t a, b;
    SAP_HIP_CHECK(hipEventCreate(&a));
    SAP_HIP_CHECK(hipEventCreate(&b));
    double best = -1.0;
    for (int r = 0; r < reps + 1; ++r) {   // the first launch is a warm-up
        SAP_HIP_CHECK(hipEventRecord(a, st));
        hipLaunchKernelGGL(stream_read_kernel, dim3(blocks), dim3(256), 0, st, n / 2,
                           reinterpret_cast<const dvec2*>(x), part.get());
        SAP_HIP_CHECK(hipEventRecord(b, st));
        SAP_HIP_CHECK(hipEventSynchronize(b));
        float ms = 0.f;
        SAP_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
        if (r > 0 && (best < 0 || ms < best)) best = ms;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return best;
}

// the batched grids: up to 8192 / cols workgroups per column (at least 64)
inline unsigned bgrid_for(int64_t maxlen, int cols)
{
    int64_t g = (maxlen + 255) / 256;
    const int64_t cap = std::max<int64_t>(64, 8192 / std::max(cols, 1));
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (unsigned)g;
}

void launch_bdots(const BatchCols& bc, int nd, const double* const* a, const double* const* b, const int* gate,
                  double* d_partials, double* out, int out_stride, hipStream_t st)
{
    if (nd < 1 || nd > 3) throw Error(SAP_ERR_BADARG, "launch_bdots: 1 to 3 products");
    BDotPtrs d{};
    for (int q = 0; q < nd; ++q) { d.a[q] = a[q]; d.b[q] = b[q]; }
    const dim3 grid(RED_BLOCKS, bc.cols);
    if (nd == 1) hipLaunchKernelGGL(bdot_partial_kernel<1>, grid, dim3(RED_THREADS), 0, st, bc.off, bc.len, d, gate, d_partials);
    else if (nd == 2) hipLaunchKernelGGL(bdot_partial_kernel<2>, grid, dim3(RED_THREADS), 0, st, bc.off, bc.len, d, gate, d_partials);
    else hipLaunchKernelGGL(bdot_partial_kernel<3>, grid, dim3(RED_THREADS), 0, st, bc.off, bc.len, d, gate, d_partials);
    SAP_LAUNCH_CHECK();
    hipLaunchKernelGGL(bdot_final_kernel, dim3(bc.cols), dim3(RED_THREADS), 0, st, nd, d_partials, gate, out,
                       out_stride);
    SAP_LAUNCH_CHECK();
}

void launch_baxpby(const BatchCols& bc, BCoef ka, const double* x, BCoef kb, const double* y, double* z,
                   const double* S, int s_stride, const int* gate, hipStream_t st)
{
    hipLaunchKernelGGL(baxpby_kernel, dim3(bgrid_for(bc.maxlen, bc.cols), bc.cols), dim3(256), 0, st, bc.off, bc.len,
                       ka, x, kb, y, z, S, s_stride, gate);
    SAP_LAUNCH_CHECK();
}

void launch_bcopy(const BatchCols& bc, const double* x, double* y, const int* gate, hipStream_t st)
{
    hipLaunchKernelGGL(bcopy_kernel, dim3(bgrid_for(bc.maxlen, bc.cols), bc.cols), dim3(256), 0, st, bc.off, bc.len,
                       x, y, gate);
    SAP_LAUNCH_CHECK();
}

void launch_bspread(const BatchCols& bc, const double* src, double* const* dst, int nd, double* const* zero, int nz,
                    const int* gate, hipStream_t st)
{
    if (nd < 0 || nd > 4 || nz < 0 || nz > 4 || (nd > 0 && !src))
        throw Error(SAP_ERR_BADARG, "launch_bspread: at most four copies and four zero fills");
    SpreadArgs a{};
    a.src = src;
    for (int q = 0; q < nd; ++q) a.dst[q] = dst[q];
    for (int q = 0; q < nz; ++q) a.zero[q] = zero[q];
    a.nd = nd;
    a.nz = nz;
    hipLaunchKernelGGL(bspread_kernel, dim3(bgrid_for(bc.maxlen, bc.cols), bc.cols), dim3(256), 0, st, bc.off,
                       bc.len, a, gate);
    SAP_LAUNCH_CHECK();
}

void launch_dots(int64_t n, int nd, const double* const* a, const double* const* b, double* d_partials,
                 double* out, hipStream_t st)
{
    if (nd < 1 || nd > 3) throw Error(SAP_ERR_BADARG, "launch_dots: 1 to 3 products");
    DotArgs d{};
    for (int q = 0; q < nd; ++q) { d.a[q] = a[q]; d.b[q] = b[q]; }
    if (nd == 1) hipLaunchKernelGGL(dot_partial_kernel<1>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, n, d, d_partials);
    else if (nd == 2) hipLaunchKernelGGL(dot_partial_kernel<2>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, n, d, d_partials);
    else hipLaunchKernelGGL(dot_partial_kernel<3>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, n, d, d_partials);
    SAP_LAUNCH_CHECK();
    hipLaunchKernelGGL(reduce_final_kernel, dim3(1), dim3(RED_THREADS), 0, st, RED_BLOCKS, nd, d_partials, out);
    SAP_LAUNCH_CHECK();
}

void launch_dot(int64_t n, const double* a, const double* b, double* d_partials, double* out, hipStream_t st)
{
    launch_dots(n, 1, &a, &b, d_partials, out, st);
}

void launch_axpby(int64_t n, double alpha, const double* x, double beta, const double* y, double* z, hipStream_t st)
{
    hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, alpha, x, beta, y, z);
    SAP_LAUNCH_CHECK();
}

void launch_axpy_ratio(int64_t n, double num, const double* den, double sign, const double* x, double* y,
                       hipStream_t st)
{
    hipLaunchKernelGGL(axpby_ratio_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, num, den, sign, x, 1.0, y, y);
    SAP_LAUNCH_CHECK();
}

void launch_copy(int64_t n, const double* x, double* y, hipStream_t st)
{
    hipLaunchKernelGGL(copy_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, x, y);
    SAP_LAUNCH_CHECK();
}

void launch_spread(int64_t n, const double* src, double* const* dst, int nd, double* const* zero, int nz,
                   hipStream_t st)
{
    if (nd < 0 || nd > 4 || nz < 0 || nz > 4 || (nd > 0 && !src))
        throw Error(SAP_ERR_BADARG, "launch_spread: at most four copies and four zero fills");
    if (nd + nz == 0) return;
    SpreadArgs a{};
    a.src = src;
    for (int q = 0; q < nd; ++q) a.dst[q] = dst[q];
    for (int q = 0; q < nz; ++q) a.zero[q] = zero[q];
    a.nd = nd;
    a.nz = nz;
    hipLaunchKernelGGL(spread_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, a);
    SAP_LAUNCH_CHECK();
}

void launch_fill(int64_t n, double v, double* x, hipStream_t st)
{
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, v, x);
    SAP_LAUNCH_CHECK();
}

}  // namespace sap_amd
