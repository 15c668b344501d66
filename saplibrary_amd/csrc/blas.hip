// FP64 vector kernels for the Krylov loops (the cusp::blas calls of
// sap/bicgstab2.h, sap/minres.h and sap/monitor.h).  Reductions are
// deterministic: a fixed grid writes per-block partial sums, and a single
// block combines them in a fixed order.
#include "sap_internal.hpp"

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

__global__ void axpby_kernel(int64_t n, double alpha, const double* __restrict__ x, double beta,
                             const double* __restrict__ y, double* __restrict__ z)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        z[i] = alpha * x[i] + beta * y[i];
}

// z = alpha x + beta y with alpha = sign * (num / *den) read on the device
// (the same division the host would make, so the same alpha)
__global__ void axpby_ratio_kernel(int64_t n, double num, const double* __restrict__ den, double sign,
                                   const double* __restrict__ x, double beta, const double* __restrict__ y,
                                   double* __restrict__ z)
{
    const double alpha = sign * (num / *den);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        z[i] = alpha * x[i] + beta * y[i];
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

struct SpreadArgs {
    const double* src;
    double* dst[4];
    double* zero[4];
    int nd, nz;
};

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

inline unsigned grid_for(int64_t n)
{
    int64_t g = (n + 255) / 256;
    if (g > 8192) g = 8192;
    if (g < 1) g = 1;
    return (unsigned)g;
}

}  // namespace

int dot_partials_size() { return 3 * RED_BLOCKS; }

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
