// Variable-bandwidth preconditioner kernels (the reference's
// variableBandwidth path, sap/precond.h:1785-1800 setup, :1457-1470 apply).
//
// After the first-level reordering (DB, RCM, drop-off) the matrix is cut into
// P partitions; each partition's diagonal block is reordered again by its own
// RCM (second level, sap/graph.h:908-1008) and stored in a band of its own
// half-bandwidth K_p: partition p occupies Bv[boff_p, boff_p + n_p(2K_p + 1))
// with A(r, c) (partition-local, second-level order) at
// boff_p + c (2K_p + 1) + K_p + r - c (the reference's var layout,
// sap/device/data_transfer.cuh:475-494).  The coupling blocks B_i, C_{i+1}
// stay in first-level order (k x k, column-major, after the band).
#include "sap_internal.hpp"

namespace sap_amd {

namespace {

// Diagonal-block entries (CSR in second-level order, global indices) into
// the zero-filled variable-width band (var::copyFromCOOMatrixToBandedMatrix,
// sap/precond.h:1818-1823).  One thread per row.
template <class T>
__global__ void csr_to_varband_kernel(PartInfo pi, const int* __restrict__ rp, const int* __restrict__ ci,
                                      const T* __restrict__ val, const int* __restrict__ ks,
                                      const int64_t* __restrict__ boff, T* __restrict__ Bv)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= pi.n) return;
    const int p = pi.part_of(r);
    int64_t s, len;
    pi.bounds(p, s, len);
    const int kp = ks[p];
    const int64_t w = 2 * kp + 1;
    for (int e = rp[r]; e < rp[r + 1]; ++e) {
        const int64_t c = ci[e];
        Bv[boff[p] + (c - s) * w + kp + (r - c)] = val[e];
    }
}

// Spike tips of the variable-bandwidth path (calculateSpikes_var,
// sap/precond.h:5540-5790, with fwdElim/bckElim_*Spike_per_partition of
// sap/device/sweep_band_var.cuh), one column of a coupling block per job:
//   side 0: V_i = bottom k rows (first-level order) of A_i^{-1} [0; B_i]
//           over the packed columns of B_i (offDiagPerms_right: the columns
//           of B_i that hold an entry, widths_right of them);
//   side 1: W_{i+1} = top k rows of A_{i+1}^{-1} [C_{i+1}; 0] over the
//           packed columns of C_{i+1} (offDiagPerms_left, packed from the
//           right end).
// var_rhs_kernel scatters each packed column to the second-level rows of
// its coupling rows (a full-length right-hand side of the partition), the
// partition's full sweep (the staged job-table sweep, one job per column)
// solves it on the partition's own band, and var_tips_kernel gathers the
// tip rows back into first-level order and unpacks the column into the
// k x k block of WV (WV[(2i + side) k^2 + col k + t]; WV zero-filled, so
// columns without entries stay zero).
__device__ inline void var_col(const VarSpikeCol& c, int k, const int* __restrict__ perms_r,
                               const int* __restrict__ perms_l, const int* __restrict__ widths, int& col)
{
    col = c.side == 0 ? perms_r[(size_t)c.i * k + c.q]
                      : perms_l[(size_t)c.i * k + (k - widths[2 * c.i + 1]) + c.q];
}

template <class T>
__global__ void var_rhs_kernel(PartInfo pi, int k, const int* __restrict__ sp, const T* __restrict__ offB,
                               const T* __restrict__ offC, const int* __restrict__ perms_r,
                               const int* __restrict__ perms_l, const int* __restrict__ widths,
                               const VarSpikeCol* __restrict__ cols, int64_t stride, double* __restrict__ V)
{
    const VarSpikeCol c = cols[blockIdx.x];
    const int p = c.i + c.side;
    int64_t s, len;
    pi.bounds(p, s, len);
    double* v = V + (int64_t)blockIdx.x * stride;
    for (int64_t r = threadIdx.x; r < len; r += blockDim.x) v[r] = 0.0;
    __syncthreads();
    int col;
    var_col(c, k, perms_r, perms_l, widths, col);
    const T* src = (c.side == 0 ? offB : offC) + (size_t)c.i * k * k + (size_t)col * k;
    const int64_t trow0 = c.side == 0 ? s + len - k : s;   // first coupling row (first-level order)
    for (int t = threadIdx.x; t < k; t += blockDim.x) v[sp[trow0 + t] - s] = (double)src[t];
}

template <class T>
__global__ void var_tips_kernel(PartInfo pi, int k, const int* __restrict__ sp, const int* __restrict__ perms_r,
                                const int* __restrict__ perms_l, const int* __restrict__ widths,
                                const VarSpikeCol* __restrict__ cols, int64_t stride, const double* __restrict__ Z,
                                T* __restrict__ WV)
{
    const VarSpikeCol c = cols[blockIdx.x];
    const int p = c.i + c.side;
    int64_t s, len;
    pi.bounds(p, s, len);
    const double* z = Z + (int64_t)blockIdx.x * stride;
    int col;
    var_col(c, k, perms_r, perms_l, widths, col);
    T* dst = WV + (size_t)(2 * c.i + c.side) * k * k + (size_t)col * k;
    const int64_t trow0 = c.side == 0 ? s + len - k : s;
    for (int t = threadIdx.x; t < k; t += blockDim.x) dst[t] = (T)z[sp[trow0 + t] - s];
}

// Interface values for the reduced solve: g (first-level positions, type T)
// of rows s_{i+1} - k .. s_{i+1} + k - 1 of every boundary i, from the
// second-level result gd of the first sweep pass.  Column blockIdx.y.
template <class T>
__global__ void var_gather_kernel(PartInfo pi, int k, const int* __restrict__ sp, const double* __restrict__ gd,
                                  int64_t ldgd, T* __restrict__ g, const int* __restrict__ gate)
{
    if (gate && !gate[blockIdx.y]) return;
    gd += blockIdx.y * ldgd;
    g += blockIdx.y * pi.n;
    const int i = blockIdx.x;
    int64_t s1, l1;
    pi.bounds(i + 1, s1, l1);
    for (int t = threadIdx.x; t < 2 * k; t += blockDim.x) {
        const int64_t j = s1 - k + t;
        g[j] = (T)gd[sp[j]];
    }
}

// Right-hand side of the second sweep pass in second-level order:
// out[r2] = T(v[r2]) - correction of its first-level row j = sr[r2] (rows
// next to a boundary, purifyRHS, sap/precond.h:5494-5530), rounded in T as
// the constant-band sweep rounds it; other rows are copied.
template <class T>
__global__ void var_purify_kernel(PartInfo pi, int k, const int* __restrict__ sr, const double* __restrict__ v,
                                  int64_t ldv, const T* __restrict__ corr, double* __restrict__ out, int64_t ldo,
                                  const int* __restrict__ gate)
{
    if (gate && !gate[blockIdx.y]) return;
    v += blockIdx.y * ldv;
    out += blockIdx.y * ldo;
    corr += blockIdx.y * (int64_t)2 * k * (pi.P - 1);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r2 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r2 < pi.n; r2 += stride) {
        const int64_t j = sr[r2];
        const int p = pi.part_of(j);
        int64_t s, len;
        pi.bounds(p, s, len);
        const int64_t lr = j - s;
        const double x = v[r2];
        if (p > 0 && lr < k) {
            T val = (T)x;
            val -= corr[(size_t)(p - 1) * 2 * k + k + lr];
            out[r2] = (double)val;
        } else if (p < pi.P - 1 && lr >= len - k) {
            T val = (T)x;
            val -= corr[(size_t)p * 2 * k + (lr - (len - k))];
            out[r2] = (double)val;
        } else {
            out[r2] = (double)(T)x;
        }
    }
}

}  // namespace

template <class T>
void launch_csr_to_varband(PartInfo pi, const int* rp, const int* ci, const T* val, const int* ks,
                           const int64_t* boff, T* Bv, hipStream_t st)
{
    const int bs = 256;
    hipLaunchKernelGGL(csr_to_varband_kernel<T>, dim3((unsigned)((pi.n + bs - 1) / bs)), dim3(bs), 0, st, pi, rp, ci,
                       val, ks, boff, Bv);
    SAP_LAUNCH_CHECK();
}

template <class T>
void launch_var_rhs(PartInfo pi, int k, const int* sp, const T* offB, const T* offC, const int* perms_r,
                    const int* perms_l, const int* widths, const VarSpikeCol* cols, int ncol, int64_t stride,
                    double* V, hipStream_t st)
{
    if (ncol <= 0) return;
    hipLaunchKernelGGL(var_rhs_kernel<T>, dim3(ncol), dim3(256), 0, st, pi, k, sp, offB, offC, perms_r, perms_l,
                       widths, cols, stride, V);
    SAP_LAUNCH_CHECK();
}

template <class T>
void launch_var_tips(PartInfo pi, int k, const int* sp, const int* perms_r, const int* perms_l, const int* widths,
                     const VarSpikeCol* cols, int ncol, int64_t stride, const double* Z, T* WV, hipStream_t st)
{
    if (ncol <= 0) return;
    hipLaunchKernelGGL(var_tips_kernel<T>, dim3(ncol), dim3(256), 0, st, pi, k, sp, perms_r, perms_l, widths, cols,
                       stride, Z, WV);
    SAP_LAUNCH_CHECK();
}

template <class T>
void launch_var_gather(PartInfo pi, int k, const int* sp, const double* gd, int64_t ldgd, T* g, int nr,
                       const int* gate, hipStream_t st)
{
    if (pi.P < 2) return;
    hipLaunchKernelGGL(var_gather_kernel<T>, dim3(pi.P - 1, nr), dim3(256), 0, st, pi, k, sp, gd, ldgd, g, gate);
    SAP_LAUNCH_CHECK();
}

template <class T>
void launch_var_purify(PartInfo pi, int k, const int* sr, const double* v, int64_t ldv, const T* corr, double* out,
                       int64_t ldo, int nr, const int* gate, hipStream_t st)
{
    int64_t g = (pi.n + 255) / 256;
    if (g > 8192) g = 8192;
    hipLaunchKernelGGL(var_purify_kernel<T>, dim3((unsigned)g, nr), dim3(256), 0, st, pi, k, sr, v, ldv, corr, out,
                       ldo, gate);
    SAP_LAUNCH_CHECK();
}

#define SAP_INST(T)                                                                                               \
    template void launch_csr_to_varband<T>(PartInfo, const int*, const int*, const T*, const int*, const int64_t*, \
                                           T*, hipStream_t);                                                       \
    template void launch_var_rhs<T>(PartInfo, int, const int*, const T*, const T*, const int*, const int*,        \
                                    const int*, const VarSpikeCol*, int, int64_t, double*, hipStream_t);          \
    template void launch_var_tips<T>(PartInfo, int, const int*, const int*, const int*, const int*,               \
                                     const VarSpikeCol*, int, int64_t, const double*, T*, hipStream_t);           \
    template void launch_var_gather<T>(PartInfo, int, const int*, const double*, int64_t, T*, int, const int*,     \
                                       hipStream_t);                                                               \
    template void launch_var_purify<T>(PartInfo, int, const int*, const double*, int64_t, const T*, double*,       \
                                       int64_t, int, const int*, hipStream_t);
SAP_INST(float)
SAP_INST(double)
#undef SAP_INST

}  // namespace sap_amd
