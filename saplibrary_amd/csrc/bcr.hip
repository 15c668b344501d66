// Block cyclic reduction (BCR) of the partitions (the reference's useBCR
// variant: bcrBlockedBandedLU, SegmentedMatrix sweeps and products, and
// updateRHS, sap/precond.h:981-1210, :6485-7076, sap/segmented_matrix.h).
//
// Each partition (band half-width k) is cut into blocks of at least 2k rows,
// making it block tridiagonal: D_j on the diagonal, and the couplings
// CL_j (the first k rows of block j against the last k columns of its left
// neighbour) and CU_j (the last k rows of j against the first k columns of
// its right neighbour), k x k each.  Level l (stride s = 2^l) eliminates the
// active blocks e = s, 3s, 5s, ... of every partition at once: D_e is
// factored (LU in place in the band), the tips of X = D_e^{-1} CL_e (first
// and last k rows) and Y = D_e^{-1} CU_e are formed, and with a = e - s and
// c = e + s
//     D_a[last k, last k]   -= CU_a X_top        CU_a <- -CU_a Y_top
//     D_c[first k, first k] -= CL_c Y_bot        CL_c <- -CL_c X_bot
// (the new couplings link a and c directly).  Blocks of >= 2k rows keep the
// two corners of a block disjoint and every update inside the band.  The
// block left at the end (block 0) is factored last.  The apply runs the
// levels forward (t = D_e^{-1} w_e; w_a[last k] -= CU_a t_top,
// w_c[first k] -= CL_c t_bot with the couplings as they were at that
// level), solves block 0, and back-substitutes level by level
// (x_e = D_e^{-1} (w_e - CL_e x_a[last k] - CU_e x_c[first k])).
// All k x k blocks column-major.
#include "sap_internal.hpp"

namespace sap_amd {

namespace {

// Coupling blocks of every block from the band (before any factorization).
// blockIdx.x = block, blockIdx.y = 0 (CL) / 1 (CU).
template <class T>
__global__ void bcr_extract_kernel(const T* __restrict__ B, int k, const BcrBlock* __restrict__ blk,
                                   T* __restrict__ CL, T* __restrict__ CU)
{
    const int j = blockIdx.x, side = blockIdx.y;
    const BcrBlock b = blk[j];
    const int64_t w = 2 * k + 1;
    const size_t kk = (size_t)k * k;
    T* dst = (side == 0 ? CL : CU) + (size_t)j * kk;
    const bool has = side == 0 ? b.has_left : b.has_right;
    for (int e = threadIdx.x; e < k * k; e += blockDim.x) {
        const int a = e % k, c = e / k;   // row a, column c of the k x k block
        T v = (T)0;
        if (has) {
            int64_t r, col;
            if (side == 0) { r = b.start + a; col = b.start - k + c; }              // needs c >= a
            else           { r = b.start + b.len - k + a; col = b.start + b.len + c; }   // needs c <= a
            if ((side == 0 && c >= a) || (side == 1 && c <= a)) v = B[col * w + k + r - col];
        }
        dst[(size_t)c * k + a] = v;
    }
}

// Tips of X = D_e^{-1} [CL_e; 0] (side 0) and Y = D_e^{-1} [0; CU_e]
// (side 1, when e has a right neighbour) from the LU factors of D_e (L unit,
// D, U' unit in the band): full sweeps of the block with BT right-hand
// sides per workgroup, working vectors in global scratch (len x BT,
// row-major).  tips[4q + 0..3] = X_top, X_bot, Y_top, Y_bot of elimination q.
constexpr int BT = 32;

template <class T>
__global__ __launch_bounds__(256) void bcr_tips_kernel(const T* __restrict__ B, int k,
                                                       const BcrBlock* __restrict__ blk,
                                                       const BcrElim* __restrict__ el, const T* __restrict__ CL,
                                                       const T* __restrict__ CU, T* __restrict__ tips,
                                                       T* __restrict__ scratch, int64_t scratch_stride)
{
    const int q = blockIdx.x, side = blockIdx.y;
    const BcrElim E = el[q];
    if (side == 1 && E.c < 0) return;
    const int r0 = blockIdx.z * BT;
    const int nrhs = min(BT, k - r0);
    const BcrBlock b = blk[E.e];
    const int len = b.len;
    const int64_t w = 2 * k + 1;
    const size_t kk = (size_t)k * k;
    const T* F = B + b.start * w;   // column c of the block at F + c*w, A(r, c) at F[c*w + k + r - c]
    T* __restrict__ Y = scratch + (size_t)((2 * q + side) * gridDim.z + blockIdx.z) * scratch_stride;
    const int lo = side == 0 ? 0 : len - k;   // first row with a nonzero right-hand side
    for (int e = threadIdx.x; e < len * BT; e += blockDim.x) Y[e] = (T)0;
    __syncthreads();
    const T* src = (side == 0 ? CL : CU) + (size_t)E.e * kk;
    for (int e = threadIdx.x; e < k * nrhs; e += blockDim.x) {
        const int a = e % k, r = e / k;
        Y[(int64_t)(lo + a) * BT + r] = src[(size_t)(r0 + r) * k + a];
    }
    __syncthreads();
    for (int j = lo; j < len - 1; ++j) {
        const int nt = min(k, len - 1 - j);
        const T* col = F + (int64_t)j * w + k;
        for (int e = threadIdx.x; e < nt * nrhs; e += blockDim.x) {
            const int dt = 1 + e % nt, r = e / nt;
            Y[(int64_t)(j + dt) * BT + r] -= col[dt] * Y[(int64_t)j * BT + r];
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < len * nrhs; e += blockDim.x) {
        const int t = e / nrhs, r = e % nrhs;
        Y[(int64_t)t * BT + r] /= F[(int64_t)t * w + k];
    }
    __syncthreads();
    for (int j = len - 1; j > 0; --j) {
        const int nt = min(k, j);
        const T* col = F + (int64_t)j * w + k;
        for (int e = threadIdx.x; e < nt * nrhs; e += blockDim.x) {
            const int dt = 1 + e % nt, r = e / nt;
            Y[(int64_t)(j - dt) * BT + r] -= col[-dt] * Y[(int64_t)j * BT + r];
        }
        __syncthreads();
    }
    T* top = tips + (size_t)(4 * q + 2 * side) * kk;
    T* bot = top + kk;
    for (int e = threadIdx.x; e < k * nrhs; e += blockDim.x) {
        const int a = e % k, r = e / k;
        top[(size_t)(r0 + r) * k + a] = Y[(int64_t)a * BT + r];
        bot[(size_t)(r0 + r) * k + a] = Y[(int64_t)(len - k + a) * BT + r];
    }
}

// A Bm for k x k column-major blocks, 32 x 32 output tiles through LDS (all
// 256 threads of the workgroup; every loop trip count is uniform, so the
// barriers are too); store(r, c, v) receives each entry of the product.
template <class T, class Store>
__device__ void gemm_kk(const T* __restrict__ A, const T* __restrict__ Bm, int k, T* lA, T* lB, Store store)
{
    constexpr int TS = 32;
    const int nt = (k + TS - 1) / TS;
    const int tx = threadIdx.x % TS, ty = threadIdx.x / TS;   // 32 x 8 threads, 4 outputs each
    for (int ti = 0; ti < nt; ++ti)
        for (int tj = 0; tj < nt; ++tj) {
            T acc[4] = {(T)0, (T)0, (T)0, (T)0};
            for (int tk = 0; tk < nt; ++tk) {
                for (int e = threadIdx.x; e < TS * TS; e += blockDim.x) {
                    const int r = e % TS, c = e / TS;
                    const int ar = ti * TS + r, ac = tk * TS + c;
                    lA[c * (TS + 1) + r] = (ar < k && ac < k) ? A[(size_t)ac * k + ar] : (T)0;
                    const int br = tk * TS + r, bc = tj * TS + c;
                    lB[c * (TS + 1) + r] = (br < k && bc < k) ? Bm[(size_t)bc * k + br] : (T)0;
                }
                __syncthreads();
#pragma unroll 4
                for (int m = 0; m < TS; ++m) {
                    const T av = lA[m * (TS + 1) + tx];
#pragma unroll
                    for (int u = 0; u < 4; ++u) acc[u] += av * lB[(ty + 8 * u) * (TS + 1) + m];
                }
                __syncthreads();
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int r = ti * TS + tx, c = tj * TS + ty + 8 * u;
                if (r < k && c < k) store(r, c, acc[u]);
            }
        }
}

// Schur updates of one level, one workgroup per elimination (see the top of
// the file).  The neighbours' couplings as they are before this level are
// kept in UA[e] / LC[e] for the apply.
template <class T>
__global__ __launch_bounds__(256) void bcr_schur_kernel(T* __restrict__ B, int k, const BcrBlock* __restrict__ blk,
                                                        const BcrElim* __restrict__ el, T* __restrict__ CL,
                                                        T* __restrict__ CU, const T* __restrict__ tips,
                                                        T* __restrict__ UA, T* __restrict__ LC)
{
    __shared__ T lA[32 * 33], lB[32 * 33];
    const int q = blockIdx.x;
    const BcrElim E = el[q];
    const size_t kk = (size_t)k * k;
    const int64_t w = 2 * k + 1;
    T* ua = UA + (size_t)E.e * kk;
    T* lc = LC + (size_t)E.e * kk;
    for (int e = threadIdx.x; e < k * k; e += blockDim.x) {
        ua[e] = CU[(size_t)E.a * kk + e];
        lc[e] = E.c >= 0 ? CL[(size_t)E.c * kk + e] : (T)0;
    }
    __syncthreads();
    const T* X_top = tips + (size_t)(4 * q) * kk;
    const T* X_bot = X_top + kk;
    const T* Y_top = X_bot + kk;
    const T* Y_bot = Y_top + kk;
    const BcrBlock ba = blk[E.a];
    // D_a[last k, last k] -= CU_a X_top
    {
        const int64_t o = ba.start + ba.len - k;
        gemm_kk<T>(ua, X_top, k, lA, lB, [&](int r, int c, T v) {
            const int64_t gr = o + r, gc = o + c;
            B[gc * w + k + gr - gc] -= v;
        });
    }
    T* cua = CU + (size_t)E.a * kk;
    if (E.c >= 0) {
        gemm_kk<T>(ua, Y_top, k, lA, lB, [&](int r, int c, T v) { cua[(size_t)c * k + r] = -v; });
        const BcrBlock bc = blk[E.c];
        const int64_t o = bc.start;
        gemm_kk<T>(lc, Y_bot, k, lA, lB, [&](int r, int c, T v) {
            const int64_t gr = o + r, gc = o + c;
            B[gc * w + k + gr - gc] -= v;
        });
        T* clc = CL + (size_t)E.c * kk;
        gemm_kk<T>(lc, X_bot, k, lA, lB, [&](int r, int c, T v) { clc[(size_t)c * k + r] = -v; });
    } else {
        for (int e = threadIdx.x; e < k * k; e += blockDim.x) cua[e] = (T)0;
    }
}

// The apply's coupling products of one level for column blockIdx.y, one
// workgroup per elimination, rounded in T:
//   forward (mode 0):  w_a[last k] -= UA_e t_e[first k];  w_c[first k] -= LC_e t_e[last k]
//   backward (mode 1): w_e[first k] -= CL_e x_a[last k];  w_e[last k] -= CU_e x_c[first k]
template <class T>
__global__ void bcr_gemv_kernel(int mode, int k, const BcrBlock* __restrict__ blk, const BcrElim* __restrict__ el,
                                const T* __restrict__ M0, const T* __restrict__ M1, const double* __restrict__ src,
                                int64_t lds, double* __restrict__ wv, int64_t ldw, const int* __restrict__ gate)
{
    if (gate && !gate[blockIdx.y]) return;
    src += blockIdx.y * lds;
    wv += blockIdx.y * ldw;
    const BcrElim E = el[blockIdx.x];
    const size_t kk = (size_t)k * k;
    const BcrBlock be = blk[E.e], ba = blk[E.a];
    for (int side = 0; side < 2; ++side) {
        if (side == 1 && E.c < 0) break;
        const T* M = (side == 0 ? M0 : M1) + (size_t)E.e * kk;
        int64_t in0, out0;
        if (mode == 0) {
            in0 = side == 0 ? be.start : be.start + be.len - k;
            out0 = side == 0 ? ba.start + ba.len - k : blk[E.c].start;
        } else {
            in0 = side == 0 ? ba.start + ba.len - k : blk[E.c].start;
            out0 = side == 0 ? be.start : be.start + be.len - k;
        }
        for (int a = threadIdx.x; a < k; a += blockDim.x) {
            T acc = (T)0;
            for (int c = 0; c < k; ++c) acc += M[(size_t)c * k + a] * (T)src[in0 + c];
            T v = (T)wv[out0 + a];
            v -= acc;
            wv[out0 + a] = (double)v;
        }
    }
}

}  // namespace

template <class T>
void launch_bcr_extract(const T* B, int k, const BcrBlock* blk, int nblk, T* CL, T* CU, hipStream_t st)
{
    if (nblk <= 0) return;
    hipLaunchKernelGGL(bcr_extract_kernel<T>, dim3(nblk, 2), dim3(256), 0, st, B, k, blk, CL, CU);
    SAP_LAUNCH_CHECK();
}

template <class T>
void launch_bcr_level_setup(T* B, int k, const BcrBlock* blk, const BcrElim* el, int nel, int64_t max_len, T* CL,
                            T* CU, T* UA, T* LC, hipStream_t st)
{
    if (nel <= 0) return;
    const size_t kk = (size_t)k * k;
    const unsigned chunks = (unsigned)((k + BT - 1) / BT);
    const int64_t stride = max_len * BT;
    DevBuf<T> tips(4 * kk * nel);
    DevBuf<T> scratch((size_t)stride * 2 * nel * chunks);
    hipLaunchKernelGGL(bcr_tips_kernel<T>, dim3(nel, 2, chunks), dim3(256), 0, st, B, k, blk, el, CL, CU, tips.get(),
                       scratch.get(), stride);
    SAP_LAUNCH_CHECK();
    hipLaunchKernelGGL(bcr_schur_kernel<T>, dim3(nel), dim3(256), 0, st, B, k, blk, el, CL, CU, tips.get(), UA, LC);
    SAP_LAUNCH_CHECK();
    SAP_HIP_CHECK(hipStreamSynchronize(st));   // the work buffers are released on return
}

template <class T>
void launch_bcr_gemv(int mode, int k, const BcrBlock* blk, const BcrElim* el, int nel, const T* M0, const T* M1,
                     const double* src, int64_t lds, double* wv, int64_t ldw, int nr, const int* gate,
                     hipStream_t st)
{
    if (nel <= 0) return;
    hipLaunchKernelGGL(bcr_gemv_kernel<T>, dim3(nel, nr), dim3(256), 0, st, mode, k, blk, el, M0, M1, src, lds, wv,
                       ldw, gate);
    SAP_LAUNCH_CHECK();
}

#define SAP_INST(T)                                                                                               \
    template void launch_bcr_extract<T>(const T*, int, const BcrBlock*, int, T*, T*, hipStream_t);                 \
    template void launch_bcr_level_setup<T>(T*, int, const BcrBlock*, const BcrElim*, int, int64_t, T*, T*, T*,   \
                                            T*, hipStream_t);                                                      \
    template void launch_bcr_gemv<T>(int, int, const BcrBlock*, const BcrElim*, int, const T*, const T*,          \
                                     const double*, int64_t, double*, int64_t, int, const int*, hipStream_t);
SAP_INST(float)
SAP_INST(double)
#undef SAP_INST

}  // namespace sap_amd
