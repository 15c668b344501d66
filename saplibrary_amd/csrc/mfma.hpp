// 16 x 16 x 4 matrix-core step for FP32 and FP64 (v_mfma_f32_16x16x4f32,
// v_mfma_f64_16x16x4f64): D += A B with A 16 x 4 and B 4 x 16.  Lane l
// supplies a = A[l % 16][l / 16] and b = B[l / 16][l % 16]; accumulator
// register q of lane l holds D[row(l, q)][l % 16].
#pragma once

#include <hip/hip_runtime.h>

namespace sap_amd {

template <class T>
struct Mfma16;

template <>
struct Mfma16<float> {
    typedef float acc_t __attribute__((ext_vector_type(4)));
    __device__ static inline acc_t mma(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
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

}  // namespace sap_amd
