// 16x16x4 matrix-core operations of gfx950 with their lane layouts (checked on
// the device with integer operands):
//   A (16 x 4):  lane l holds A[l % 16][l / 16]
//   B (4 x 16):  lane l holds B[l / 16][l % 16]
//   D (16 x 16): lane l holds D[row(l, e)][l % 16] in element e = 0..3, with
//                row = 4 (l / 16) + e for FP32 and row = l / 16 + 4 e for FP64.
#pragma once
#include <hip/hip_runtime.h>

typedef float mf4 __attribute__((ext_vector_type(4)));
typedef double md4 __attribute__((ext_vector_type(4)));

template <class T> struct MfmaT;
template <> struct MfmaT<float> {
    typedef mf4 V;
    static __device__ __forceinline__ V mma(float a, float b, V c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int e) { return 4 * (l >> 4) + e; }
};
template <> struct MfmaT<double> {
    typedef md4 V;
    static __device__ __forceinline__ V mma(double a, double b, V c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int e) { return (l >> 4) + 4 * e; }
};
