// Matrix-core (MFMA) helpers for gfx950 kernels: one 16 x 16 x 4 step
// D = A B + C in FP32 or FP64, with the operand and accumulator lane layouts
// of the instructions (checked against plain products in the tests):
//   A (16 x 4):  lane l holds A[l % 16][l / 16]
//   B (4 x 16):  lane l holds B[l / 16][l % 16]
//   D (16 x 16): lane l holds D[row(l, v)][l % 16], v = 0..3, with
//                row = 4 (l / 16) + v (FP32) or l / 16 + 4 v (FP64).
#pragma once

#include <hip/hip_runtime.h>

namespace ems {

typedef float ems_f4 __attribute__((ext_vector_type(4)));
typedef double ems_d4 __attribute__((ext_vector_type(4)));

template <typename T> struct MfmaAcc;
template <> struct MfmaAcc<float> {
    typedef ems_f4 V;
    static __device__ __forceinline__ V step(float a, float b, V c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};
template <> struct MfmaAcc<double> {
    typedef ems_d4 V;
    static __device__ __forceinline__ V step(double a, double b, V c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};

}  // namespace ems
