// 16x16x4 matrix-core step in the arithmetic type of the kernel (device code only).
//
// Lane (j, g) = (lane % 16, lane / 16) supplies A[row j][k g] and B[k g][column j]; the 16x16
// accumulator's rows are laid out as mfma_row<T>(g, i) for register i of lane group g.
#pragma once

#include <hip/hip_runtime.h>

namespace tdg {

template <class T> struct Mma;
template <> struct Mma<double> {
    using acc = double __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc mma(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <> struct Mma<float> {
    using acc = float __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc mma(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

} // namespace tdg
