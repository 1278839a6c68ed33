// bh_mfma.hpp -- the 16x16x4 matrix-core instructions of gfx950 used by the
// blocked band kernels, FP32 and FP64.
#pragma once

#include "bh_common.hpp"

namespace bh {

template <typename T>
struct Mf;
template <>
struct Mf<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};
template <>
struct Mf<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
// Result layout of the 16x16x4 MFMA (checked on gfx950): lane l holds
// column l%16 of D, element v of its result vector being row drow(l/16, v):
// 4(l/16)+v for the f32 instruction, l/16+4v for the f64 one.  The operands
// are A(i, k) from lane i+16k and B(k, j) from lane j+16k for both.
template <typename T>
__device__ __forceinline__ int drow(int g, int v) {
    return sizeof(T) == 4 ? 4 * g + v : g + 4 * v;
}

// Raw buffer resource (word 3 for gfx9-family parts) and the offset that
// lies beyond any resource: loads there return 0, stores there are dropped.
constexpr int kQRsrcWord3 = 0x00020000;
constexpr int kQOob = 0x7ffffff0;
template <typename T>
__device__ __forceinline__ T qbload(__amdgpu_buffer_rsrc_t r, int off);
template <>
__device__ __forceinline__ float qbload<float>(__amdgpu_buffer_rsrc_t r, int off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <>
__device__ __forceinline__ double qbload<double>(__amdgpu_buffer_rsrc_t r, int off) {
    typedef unsigned v2u __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
template <typename T>
__device__ __forceinline__ void qbstore(__amdgpu_buffer_rsrc_t r, int off, T v);
template <>
__device__ __forceinline__ void qbstore<float>(__amdgpu_buffer_rsrc_t r, int off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 0);
}
template <>
__device__ __forceinline__ void qbstore<double>(__amdgpu_buffer_rsrc_t r, int off, double v) {
    typedef unsigned v2u __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, v), r, off, 0, 0);
}

}  // namespace bh
