// Device helpers of the BF16 MFMA multilayer perceptron (mlp.hip, surrogate.hip).
//
// bfloat16 values live in memory as uint16_t and reach the matrix cores as
// 8-element vectors.  The operand layouts of v_mfma_f32_32x32x16_bf16 on
// gfx950 (64 lanes, 16 FP32 accumulators per lane):
//   A (32 x 16): lane l holds A[l % 32][8 (l / 32) + j], j = 0..7
//   B (16 x 32): lane l holds B[8 (l / 32) + j][l % 32], j = 0..7
//   C (32 x 32): accumulator r of lane l is C[8 (r / 4) + 4 (l / 32) + r % 4][l % 32]
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bihip {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32v4 __attribute__((ext_vector_type(4)));

// float -> bfloat16, round to nearest even (NaN stays NaN)
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// two floats -> packed bfloat16 pair (low half = a), round to nearest even.
// (The integer form keeps the fused kernels' register allocation free of
// spills, which the compiler's v_cvt_pk_bf16_f32 lowering did not.)
__device__ __forceinline__ uint32_t pack_bf2(float a, float b) {
  return (uint32_t)f2bf(a) | ((uint32_t)f2bf(b) << 16);
}

// GELU in its sigmoid form z sigma(1.702 z) (Hendrycks & Gimpel 2016, sec. 2):
// one exponential and one reciprocal per activation, which keeps the hidden
// layers' vector work below the matrix-core time.  Derivative
// s + 1.702 z s (1 - s) with s = sigma(1.702 z).
__device__ __forceinline__ float gelu_sigmoid(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __expf(-1.702f * z));
}
__device__ __forceinline__ float gelu_f(float z) { return z * gelu_sigmoid(z); }
__device__ __forceinline__ float gelu_grad_f(float z) {
  const float s = gelu_sigmoid(z);
  return s + 1.702f * z * s * (1.0f - s);
}

__device__ __forceinline__ f32x16 mfma32(const u16x8& a, const u16x8& b, const f32x16& c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                 __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}
__device__ __forceinline__ f32x16 mfma32(const u16x8& a, const u32v4& b, const f32x16& c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                 __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// row of accumulator r of lane l in a 32 x 32 tile
__device__ __forceinline__ int acc_row(int r, int lane) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

}  // namespace bihip
