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

// float -> bfloat16 in the high half of a word, round to nearest even; a NaN
// stays a NaN (quiet, sign and payload kept).  Branch-free: the NaN case is a
// select, so the fused kernels carry no exec-mask branches per activation.
__device__ __forceinline__ uint32_t f2bf_hi(float f) {
  const uint32_t u = __float_as_uint(f);
  const uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
  return __builtin_isnan(f) ? (u | 0x400000u) : r;
}

// float -> bfloat16, round to nearest even (NaN stays NaN)
__device__ __forceinline__ uint16_t f2bf(float f) { return (uint16_t)(f2bf_hi(f) >> 16); }

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// two floats -> packed bfloat16 pair (low half = a), round to nearest even:
// the two high halves joined by one byte permute.  (The integer form keeps
// the fused kernels' register allocation free of spills, which the
// compiler's v_cvt_pk_bf16_f32 lowering did not.)
__device__ __forceinline__ uint32_t pack_bf2(float a, float b) {
  return __builtin_amdgcn_perm(f2bf_hi(b), f2bf_hi(a), 0x07060302u);
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

// ---------------------------------------------------------------- fused inference
// Layers computed transposed, out^T = W in^T, for 32 particles per wave: the
// weights are the A operand (rows staged in LDS with row stride LDSK), the
// activations the B operand held in registers as packed bfloat16 pairs
// (act[c][0..3] = inputs 16 c + 8 h + 0..7 of the lane's particle, h = lane / 32).
// An output tile's accumulator lands in the lane that owns the particle, with
// the rows of each 16-block in the order [0-3, 8-11, 4-7, 12-15] across the
// two half-waves; layers fed by such outputs store their input columns in that
// order (the inference weights).

// rows [row0, row0 + NROWS) of a row-major [.][K] bf16 matrix into Ws
template <int NROWS, int K, int LDSK, int THREADS>
__device__ __forceinline__ void stage_rows(const uint16_t* __restrict__ W, int row0,
                                           uint16_t* Ws) {
  constexpr int PER_ROW = K / 8, V = NROWS * PER_ROW;
  for (int v = threadIdx.x; v < V; v += THREADS) {
    const int row = v / PER_ROW, c8 = (v - row * PER_ROW) * 8;
    *reinterpret_cast<u16x8*>(&Ws[row * LDSK + c8]) =
        *reinterpret_cast<const u16x8*>(&W[(size_t)(row0 + row) * K + c8]);
  }
}

// rows [row0, row0 + nrows) (a run-time count) of a row-major [.][K] bf16 matrix
template <int K, int LDSK, int THREADS>
__device__ __forceinline__ void stage_rows_n(const uint16_t* __restrict__ W, int row0, int nrows,
                                             uint16_t* Ws) {
  constexpr int PER_ROW = K / 8;
  const int V = nrows * PER_ROW;
  for (int v = threadIdx.x; v < V; v += THREADS) {
    const int row = v / PER_ROW, c8 = (v - row * PER_ROW) * 8;
    *reinterpret_cast<u16x8*>(&Ws[row * LDSK + c8]) =
        *reinterpret_cast<const u16x8*>(&W[(size_t)(row0 + row) * K + c8]);
  }
}

// output tiles [T0, T1) of a GELU layer with KC input chunks: Ws holds rows
// from 32 T0 on; out[2 t + s] receives the packed activations of tile t
template <int KC, int NOUTC, int LDSK, int T0, int T1>
__device__ __forceinline__ void gelu_tiles(const uint16_t* Ws, const float* __restrict__ bias,
                                           const uint32_t (&in)[KC][4],
                                           uint32_t (&out)[NOUTC][4], int lane) {
  const int h = lane >> 5, row = lane & 31;
#pragma unroll
  for (int t = T0; t < T1; ++t) {
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
#pragma unroll
    for (int c = 0; c < KC; ++c) {
      const u16x8 wa =
          *reinterpret_cast<const u16x8*>(&Ws[(32 * (t - T0) + row) * LDSK + 16 * c + 8 * h]);
      const u32v4 bv = {in[c][0], in[c][1], in[c][2], in[c][3]};
      acc = mfma32(wa, bv, acc);
    }
#pragma unroll
    for (int r = 0; r < 16; r += 2) {
      const int n = 32 * t + acc_row(r, lane);
      out[2 * t + (r >> 3)][(r & 7) >> 1] =
          pack_bf2(gelu_f(acc[r] + bias[n]), gelu_f(acc[r + 1] + bias[n + 1]));
    }
  }
}

// one linear output tile t of a layer with KC input chunks (Ws holds rows from 32 t0)
template <int KC, int LDSK>
__device__ __forceinline__ f32x16 linear_tile(const uint16_t* Ws, int t, int t0,
                                              const float* __restrict__ bias,
                                              const uint32_t (&in)[KC][4], int lane) {
  const int h = lane >> 5, row = lane & 31;
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = bias[32 * t + acc_row(r, lane)];
#pragma unroll
  for (int c = 0; c < KC; ++c) {
    const u16x8 wa =
        *reinterpret_cast<const u16x8*>(&Ws[(32 * (t - t0) + row) * LDSK + 16 * c + 8 * h]);
    const u32v4 bv = {in[c][0], in[c][1], in[c][2], in[c][3]};
    acc = mfma32(wa, bv, acc);
  }
  return acc;
}

}  // namespace bihip
