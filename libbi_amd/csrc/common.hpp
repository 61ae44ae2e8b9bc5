// Shared device/host definitions of the LibBi particle-filter library for
// gfx950 (MI355X): Philox4x32-10, variate conversions, wavefront and block
// reductions, launch error handling.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#include "../../include/bihip.h"
#include "internal.hpp"

namespace bihip {

// ---------------------------------------------------------------- random streams
// Counter layout shared with the CPU oracle:
//   counter = { index, step, tag | (batch << 8), block }, key = { seed, seed >> 32 }.
enum : uint32_t {
  TAG_INIT = 1,
  TAG_NOISE = 2,
  TAG_SYSTEMATIC = 3,
  TAG_STRATIFIED = 4,
  TAG_MULTINOMIAL = 5,
  TAG_OBS = 6
};

struct u32x4 {
  uint32_t x, y, z, w;
};

// Philox4x32-10 (Salmon et al., SC'11).  Each 32x32->64 product is one
// 64-bit multiply (a single v_mad_u64_u32 rather than a low and a high
// multiply); the key schedule is uniform per launch.
__device__ __forceinline__ u32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                               uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
    const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
    const uint32_t n0 = hi1 ^ c1 ^ k0;
    const uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return u32x4{c0, c1, c2, c3};
}

__device__ __forceinline__ u32x4 philox_block(uint64_t seed, uint32_t index, uint32_t step,
                                              uint32_t tag, uint32_t batch, uint32_t block) {
  return philox4x32_10(index, step, tag | (batch << 8), block, (uint32_t)seed,
                       (uint32_t)(seed >> 32));
}

// uniform on [0,1) from the top 53 bits of (r0:r1)
__device__ __forceinline__ double u53(uint32_t r0, uint32_t r1) {
  const uint64_t m = ((((uint64_t)r0) << 32) | r1) >> 11;
  return (double)m * 0x1.0p-53;
}

// Box-Muller pair.  double: u1 = 1 - r0*2^-32 in (0,1], u2 = r1*2^-32;
// float: the top 24 bits of each word (exact in float).
template <class T>
struct Normal;

template <>
struct Normal<double> {
  __device__ __forceinline__ static void pair(uint32_t r0, uint32_t r1, double& z0,
                                              double& z1) {
    const double u1 = 1.0 - (double)r0 * 0x1.0p-32;
    const double u2 = (double)r1 * 0x1.0p-32;
    const double rad = sqrt(-2.0 * log(u1));
    double s, c;
    sincospi(2.0 * u2, &s, &c);
    z0 = rad * c;
    z1 = rad * s;
  }
};

template <>
struct Normal<float> {
  __device__ __forceinline__ static void pair(uint32_t r0, uint32_t r1, float& z0, float& z1) {
    const float u1 = 1.0f - (float)(r0 >> 8) * 0x1.0p-24f;
    const float u2 = (float)(r1 >> 8) * 0x1.0p-24f;
    // native v_log_f32 (log2), v_sqrt_f32, and v_sin/v_cos_f32, which take
    // their argument in revolutions: sin(2 pi u2) is v_sin_f32(u2)
    const float rad = __builtin_sqrtf(-2.0f * 0.69314718055994531f * __builtin_amdgcn_logf(u1));
    z0 = rad * __builtin_amdgcn_cosf(u2);
    z1 = rad * __builtin_amdgcn_sinf(u2);
  }
};

// The four normals of one Philox block.
template <class T>
__device__ __forceinline__ void normals4(uint64_t seed, uint32_t index, uint32_t step,
                                         uint32_t tag, uint32_t batch, uint32_t block,
                                         T (&z)[4]) {
  const u32x4 r = philox_block(seed, index, step, tag, batch, block);
  Normal<T>::pair(r.x, r.y, z[0], z[1]);
  Normal<T>::pair(r.z, r.w, z[2], z[3]);
}

// ---------------------------------------------------------------- NaN rules
// nanexp (NaN -> 0) as the reference's bi::nanexp
__device__ __forceinline__ double nanexp(double x) { return isnan(x) ? 0.0 : exp(x); }

// max with NaN below every number
__device__ __forceinline__ double nan_max(double a, double b) {
  if (isnan(a)) return b;
  if (isnan(b)) return a;
  return a < b ? b : a;
}

// ---------------------------------------------------------------- XCD-aware blocks
// The chip's 8 XCDs each have their own L2, and consecutive workgroups are
// dispatched to them in turn (block b runs on XCD b % 8).  xcd_block maps the
// dispatched block b of nb to a logical block so that each XCD takes one
// contiguous range of logical blocks: neighbouring particles (for instance
// the consecutive copies of one ancestor after a resample) then share one
// L2 instead of being fetched into eight.  A bijection on [0, nb).
constexpr int NUM_XCD = 8;
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int q = nb / NUM_XCD, r = nb % NUM_XCD;
  const int k = b % NUM_XCD, j = b / NUM_XCD;
  return k * q + (k < r ? k : r) + j;
}

// ---------------------------------------------------------------- wave helpers
constexpr int WAVE = 64;

template <class T>
__device__ __forceinline__ T shfl_xor(T v, int mask) {
  return __shfl_xor(v, mask, WAVE);
}

}  // namespace bihip
