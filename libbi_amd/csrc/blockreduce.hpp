// Deterministic wavefront/block reductions (fixed association order).
#pragma once

#include "common.hpp"

namespace bihip {

// Butterfly reductions over the wave.  FROM > 1: only lanes that are
// multiples of FROM hold values (the others hold the reduction's identity),
// so the exchanges below FROM are skipped; lane 0 and every lane that holds
// a value end with the same result as the full butterfly.
template <int FROM = 1>
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = FROM; o < WAVE; o <<= 1) v = nan_max(v, __shfl_xor(v, o, WAVE));
  return v;
}

template <class T, int FROM = 1>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = FROM; o < WAVE; o <<= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}

// Block-wide max (NaN below all); every thread gets the result.
template <int BLOCK, int FROM = 1>
__device__ __forceinline__ double block_max(double v, double* lds) {
  constexpr int NW = BLOCK / WAVE;
  v = wave_max<FROM>(v);
  const int w = threadIdx.x / WAVE, l = threadIdx.x & (WAVE - 1);
  __syncthreads();
  if (l == 0) lds[w] = v;
  __syncthreads();
  double r = lds[0];
#pragma unroll
  for (int i = 1; i < NW; ++i) r = nan_max(r, lds[i]);
  return r;
}

// Block-wide sum; every thread gets the result.
template <int BLOCK, class T, int FROM = 1>
__device__ __forceinline__ T block_sum(T v, T* lds) {
  constexpr int NW = BLOCK / WAVE;
  v = wave_sum<T, FROM>(v);
  const int w = threadIdx.x / WAVE, l = threadIdx.x & (WAVE - 1);
  __syncthreads();
  if (l == 0) lds[w] = v;
  __syncthreads();
  T r = lds[0];
#pragma unroll
  for (int i = 1; i < NW; ++i) r += lds[i];
  return r;
}

// (max, sum exp(lw - max), sum exp(2(lw - max))) over the block's values;
// NaN values (also used for "no value") weigh zero, as bi::nanexp does.
// FROM: values only in lanes that are multiples of FROM (one per lane group).
template <int BLOCK, int FROM = 1>
__device__ __forceinline__ void block_weight_partial(double lw, double* out) {
  __shared__ double lds[BLOCK / WAVE];
  const double m = block_max<BLOCK, FROM>(lw, lds);
  const double z = nanexp(lw - m);
  const double s1 = block_sum<BLOCK, double, FROM>(z, lds);
  const double s2 = block_sum<BLOCK, double, FROM>(z * z, lds);
  if (threadIdx.x == 0) {
    out[0] = m;
    out[1] = s1;
    out[2] = s2;
  }
}

}  // namespace bihip
