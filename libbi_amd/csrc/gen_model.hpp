// Models generated from LibBi model files (libbi_amd/bi/*.bi) by
// libbi_amd/dsl: what the generated functors use, and the generated header
// itself (written into the build folder when the library is built).
//
// A generated functor has the integrators' interface (rhs, D, DL, LPP = 1)
// plus the whole transition of one interval, the observation log-density,
// observation samples and initial samples; the filter kernels call those in
// place of their hand-written model code when is_generated<M> holds.
#pragma once

#include <type_traits>

#include "internal.hpp"
#include "model_ops.hpp"
#include "gen_samplers.hpp"

namespace bihip {

template <class M, class = void>
struct is_generated : std::false_type {};
template <class M>
struct is_generated<M, std::void_t<decltype(M::GENERATED)>> : std::bool_constant<M::GENERATED> {};

// Time in units of the transition's delta, rounded to a multiple of delta
// within 1e-4 (User2Scaled, share/src/bi/state/Schedule.hpp:30-42).
__device__ __forceinline__ double gen_scaled(double t, double delta) {
  const double st = t / delta;
  const double r = round(st);
  return fabs(st - r) < 1.0e-4 ? r : st;
}

// Uniform on [0,1) from one Philox block: 53 bits in double, 24 in float.
template <class T>
__device__ __forceinline__ T uniform_word(const u32x4& r);
template <>
__device__ __forceinline__ double uniform_word<double>(const u32x4& r) {
  return u53(r.x, r.y);
}
template <>
__device__ __forceinline__ float uniform_word<float>(const u32x4& r) {
  return (float)(r.x >> 8) * 0x1.0p-24f;
}

template <class M, std::enable_if_t<is_generated<M>::value, int> = 0>
__device__ __forceinline__ double obs_of(const M& m, const typename M::real_t (&x)[M::DL],
                                         const double* y, int lo, const ModelArgs* = nullptr) {
  return m.obs(x, y, lo);  // the terms of the lane's state elements [lo, lo + DL)
}

}  // namespace bihip

#include "generated_models.hpp"
