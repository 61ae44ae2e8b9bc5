// Model wrappers shared by the single-filter and the batched kernels:
// observation log-density terms, transition noise, and the model types the
// kernels are instantiated on.
#pragma once

#include <cmath>

#include "models.hpp"

namespace bihip {

inline ModelArgs to_args(const bihip_model* m) {
  ModelArgs a;
  a.model = m->model;
  a.d = m->d;
  a.integrator = m->integrator;
  a.nsteps = m->nsteps;
  a.host_quirk = m->host_quirk;
  a.precision = m->precision;
  a.h0 = m->h0;
  a.atol = m->atol;
  a.rtol = m->rtol;
  for (int i = 0; i < 16; ++i) a.par[i] = m->par[i];
  a.obs_pre = 0;
  a.obs_inv = a.obs_c = 0.0;
  if (m->model == BIHIP_MODEL_L96 && m->par[2] > 0.0) {  // obs_sd of the Lorenz-96 model
    a.obs_pre = 1;
    a.obs_inv = 1.0 / m->par[2];
    a.obs_c = -0.91893853320467274178 - std::log(m->par[2]);
  }
  return a;
}

// ---------------------------------------------------------------- observation terms
// Gaussian / log-normal log-density terms of the lane's slice
// (share/tt/cpp/action/gaussian.hpp.tt:48-70), in double.
template <class T>
__device__ __forceinline__ double obs_terms(const ModelPZ<T>& m, const T (&x)[2],
                                            const double* y, int) {
  const double yy = y[0];
  const double mu = log((double)x[0]);
  const double z = (log(yy) - mu) / m.obs_sd;
  return -0.5 * (z * z) - HALF_LOG_TWO_PI - log(m.obs_sd * yy);
}

template <class T, int D, int LPP>
__device__ __forceinline__ double obs_terms(const ModelL96<T, D, LPP>& m,
                                            const T (&x)[ModelL96<T, D, LPP>::DL],
                                            const double* y, int lo,
                                            const ModelArgs* pre = nullptr) {
  constexpr int DL = ModelL96<T, D, LPP>::DL;
  double inv, c;
  if (pre != nullptr && pre->obs_pre) {
    inv = pre->obs_inv;
    c = pre->obs_c;
  } else {
    inv = 1.0 / m.obs_sd;
    c = -HALF_LOG_TWO_PI - log(m.obs_sd);
  }
  double s = 0.0;
  if constexpr (DL % 2 == 0) {  // even slices start on an observed dimension
#pragma unroll
    for (int i = 0; i < DL; i += 2) {
      const double z = (y[(lo + i) >> 1] - (double)x[i]) * inv;
      s += -0.5 * (z * z) + c;
    }
  } else {  // odd slices: the observed (even) dimensions alternate with the lane
    const int o = lo & 1;
#pragma unroll
    for (int i = 0; i + 1 < DL; i += 2) {
      const int j = i + o;  // local index of an observed dimension
      const double xv = (double)(o ? x[i + 1] : x[i]);
      const double z = (y[(lo + j) >> 1] - xv) * inv;
      s += -0.5 * (z * z) + c;
    }
    if (o == 0) {  // the last element is observed only for even starts
      const double z = (y[(lo + DL - 1) >> 1] - (double)x[DL - 1]) * inv;
      s += -0.5 * (z * z) + c;
    }
  }
  return s;
}

template <class T>
__device__ __forceinline__ double obs_terms(const ModelLG<T>& m, const T (&x)[1],
                                            const double* y, int) {
  const double z = (y[0] - (double)x[0]) / m.obs_sd;
  return -0.5 * (z * z) - HALF_LOG_TWO_PI - log(m.obs_sd);
}

// ---------------------------------------------------------------- noise
template <class T>
__device__ __forceinline__ void noise_first(ModelPZ<T>& m, uint64_t seed, uint32_t batch, int p,
                                            uint32_t step) {
  T z[4];
  normals4<T>(seed, (uint32_t)p, step, TAG_NOISE, batch, 0, z);
  m.alpha = (T)m.mu + (T)m.sigma * z[0];
}

// zout (nullable): the standard normals w of the lane's slice, the model
// file's noise variable (x <- x + noise_sd w), at zout[(lo + i) * ld + p].
template <class M>
__device__ __forceinline__ void noise_after(const M& m, uint64_t seed, uint32_t batch, int p,
                                            int lo, uint32_t step,
                                            typename M::real_t (&x)[M::DL],
                                            typename M::real_t* zout = nullptr, int ld = 0) {
  using T = typename M::real_t;
  T z[M::DL];
  slice_normals<T, M::DL>(seed, (uint32_t)p, step, TAG_NOISE, batch, lo, z);
  const T sd = (T)m.noise_sd;
#pragma unroll
  for (int i = 0; i < M::DL; ++i) x[i] = x[i] + sd * z[i];
  if (zout != nullptr) {
#pragma unroll
    for (int i = 0; i < M::DL; ++i) zout[(size_t)(lo + i) * ld + p] = z[i];
  }
}

template <class T>
struct PZ : ModelPZ<T> {
  using real_t = T;
  using ModelPZ<T>::ModelPZ;
};
template <class T, int D, int LPP>
struct L96 : ModelL96<T, D, LPP> {
  using real_t = T;
  using ModelL96<T, D, LPP>::ModelL96;
};
template <class T>
struct LG : ModelLG<T> {
  using real_t = T;
  using ModelLG<T>::ModelLG;
};

template <class T>
__device__ __forceinline__ double obs_of(const PZ<T>& m, const T (&x)[2], const double* y,
                                         int lo, const ModelArgs* = nullptr) {
  return obs_terms(static_cast<const ModelPZ<T>&>(m), x, y, lo);
}
template <class T, int D, int LPP>
__device__ __forceinline__ double obs_of(const L96<T, D, LPP>& m,
                                         const T (&x)[ModelL96<T, D, LPP>::DL],
                                         const double* y, int lo,
                                         const ModelArgs* pre = nullptr) {
  return obs_terms(static_cast<const ModelL96<T, D, LPP>&>(m), x, y, lo, pre);
}
template <class T>
__device__ __forceinline__ double obs_of(const LG<T>& m, const T (&x)[1], const double* y,
                                         int lo, const ModelArgs* = nullptr) {
  return obs_terms(static_cast<const ModelLG<T>&>(m), x, y, lo);
}

}  // namespace bihip
