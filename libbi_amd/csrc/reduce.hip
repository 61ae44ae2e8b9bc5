// Log-weight reductions: max, log-sum-exp and ESS in one read of lws,
// as per-tile (max, sum w, sum w^2) partials combined in a fixed order.
//
// Replaces ess_reduce / logsumexp_reduce / max_reduce
// (share/src/bi/primitive/vector_primitive.hpp:860-974) and the
// bookkeeping of BootstrapPF::correct (share/src/bi/filter/BootstrapPF.hpp:171-181)
// with Resampler::resample's trigger (share/src/bi/resampler/Resampler.hpp:168).
#include "blockreduce.hpp"
#include "internal.hpp"

namespace bihip {

// One tile of RED_TILE log-weights per block: two-phase (block max, then sums).
__global__ __launch_bounds__(RED_BLOCK) void k_lws_partials(const double* lws, int P,
                                                            double* partials) {
  __shared__ double lds[RED_BLOCK / WAVE];
  const long base = (long)blockIdx.x * RED_TILE;
  double v[RED_ITEMS];
  double m = __builtin_nan("");
#pragma unroll
  for (int i = 0; i < RED_ITEMS; ++i) {
    const long idx = base + (long)i * RED_BLOCK + threadIdx.x;
    v[i] = idx < P ? lws[idx] : __builtin_nan("");
    m = nan_max(m, v[i]);
  }
  const double M = block_max<RED_BLOCK>(m, lds);
  double s1 = 0.0, s2 = 0.0;
#pragma unroll
  for (int i = 0; i < RED_ITEMS; ++i) {
    const double z = nanexp(v[i] - M);
    s1 += z;
    s2 += z * z;
  }
  s1 = block_sum<RED_BLOCK>(s1, lds);
  s2 = block_sum<RED_BLOCK>(s2, lds);
  if (threadIdx.x == 0) {
    partials[3 * blockIdx.x] = M;
    partials[3 * blockIdx.x + 1] = s1;
    partials[3 * blockIdx.x + 2] = s2;
  }
}

// Combine nb partials (one block): M = max m_b, S1 = sum s1_b e^{m_b - M},
// S2 = sum s2_b e^{2(m_b - M)}.  Every thread returns the totals.
__device__ void combine_partials(const double* partials, int nb, double& M, double& S1,
                                 double& S2) {
  __shared__ double lds[FINAL_BLOCK / WAVE];
  double m = __builtin_nan("");
  for (int b = threadIdx.x; b < nb; b += FINAL_BLOCK) m = nan_max(m, partials[3 * b]);
  M = block_max<FINAL_BLOCK>(m, lds);
  double s1 = 0.0, s2 = 0.0;
  for (int b = threadIdx.x; b < nb; b += FINAL_BLOCK) {
    const double a1 = partials[3 * b + 1], a2 = partials[3 * b + 2];
    if (a1 != 0.0) {
      const double f = exp(partials[3 * b] - M);
      s1 += a1 * f;
      s2 += a2 * (f * f);
    }
  }
  S1 = block_sum<FINAL_BLOCK>(s1, lds);
  S2 = block_sum<FINAL_BLOCK>(s2, lds);
}

__global__ __launch_bounds__(FINAL_BLOCK) void k_finalize(const double* partials, int nb, int P,
                                                          double* out) {
  double M, S1, S2;
  combine_partials(partials, nb, M, S1, S2);
  if (threadIdx.x == 0) {
    out[0] = M;
    out[1] = S1;
    out[2] = S2;
    out[3] = S1 * S1 / S2;                    // ESS
    out[4] = M + log(S1) - log((double)P);    // lW
  }
}

// BootstrapPF::correct bookkeeping on the device:
//   ess, lW = ess_reduce(lws); logIncrements[k] = lW - logLikelihood;
//   logLikelihood = lW; resample at the next time iff ess < essRel * P.
__global__ __launch_bounds__(FINAL_BLOCK) void k_pf_correct_finalize(const double* partials,
                                                                     int nb, int P,
                                                                     double essRel, PFDevice* dev,
                                                                     PFHistory hist) {
  double M, S1, S2;
  combine_partials(partials, nb, M, S1, S2);
  if (threadIdx.x == 0) {
    const int k = dev->k;
    const double ess = S1 * S1 / S2;
    const double lW = M + log(S1) - log((double)P);
    if (hist.resampled) hist.resampled[k] = dev->resample != 0 && dev->status == BIHIP_OK;
    if (hist.logIncrements) hist.logIncrements[k] = lW - dev->logLikelihood;
    if (hist.ess) hist.ess[k] = ess;
    dev->logLikelihood = lW;
    dev->ess = ess;
    dev->mx = M;
    dev->lW = lW;
    dev->resample = ess < essRel * (double)P ? 1 : 0;
    dev->k = k + 1;
  }
}

int launch_lws_partials(const double* lws, int P, double* partials, hipStream_t s) {
  hipLaunchKernelGGL(k_lws_partials, dim3(ceil_div(P, RED_TILE)), dim3(RED_BLOCK), 0, s, lws, P,
                     partials);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

int launch_finalize(const double* partials, int nb, int P, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(FINAL_BLOCK), 0, s, partials, nb, P, out);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

int launch_pf_correct_finalize(const double* partials, int nb, int P, double essRel,
                               PFDevice* dev, PFHistory hist, hipStream_t s) {
  hipLaunchKernelGGL(k_pf_correct_finalize, dim3(1), dim3(FINAL_BLOCK), 0, s, partials, nb, P,
                     essRel, dev, hist);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

}  // namespace bihip
