// Log-weight reductions: max, log-sum-exp and ESS in one read of lws,
// as per-tile (max, sum w, sum w^2) partials combined in a fixed order.
//
// Replaces ess_reduce / logsumexp_reduce / max_reduce
// (share/src/bi/primitive/vector_primitive.hpp:860-974) and the
// bookkeeping of BootstrapPF::correct (share/src/bi/filter/BootstrapPF.hpp:171-181)
// with Resampler::resample's trigger (share/src/bi/resampler/Resampler.hpp:168).
#include "scan.hpp"

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

// (m, s1, s2) triples -> (M, S1, S2) over [lo, hi) in a fixed order; every
// thread of the block gets the totals.  `load(i, m, s1, s2)` fetches triple i.
template <class Load>
__device__ __forceinline__ void combine(int lo, int hi, Load load, double* lds, double& M, double& S1,
                        double& S2) {
  double m = __builtin_nan("");
  for (int i = lo + threadIdx.x; i < hi; i += COMB_BLOCK) {
    double a, b, c;
    load(i, a, b, c);
    m = nan_max(m, a);
  }
  M = block_max<COMB_BLOCK>(m, lds);
  double s1 = 0.0, s2 = 0.0;
  for (int i = lo + threadIdx.x; i < hi; i += COMB_BLOCK) {
    double a, b, c;
    load(i, a, b, c);
    if (b != 0.0) {
      const double f = exp(a - M);
      s1 += b * f;
      s2 += c * (f * f);
    }
  }
  S1 = block_sum<COMB_BLOCK>(s1, lds);
  S2 = block_sum<COMB_BLOCK>(s2, lds);
}

// BootstrapPF::correct's bookkeeping (share/src/bi/filter/BootstrapPF.hpp:171-181)
// from the totals (M, S1, S2) of the step's log-weights, with Resampler::
// resample's decision (Resampler.hpp:168) for the next time; with out, the
// totals {max, sum w, sum w^2, ess, lW}.  One thread.
__device__ __forceinline__ void pf_bookkeeping(double M, double S1, double S2, int P,
                                               double essRel, PFDevice* dev, PFHistory hist,
                                               double* out, const int* mask) {
  const double ess = S1 * S1 / S2;
  const double lW = M + log(S1) - log((double)P);
  if (out != nullptr) {
    out[0] = M;
    out[1] = S1;
    out[2] = S2;
    out[3] = ess;
    out[4] = lW;
  }
  if (dev != nullptr) {
    const int k = dev->k;
    if (hist.resampled) hist.resampled[k] = dev->resample != 0 && dev->status == BIHIP_OK;
    if (mask == nullptr || mask[k] != 0) {
      if (hist.logIncrements) hist.logIncrements[k] = lW - dev->logLikelihood;
      if (hist.ess) hist.ess[k] = ess;
      dev->logLikelihood = lW;
      dev->ess = ess;
      dev->mx = M;
      dev->lW = lW;
      dev->resample = ess < essRel * (double)P ? 1 : 0;
    } else {
      // no observation: nothing to correct, and Resampler::resample acts only
      // at observed elements (share/src/bi/resampler/Resampler.hpp:168)
      if (hist.logIncrements) hist.logIncrements[k] = 0.0;
      if (hist.ess) hist.ess[k] = dev->ess;
      dev->resample = 0;
    }
    dev->k = k + 1;
  }
}

// Two-level combination of nb partials: block b combines partials
// [b*COMB_BLOCK, (b+1)*COMB_BLOCK) into a level-2 triple, published as
// epoch-tagged words; the block that takes the last ticket combines the
// level-2 triples (fixed order, so the result does not depend on which block
// is last) and finishes:
//   - with dev: BootstrapPF::correct's bookkeeping -- ess, lW = ess_reduce(lws),
//     logIncrements[k] = lW - logLikelihood, logLikelihood = lW, and the
//     resampling decision ess < essRel * P for the next time;
//   - with out: {max, sum w, sum w^2, ess, lW}.
__global__ __launch_bounds__(COMB_BLOCK) void k_reduce(const double* partials, int nb, int P,
                                                       double essRel, PFDevice* dev,
                                                       PFHistory hist, double* out,
                                                       ReduceWork w, const int* mask) {
  __shared__ double lds[COMB_BLOCK / WAVE];
  __shared__ int last;
  __shared__ uint32_t s_epoch;
  if (threadIdx.x == 0)
    s_epoch = __hip_atomic_load(&w.state->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u;
  __syncthreads();
  const uint32_t epoch = s_epoch;
  const int lo = blockIdx.x * COMB_BLOCK, hi = min(nb, lo + COMB_BLOCK);
  double M, S1, S2;
  combine(lo, hi,
          [&](int i, double& a, double& b, double& c) {
            a = partials[3 * i];
            b = partials[3 * i + 1];
            c = partials[3 * i + 2];
          },
          lds, M, S1, S2);
  if (threadIdx.x == 0) {
    uint64_t* my = w.words + 6 * (size_t)blockIdx.x;
    publish(my, epoch, M);
    publish(my + 2, epoch, S1);
    publish(my + 4, epoch, S2);
    const uint32_t t =
        __hip_atomic_fetch_add(&w.state->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    last = t == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;

  combine(0, (int)gridDim.x,
          [&](int i, double& a, double& b, double& c) {
            uint64_t* q = w.words + 6 * (size_t)i;
            a = await_value<double>(q, epoch, &w.state->error);
            b = await_value<double>(q + 2, epoch, &w.state->error);
            c = await_value<double>(q + 4, epoch, &w.state->error);
          },
          lds, M, S1, S2);
  if (threadIdx.x != 0) return;
  __hip_atomic_store(&w.state->ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(&w.state->epoch, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  pf_bookkeeping(M, S1, S2, P, essRel, dev, hist, out, mask);
}

// One workgroup per filter of a bank (filter blockIdx.x): the filter's nb
// partials (after the nb of each filter before it), thread t taking partials
// t, t + BANK_RED, ... in a fixed order -- maximum first, then the rescaled
// sums -- and the filter's bookkeeping, with its PFDevice at dev + f and its
// histories at f * T.
constexpr int BANK_RED = 1024;

__global__ __launch_bounds__(BANK_RED) void k_correct_bank(const double* partials, int nb, int P,
                                                           double essRel, PFDevice* dev,
                                                           PFHistory hist, int T) {
  __shared__ double lds[BANK_RED / WAVE];
  const long f = blockIdx.x;
  const double* q = partials + 3 * f * nb;
  double m = __builtin_nan("");
#pragma unroll 4
  for (int i = threadIdx.x; i < nb; i += BANK_RED) m = nan_max(m, q[3 * i]);
  const double M = block_max<BANK_RED>(m, lds);
  double s1 = 0.0, s2 = 0.0;
#pragma unroll 4
  for (int i = threadIdx.x; i < nb; i += BANK_RED) {
    const double a = q[3 * i], b = q[3 * i + 1], c = q[3 * i + 2];
    if (b != 0.0) {
      const double e = exp(a - M);
      s1 += b * e;
      s2 += c * (e * e);
    }
  }
  const double S1 = block_sum<BANK_RED>(s1, lds);
  const double S2 = block_sum<BANK_RED>(s2, lds);
  if (threadIdx.x != 0) return;
  PFHistory h{hist.logIncrements ? hist.logIncrements + f * T : nullptr,
              hist.ess ? hist.ess + f * T : nullptr,
              hist.resampled ? hist.resampled + f * T : nullptr};
  pf_bookkeeping(M, S1, S2, P, essRel, dev + f, h, nullptr, nullptr);
}

// The same combination in two launches without tickets or epochs: block b of
// k_combine stores the triple of partials [b COMB_BLOCK, (b + 1) COMB_BLOCK)
// with plain stores, then one block of k_finish combines the triples in order
// and finishes.  The combining order is k_reduce's, so are the results, bit
// for bit; with nb <= COMB_BLOCK k_finish alone combines the partials (one
// launch, no device-scope atomics: the small filters' correct step).  At
// 16384 partials the two launches take 4.7 + 5.0 us against k_reduce's 9 us:
// each launch's own latency, not the ticket, sets the time.
__global__ __launch_bounds__(COMB_BLOCK) void k_combine(const double* partials, int nb,
                                                        double* lvl) {
  __shared__ double lds[COMB_BLOCK / WAVE];
  const int lo = blockIdx.x * COMB_BLOCK, hi = min(nb, lo + COMB_BLOCK);
  double M, S1, S2;
  combine(lo, hi,
          [&](int i, double& a, double& b, double& c) {
            a = partials[3 * i];
            b = partials[3 * i + 1];
            c = partials[3 * i + 2];
          },
          lds, M, S1, S2);
  if (threadIdx.x == 0) {
    lvl[3 * blockIdx.x] = M;
    lvl[3 * blockIdx.x + 1] = S1;
    lvl[3 * blockIdx.x + 2] = S2;
  }
}

__global__ __launch_bounds__(COMB_BLOCK) void k_finish(const double* trip, int nb, int P,
                                                       double essRel, PFDevice* dev,
                                                       PFHistory hist, double* out,
                                                       const int* mask) {
  __shared__ double lds[COMB_BLOCK / WAVE];
  double M, S1, S2;
  combine(0, nb,
          [&](int i, double& a, double& b, double& c) {
            a = trip[3 * i];
            b = trip[3 * i + 1];
            c = trip[3 * i + 2];
          },
          lds, M, S1, S2);
  if (threadIdx.x == 0) pf_bookkeeping(M, S1, S2, P, essRel, dev, hist, out, mask);
}

// k_finish with the next step's tile weight sums, one block per tile of
// SCAN_TILE log-weights: every block combines the triples in k_finish's
// order (the same M, S1, S2 in each), block 0 does the bookkeeping, and
// each block stores its tile's sum of fix_weight(nanexp(lw - M)), the value
// k_weight_sums computes from dev->mx = M, when the step leaves the gate
// open: ess < essRel P (the decision pf_bookkeeping takes), or any step under
// a mask, whose decision needs the step number that block 0 moves on.  The
// log-weights are loaded first so that their latency overlaps the
// combination's.
__global__ __launch_bounds__(COMB_BLOCK) void k_finish_sums(const double* trip, int nb, int P,
                                                            double essRel, PFDevice* dev,
                                                            PFHistory hist, const int* mask,
                                                            const double* lws, u128* wsum) {
  static_assert(COMB_BLOCK == SCAN_BLOCK, "one thread per SCAN_ITEMS log-weights");
  __shared__ double lds[COMB_BLOCK / WAVE];
  __shared__ u128 lds_w[SCAN_BLOCK / WAVE];
  const long t0 = (long)blockIdx.x * SCAN_TILE;
  double lw[SCAN_ITEMS];
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    const long i = t0 + (long)j * SCAN_BLOCK + threadIdx.x;
    lw[j] = i < P ? lws[i] : __builtin_nan("");  // NaN weighs zero
  }
  double M, S1, S2;
  combine(0, nb,
          [&](int i, double& a, double& b, double& c) {
            a = trip[3 * i];
            b = trip[3 * i + 1];
            c = trip[3 * i + 2];
          },
          lds, M, S1, S2);
  if (blockIdx.x == 0 && threadIdx.x == 0)
    pf_bookkeeping(M, S1, S2, P, essRel, dev, hist, nullptr, mask);
  const double ess = S1 * S1 / S2;
  if (mask == nullptr && !(ess < essRel * (double)P)) return;  // uniform across the grid
  u128 t(0);
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) t = t + fix_weight(nanexp(lw[j] - M));
  t = block_sum_v(t, lds_w);
  if (threadIdx.x == 0) wsum[blockIdx.x] = t;
}

static int launch_two_level(const double* partials, int nb, int P, double essRel, PFDevice* dev,
                            PFHistory hist, double* out, const ReduceWork& w, hipStream_t s,
                            const int* mask) {
  if (nb <= COMB_BLOCK) {
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(COMB_BLOCK), 0, s, partials, nb, P, essRel, dev,
                       hist, out, mask);
    BIHIP_LAUNCH_CHECK();
    return BIHIP_OK;
  }
  const int nb2 = reduce_blocks(nb);
  if (nb2 > COMB_BLOCK || nb2 > w.max_blocks)
    return set_error(BIHIP_EINVAL, "too many partials for two levels");
  // plain triples after the epoch-tagged words (which k_reduce alone reads)
  double* lvl = reinterpret_cast<double*>(w.words + 6 * (size_t)w.max_blocks);
  hipLaunchKernelGGL(k_combine, dim3(nb2), dim3(COMB_BLOCK), 0, s, partials, nb, lvl);
  BIHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(COMB_BLOCK), 0, s, (const double*)lvl, nb2, P,
                     essRel, dev, hist, out, mask);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

int reduce_blocks(int nb) { return ceil_div(nb, COMB_BLOCK); }

int launch_lws_partials(const double* lws, int P, double* partials, hipStream_t s) {
  hipLaunchKernelGGL(k_lws_partials, dim3(ceil_div(P, RED_TILE)), dim3(RED_BLOCK), 0, s, lws, P,
                     partials);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

int launch_finalize(const double* partials, int nb, int P, double* out, ReduceWork w,
                    hipStream_t s) {
  if (reduce_blocks(nb) <= COMB_BLOCK)
    return launch_two_level(partials, nb, P, 0.0, nullptr, PFHistory{}, out, w, s, nullptr);
  hipLaunchKernelGGL(k_reduce, dim3(reduce_blocks(nb)), dim3(COMB_BLOCK), 0, s, partials, nb, P,
                     0.0, (PFDevice*)nullptr, PFHistory{}, out, w, (const int*)nullptr);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

bool correct_stores_sums(int nb) { return reduce_blocks(nb) <= COMB_BLOCK; }

int launch_pf_correct_finalize(const double* partials, int nb, int P, double essRel,
                               PFDevice* dev, PFHistory hist, ReduceWork w, hipStream_t s,
                               const int* mask, const double* lws, uint64_t* tile_sums) {
  if (tile_sums != nullptr) {
    if (!correct_stores_sums(nb) || lws == nullptr || dev == nullptr || P <= 0)
      return set_error(BIHIP_EINVAL, "the correct step stores tile sums up to %d partials",
                       COMB_BLOCK * COMB_BLOCK);
    const double* trip = partials;
    int n = nb;
    if (nb > COMB_BLOCK) {
      n = reduce_blocks(nb);
      if (n > w.max_blocks) return set_error(BIHIP_EINVAL, "too many partials for two levels");
      double* lvl = reinterpret_cast<double*>(w.words + 6 * (size_t)w.max_blocks);
      hipLaunchKernelGGL(k_combine, dim3(n), dim3(COMB_BLOCK), 0, s, partials, nb, lvl);
      BIHIP_LAUNCH_CHECK();
      trip = lvl;
    }
    hipLaunchKernelGGL(k_finish_sums, dim3(ceil_div(P, SCAN_TILE)), dim3(COMB_BLOCK), 0, s, trip,
                       n, P, essRel, dev, hist, mask, lws, reinterpret_cast<u128*>(tile_sums));
    BIHIP_LAUNCH_CHECK();
    return BIHIP_OK;
  }
  if (reduce_blocks(nb) <= COMB_BLOCK)
    return launch_two_level(partials, nb, P, essRel, dev, hist, nullptr, w, s, mask);
  hipLaunchKernelGGL(k_reduce, dim3(reduce_blocks(nb)), dim3(COMB_BLOCK), 0, s, partials, nb, P,
                     essRel, dev, hist, (double*)nullptr, w, mask);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

int launch_correct_bank(const double* partials, int nb, int P, double essRel, PFDevice* dev,
                        PFHistory hist, int T, int banks, hipStream_t s) {
  hipLaunchKernelGGL(k_correct_bank, dim3(banks), dim3(BANK_RED), 0, s, partials, nb, P, essRel,
                     dev, hist, T);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

}  // namespace bihip
