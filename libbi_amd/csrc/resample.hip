// Resampling kernels: cumulative offspring (systematic, stratified),
// multinomial ancestors, the permuted ancestor layout and the ancestor
// gather.
//
// Replaces (ON_DEVICE):
//   SystematicResampler::cumulativeOffspring  share/src/bi/resampler/SystematicResampler.hpp:110-130
//   StratifiedResamplerGPU / Host::op         share/src/bi/host/resampler/StratifiedResamplerHost.hpp:25-74
//   MultinomialResamplerGPU::ancestors        share/src/bi/cuda/resampler/MultinomialResamplerGPU.hpp:26-51
//   cumulativeOffspringToAncestorsPermute     share/src/bi/host/resampler/ResamplerHost.hpp:133-168
//   cumulativeOffspringToAncestors            share/src/bi/host/resampler/ResamplerHost.hpp:109-131
//   ancestorsToOffspring                      share/src/bi/host/resampler/ResamplerHost.hpp:55-68
//   gather_rows                               share/src/bi/cuda/primitive/matrix_primitive.cuh:67-78
//
// The permuted layout is computed without the reference GPU's atomicMin
// claims (share/src/bi/cuda/resampler/ResamplerKernel.cuh:129-165): with
// e_i = max(o_i - 1, 0) and f_k = [o_k == 0], E and F their exclusive scans,
// slot k keeps particle k when o_k > 0, and otherwise receives the particle
// i whose extra copies cover rank F_k, i.e. the largest i with E_i <= F_k.
// This is the host's serial two-pointer layout, bit for bit.
#include "scan.hpp"

namespace bihip {

constexpr int RS_BLOCK = 256;

// O_i = min(n, (int)(Ws_i / W * n + a)) with every operation rounded as
// written (SystematicResampler.hpp:47-50).
__device__ __forceinline__ int systematic_offspring(double Ws, double W, int n, double a) {
#pragma clang fp contract(off)
  const double v = Ws / W * (double)n + a;
  const int o = (int)v;
  return o < n ? o : n;
}

// StratifiedResamplerHost.hpp:68-71
__device__ __forceinline__ double stratified_reach(double Ws, double W, int n) {
#pragma clang fp contract(off)
  return Ws / W * (double)n;
}

__device__ __forceinline__ int stratified_offspring(double reach, double alpha, int n) {
#pragma clang fp contract(off)
  const int o = (int)(reach + alpha);
  return o < n ? o : n;
}

__device__ __forceinline__ bool gate_open(const PFDevice* g) {
  return g == nullptr || (g->resample != 0 && g->status == BIHIP_OK);
}

struct OffspringArgs {
  int mode;
  const double* Ws;
  int P, n;
  double a;
  const double* alphas;
  int philox;
  uint64_t seed;
  uint32_t step, batch;
  const PFDevice* gate;
  PFDevice* status;
  int* Os;
};

__global__ __launch_bounds__(RS_BLOCK) void k_cumulative_offspring(OffspringArgs a) {
  if (!gate_open(a.gate)) return;
  const double W = a.Ws[a.P - 1];
  if (!(W > 0.0)) {
    if (a.status != nullptr && blockIdx.x == 0 && threadIdx.x == 0)
      a.status->status = BIHIP_DEGENERATE;
    return;
  }
  const int i = blockIdx.x * RS_BLOCK + threadIdx.x;
  if (i >= a.P) return;
  const double Wi = a.Ws[i];
  int o;
  const uint32_t step = a.gate != nullptr ? (uint32_t)a.gate->k : a.step;
  if (a.mode == BIHIP_SYSTEMATIC) {
    double av = a.a;
    if (a.philox) {
      const u32x4 r = philox_block(a.seed, 0, step, TAG_SYSTEMATIC, a.batch, 0);
      av = u53(r.x, r.y);
    }
    o = systematic_offspring(Wi, W, a.n, av);
  } else {
    const double reach = stratified_reach(Wi, W, a.n);
    int k = (int)reach;
    if (k > a.n - 1) k = a.n - 1;
    double alpha;
    if (a.philox) {
      const u32x4 r = philox_block(a.seed, (uint32_t)k, step, TAG_STRATIFIED, a.batch, 0);
      alpha = u53(r.x, r.y);
    } else {
      alpha = a.alphas[k];
    }
    o = stratified_offspring(reach, alpha, a.n);
  }
  a.Os[i] = o;
}

int launch_cumulative_offspring(const OffspringLaunch& L, hipStream_t s) {
  if (L.P <= 0) return BIHIP_OK;
  OffspringArgs a{L.mode, L.Ws, L.P, L.n, L.a, L.alphas, L.philox ? 1 : 0, L.seed, L.step,
                  L.batch, L.gate, L.status, L.Os};
  hipLaunchKernelGGL(k_cumulative_offspring, dim3(ceil_div(L.P, RS_BLOCK)), dim3(RS_BLOCK), 0, s,
                     a);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

// ---------------------------------------------------------------- permuted layout
__device__ __forceinline__ int offspring_at(const int* in, bool cumulative, long i) {
  return cumulative ? in[i] - (i > 0 ? in[i - 1] : 0) : in[i];
}

struct PermLoader {
  const int* in;
  int cumulative;
  __device__ uint64_t operator()(long i) const {
    const int o = offspring_at(in, cumulative != 0, i);
    const uint64_t e = o > 1 ? (uint64_t)(o - 1) : 0u;
    const uint64_t f = o == 0 ? 1u : 0u;
    return e | (f << 32);
  }
};

struct StoreExclusive64 {
  uint64_t* out;
  __device__ void operator()(long i, uint64_t, uint64_t excl) const { out[i] = excl; }
};

__global__ __launch_bounds__(RS_BLOCK) void k_place(const int* in, int cumulative, int P,
                                                    const uint64_t* EF, int* as,
                                                    const PFDevice* gate) {
  if (!gate_open(gate)) return;
  const int k = blockIdx.x * RS_BLOCK + threadIdx.x;
  if (k >= P) return;
  const int o = offspring_at(in, cumulative != 0, k);
  if (o > 0) {
    as[k] = k;
    return;
  }
  const uint32_t r = (uint32_t)(EF[k] >> 32);  // rank of this free slot
  int lo = 0, hi = P;                          // first i with E_i > r
  while (lo < hi) {
    const int mid = lo + ((hi - lo) >> 1);
    if ((uint32_t)EF[mid] <= r) lo = mid + 1;
    else hi = mid;
  }
  as[k] = lo > 0 ? lo - 1 : 0;
}

int launch_ancestors_permute(const int* in, bool cumulative, int P, int* as, uint64_t* EF,
                             ScanWorkspace ws, const PFDevice* gate, hipStream_t s) {
  if (P <= 0) return BIHIP_OK;
  int rc = launch_tile_scan<uint64_t>(PermLoader{in, cumulative ? 1 : 0}, StoreExclusive64{EF},
                                     PFGate{gate}, P, ws, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_place, dim3(ceil_div(P, RS_BLOCK)), dim3(RS_BLOCK), 0, s, in,
                     cumulative ? 1 : 0, P, EF, as, gate);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

// as_j = #{i : O_i <= j}: slot j of the unpermuted ancestor vector.
__global__ __launch_bounds__(RS_BLOCK) void k_cumulative_to_ancestors(const int* Os, int P,
                                                                      int* as) {
  const int n = Os[P - 1] < P ? Os[P - 1] : P;
  const int j = blockIdx.x * RS_BLOCK + threadIdx.x;
  if (j >= n) return;
  int lo = 0, hi = P;
  while (lo < hi) {
    const int mid = lo + ((hi - lo) >> 1);
    if (Os[mid] <= j) lo = mid + 1;
    else hi = mid;
  }
  as[j] = lo < P ? lo : P - 1;
}

int launch_cumulative_to_ancestors(const int* Os, int P, int* as, hipStream_t s) {
  if (P <= 0) return BIHIP_OK;
  // the number of ancestors is Os[P-1], bounded by the caller's buffer of P
  hipLaunchKernelGGL(k_cumulative_to_ancestors, dim3(ceil_div(P, RS_BLOCK)), dim3(RS_BLOCK), 0, s,
                     Os, P, as);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

// ---------------------------------------------------------------- multinomial
struct MultinomialArgs {
  const double* Ws;
  int P, n;
  const double* U;
  uint64_t seed;
  uint32_t step, batch;
  const PFDevice* gate;
  PFDevice* status;
  int* as;
  int* os;
};

// a_j = #{i : Ws_i <= W U_j}: upper bound, never a zero-weight particle.
__global__ __launch_bounds__(RS_BLOCK) void k_multinomial(MultinomialArgs a) {
  if (!gate_open(a.gate)) return;
  const double W = a.Ws[a.P - 1];
  if (!(W > 0.0)) {
    if (a.status != nullptr && blockIdx.x == 0 && threadIdx.x == 0)
      a.status->status = BIHIP_DEGENERATE;
    return;
  }
  const int j = blockIdx.x * RS_BLOCK + threadIdx.x;
  if (j >= a.n) return;
  double U;
  if (a.U != nullptr) {
    U = a.U[j];
  } else {
    const uint32_t step = a.gate != nullptr ? (uint32_t)a.gate->k : a.step;
    const u32x4 r = philox_block(a.seed, (uint32_t)j, step, TAG_MULTINOMIAL, a.batch, 0);
    U = u53(r.x, r.y);
  }
  const double u = W * U;
  int lo = 0, hi = a.P;
  while (lo < hi) {
    const int mid = lo + ((hi - lo) >> 1);
    if (a.Ws[mid] <= u) lo = mid + 1;
    else hi = mid;
  }
  if (lo >= a.P) lo = a.P - 1;
  if (a.as != nullptr) a.as[j] = lo;
  if (a.os != nullptr) atomicAdd(&a.os[lo], 1);
}

int launch_multinomial(const MultinomialLaunch& L, hipStream_t s) {
  if (L.n <= 0) return BIHIP_OK;
  MultinomialArgs a{L.Ws, L.P, L.n, L.U, L.seed, L.step, L.batch, L.gate, L.status, L.as, L.os};
  hipLaunchKernelGGL(k_multinomial, dim3(ceil_div(L.n, RS_BLOCK)), dim3(RS_BLOCK), 0, s, a);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

__global__ void k_zero_ints(int* p, int n, const PFDevice* gate) {
  if (!gate_open(gate)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

int launch_zero_ints(int* p, int n, const PFDevice* gate, hipStream_t s) {
  if (n <= 0) return BIHIP_OK;
  hipLaunchKernelGGL(k_zero_ints, dim3(ceil_div(n, 256)), dim3(256), 0, s, p, n, gate);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

__global__ void k_histogram(const int* as, int n, int P, int* os) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int a = as[j];
  if (a >= 0 && a < P) atomicAdd(&os[a], 1);
}

int launch_ancestors_to_offspring(const int* as, int n, int P, int* os, hipStream_t s) {
  if (n <= 0) return BIHIP_OK;
  hipLaunchKernelGGL(k_histogram, dim3(ceil_div(n, 256)), dim3(256), 0, s, as, n, P, os);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

// ---------------------------------------------------------------- gather
template <class E>
__global__ void k_gather_rows(const int* as, int P, int ncols, E* X, int ld) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= P || j >= ncols) return;
  const int src = as[i];
  if (src != i && src >= 0 && src < P) X[(size_t)j * ld + i] = X[(size_t)j * ld + src];
}

int launch_gather_rows(const int* as, int P, int ncols, void* X, int ld, int elem,
                       hipStream_t s) {
  if (P <= 0 || ncols <= 0) return BIHIP_OK;
  const dim3 grid(ceil_div(P, 256), ncols);
  if (elem == 8) {
    hipLaunchKernelGGL(k_gather_rows<double>, grid, dim3(256), 0, s, as, P, ncols,
                       static_cast<double*>(X), ld);
  } else if (elem == 4) {
    hipLaunchKernelGGL(k_gather_rows<float>, grid, dim3(256), 0, s, as, P, ncols,
                       static_cast<float*>(X), ld);
  } else {
    return set_error(BIHIP_EINVAL, "gather_rows: element size must be 4 or 8");
  }
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

// Ancestry output of the step: as (resampled) or the identity.
__global__ void k_record_ancestors(const int* as, int P, const PFDevice* dev, int* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const bool r = dev->resample != 0 && dev->status == BIHIP_OK;
  out[(size_t)dev->k * P + i] = r ? as[i] : i;
}

int launch_record_ancestors(const int* as, int P, const PFDevice* gate, int* out,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_record_ancestors, dim3(ceil_div(P, 256)), dim3(256), 0, s, as, P, gate,
                     out);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

// The state after the step into slot k + 1 of the history (slot 0 is the
// initial state): (T + 1) x D columns of leading dimension ld.
template <class T>
__global__ void k_record_state(const T* X, int ld, int D, int P, const PFDevice* dev, T* hist,
                               int slot0) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long)P * D) return;
  const int c = (int)(g / P), p = (int)(g % P);
  const int slot = dev != nullptr ? dev->k + 1 : slot0;
  hist[((size_t)slot * D + c) * ld + p] = X[(size_t)c * ld + p];
}

int launch_record_state(const void* X, int ld, int D, int P, int elem, const PFDevice* dev,
                        void* hist, int slot0, hipStream_t s) {
  const dim3 grid(ceil_div((long)P * D, 256)), block(256);
  if (elem == 8)
    hipLaunchKernelGGL(k_record_state<double>, grid, block, 0, s, static_cast<const double*>(X), ld,
                       D, P, dev, static_cast<double*>(hist), slot0);
  else
    hipLaunchKernelGGL(k_record_state<float>, grid, block, 0, s, static_cast<const float*>(X), ld,
                       D, P, dev, static_cast<float*>(hist), slot0);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

// Paths of particles idx[q] at the current time K, traced back through the
// recorded ancestors: out[q][k][c] = state c at time k of the lineage
// (AncestryCache::readPath, share/src/bi/cache/AncestryCache.hpp:328-352).
template <class T>
__global__ void k_paths(const int* anc, const T* hist, int ld, int D, int P, int K, const int* idx,
                        int n, double* out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  int p = idx[q];
  for (int k = K; k >= 0; --k) {
    for (int c = 0; c < D; ++c)
      out[((size_t)q * (K + 1) + k) * D + c] = (double)hist[((size_t)k * D + c) * ld + p];
    if (k > 0) p = anc[(size_t)(k - 1) * P + p];
  }
}

int launch_paths(const int* anc, const void* hist, int ld, int D, int P, int K, int elem,
                 const int* idx, int n, double* out, hipStream_t s) {
  if (n <= 0) return BIHIP_OK;
  const dim3 grid(ceil_div(n, 64)), block(64);
  if (elem == 8)
    hipLaunchKernelGGL(k_paths<double>, grid, block, 0, s, anc, static_cast<const double*>(hist), ld,
                       D, P, K, idx, n, out);
  else
    hipLaunchKernelGGL(k_paths<float>, grid, block, 0, s, anc, static_cast<const float*>(hist), ld,
                       D, P, K, idx, n, out);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

__global__ void k_fill_f64(double* x, int n, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

int launch_fill_f64(double* x, int n, double v, hipStream_t s) {
  if (n <= 0) return BIHIP_OK;
  hipLaunchKernelGGL(k_fill_f64, dim3(ceil_div(n, 256)), dim3(256), 0, s, x, n, v);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

__global__ void k_iota(int* x, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = i;
}

int launch_iota(int* x, int n, hipStream_t s) {
  if (n <= 0) return BIHIP_OK;
  hipLaunchKernelGGL(k_iota, dim3(ceil_div(n, 256)), dim3(256), 0, s, x, n);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

}  // namespace bihip
