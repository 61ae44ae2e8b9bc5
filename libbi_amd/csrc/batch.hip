// Batched inner filters: B bootstrap particle filters with uneven particle
// counts N_j and their own parameters theta_j, advanced one observation
// interval per persistent fused launch.
//
// This is the inner loop of MarginalSIR::step (share/src/bi/sampler/MarginalSIR.hpp:346-368),
// which the reference runs serially over theta-particles with equal N
// (share/src/bi/state/MarginalSIRState.hpp:198-207).  Here every workgroup
// takes whole filters from a work queue (largest first) and runs the full
// BootstrapPF step for each: the resampling decision, the exact fixed-point
// weight scan, systematic/stratified offspring, the permuted ancestors (a
// search table in LDS), the gathered transition, the observation density and
// the ESS / log-likelihood reduction.  Filter j draws its Philox streams
// with its own batch id (stream base + j at initialisation, kept when the
// filter's content is replaced by a copy), so each inner filter equals the
// single-filter path run with that batch id.
#include <type_traits>

#include "model_ops.hpp"
#include "scan.hpp"
#include "gen_model.hpp"

namespace bihip {

// Largest inner filter: the permutation's search table E (values below N)
// is kept in LDS as 16-bit integers.
constexpr int BATCH_NMAX = 16384;
// Consecutive particles per thread in the resampling passes: the kernel's
// template parameter ITEMS, 4 unless the autotuner picks another width.  The
// hand-written predator-prey model in FP64 with RK4 (the SMC^2 workload) is
// compiled with widths 2, 4 and 8; every other model with 4.
constexpr int BITEMS_DEFAULT = 4;

struct BatchArgs {
  ModelArgs m;
  const double* theta;  // B x ntheta: par[0..ntheta) of filter j
  int ntheta;
  const int* order;     // processing order (largest filter first)
  int B;
  BatchFilter* f;
  const void* Xin;
  void* Xout;
  long ld;              // column stride of the flattened state
  double* lws;
  int* as;              // scratch: ancestors
  int* anc_rec;         // nullable: ancestors of step k at anc_rec[k * ld + off + i]
  double* hist_li;      // B x T
  double* hist_ess;     // B x T
  int* hist_res;        // B x T
  const double* y;      // T x nobs
  int nobs, T;
  double t0, dt;
  uint64_t seed;
  double essRel;
  int resampler;
  uint32_t* counter;    // work queue
};

template <class M>
__device__ __forceinline__ M make_model(const BatchArgs& a, int j) {
  ModelArgs ma = a.m;
#pragma unroll
  for (int i = 0; i < 16; ++i)  // constant indices keep par[] in registers
    if (i < a.ntheta) ma.par[i] = a.theta[(size_t)j * a.ntheta + i];
  return M(ma);
}

template <int BLOCK>
__device__ __forceinline__ u128 block_incl_scan_u128(u128 v, u128* lds, u128& total) {
  constexpr int NW = BLOCK / WAVE;
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const u128 u = shfl_up_v(v, o);
    if (lane >= o) v = u + v;
  }
  if (lane == WAVE - 1) lds[w] = v;
  __syncthreads();
  u128 pre(0), tot(0);
#pragma unroll
  for (int i = 0; i < NW; ++i) {
    if (i < w) pre = pre + lds[i];
    tot = tot + lds[i];
  }
  __syncthreads();
  total = tot;
  return pre + v;
}

template <int BLOCK>
__device__ __forceinline__ uint64_t block_excl_scan_u64(uint64_t v, uint64_t* lds,
                                                        uint64_t& total) {
  constexpr int NW = BLOCK / WAVE;
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  uint64_t incl = v;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const uint64_t u = __shfl_up(incl, o, WAVE);
    if (lane >= o) incl += u;
  }
  if (lane == WAVE - 1) lds[w] = incl;
  __syncthreads();
  uint64_t pre = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < NW; ++i) {
    if (i < w) pre += lds[i];
    tot += lds[i];
  }
  __syncthreads();
  total = tot;
  return pre + incl - v;
}

__device__ __forceinline__ int sys_offspring(double Ws, double W, int n, double a) {
#pragma clang fp contract(off)
  const double v = Ws / W * (double)n + a;
  const int o = (int)v;
  return o < n ? o : n;
}

__device__ __forceinline__ int strat_offspring(double Ws, double W, int n, uint64_t seed,
                                               uint32_t step, uint32_t batch) {
#pragma clang fp contract(off)
  const double reach = Ws / W * (double)n;
  int k = (int)reach;
  if (k > n - 1) k = n - 1;
  const u32x4 r = philox_block(seed, (uint32_t)k, step, TAG_STRATIFIED, batch, 0);
  const double alpha = u53(r.x, r.y);
  const int o = (int)(reach + alpha);
  return o < n ? o : n;
}

template <class M, int INTEG, int BLOCK, int BITEMS = BITEMS_DEFAULT>
__global__ __launch_bounds__(BLOCK) void k_batch_step(BatchArgs a) {
  using T = typename M::real_t;
  constexpr int DL = M::DL;
  static_assert(M::LPP == 1, "batched filters use one lane per particle");
  __shared__ uint16_t E[BATCH_NMAX];
  __shared__ u128 l128[BLOCK / WAVE];
  __shared__ uint64_t l64[BLOCK / WAVE];
  __shared__ double ld[BLOCK / WAVE];
  __shared__ int s_job;
  __shared__ int s_res;
  const int tid = threadIdx.x;
  const Group<1> grp;

  for (;;) {
    if (tid == 0)
      s_job = (int)__hip_atomic_fetch_add(a.counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int job = __builtin_amdgcn_readfirstlane(s_job);  // uniform: filter data in SGPRs
    if (job >= a.B) break;
    const int j = a.order[job];
    const BatchFilter F = a.f[j];
    const int N = F.N, k = F.k;
    const long off = F.off;
    const uint32_t batch = F.stream;
    if (tid == 0) s_res = F.resample != 0 && F.status == BIHIP_OK;
    __syncthreads();
    bool res = s_res != 0;

    // ---- Resampler::resample: exact weight scan, offspring, permuted ancestors.
    // Each thread takes BITEMS consecutive particles of a tile of BLOCK * BITEMS:
    // (1) the exact fixed-point total W; (2) per tile, the inclusive prefix of
    // every particle (thread-local sums plus one block scan), its cumulative
    // offspring and the one of its predecessor (P_{i-1} = P_i - w_i, exact),
    // the exclusive (e, f) sums (one more block scan), E into LDS and the free
    // slots' ranks into as[]; (3) the free slots' ancestors by searches of E
    // that run in lockstep (a uniform number of halvings) with BITEMS in flight.
    if (res) {
      u128 part(0);
      for (int base = tid * BITEMS; base < N; base += BLOCK * BITEMS) {
#pragma unroll
        for (int u = 0; u < BITEMS; ++u)
          if (base + u < N) part = part + fix_weight(nanexp(a.lws[off + base + u] - F.mx));
      }
      u128 Wfix;
      (void)block_incl_scan_u128<BLOCK>(part, l128, Wfix);
      const double W = fix_to_double(Wfix);
      if (!(W > 0.0)) {
        if (tid == 0) a.f[j].status = BIHIP_DEGENERATE;
        res = false;
      } else {
        double av = 0.0;
        if (a.resampler == BIHIP_SYSTEMATIC) {
          const u32x4 r = philox_block(a.seed, 0, (uint32_t)k, TAG_SYSTEMATIC, batch, 0);
          av = u53(r.x, r.y);
        }
        auto offspring = [&](const u128& P) {  // cumulative offspring at prefix P
          const double Ws = fix_to_double(P);
          return a.resampler == BIHIP_SYSTEMATIC
                     ? sys_offspring(Ws, W, N, av)
                     : strat_offspring(Ws, W, N, a.seed, (uint32_t)k, batch);
        };
        u128 carry(0);
        uint64_t efcarry = 0;
        for (int t0 = 0; t0 < N; t0 += BLOCK * BITEMS) {
          const int base = t0 + tid * BITEMS;
          u128 w[BITEMS];
          u128 loc(0);
#pragma unroll
          for (int u = 0; u < BITEMS; ++u) {
            w[u] = base + u < N ? fix_weight(nanexp(a.lws[off + base + u] - F.mx)) : u128(0);
            loc = loc + w[u];
          }
          u128 tot;
          const u128 incl = block_incl_scan_u128<BLOCK>(loc, l128, tot);
          u128 P = carry + (incl - loc);  // prefix before this thread's first particle
          int O0 = base > 0 && base < N ? offspring(P) : 0;
          uint64_t ef[BITEMS];
          int oo[BITEMS];
          uint64_t eloc = 0;
#pragma unroll
          for (int u = 0; u < BITEMS; ++u) {
            P = P + w[u];
            const int O1 = base + u < N ? offspring(P) : O0;
            const int o = O1 - O0;
            O0 = O1;
            oo[u] = o;
            const uint64_t e = o > 1 ? (uint64_t)(o - 1) : 0u;
            const uint64_t fz = (base + u < N && o == 0) ? 1u : 0u;
            ef[u] = eloc;  // exclusive within the thread
            eloc += e | (fz << 32);
          }
          uint64_t eftot;
          const uint64_t ex0 = efcarry + block_excl_scan_u64<BLOCK>(eloc, l64, eftot);
#pragma unroll
          for (int u = 0; u < BITEMS; ++u) {
            const int i = base + u;
            if (i < N) {
              const uint64_t ex = ex0 + ef[u];
              E[i] = (uint16_t)(uint32_t)ex;
              a.as[off + i] = oo[u] > 0 ? i : -1 - (int)(uint32_t)(ex >> 32);
            }
          }
          carry = carry + tot;
          efcarry += eftot;
        }
        __syncthreads();  // E complete
        // free slot of rank r <- the last i with E_i <= r (E is non-decreasing)
        const int top = 1 << (31 - __clz(N));  // largest power of two <= N
        for (int base = tid * BITEMS; base < N; base += BLOCK * BITEMS) {
          int v[BITEMS], pos[BITEMS];
#pragma unroll
          for (int u = 0; u < BITEMS; ++u) {
            v[u] = base + u < N ? a.as[off + base + u] : 0;
            pos[u] = -1;
          }
          for (int s = top; s > 0; s >>= 1) {
#pragma unroll
            for (int u = 0; u < BITEMS; ++u) {
              const int c = pos[u] + s;
              if (v[u] < 0 && c < N && (uint32_t)E[c] <= (uint32_t)(-1 - v[u])) pos[u] = c;
            }
          }
#pragma unroll
          for (int u = 0; u < BITEMS; ++u)
            if (base + u < N && v[u] < 0) a.as[off + base + u] = pos[u] > 0 ? pos[u] : 0;
        }
        __syncthreads();  // ancestors written by other threads than the ones that gather
      }
    }

    // ---- predict + correct (ancestor gather on read), weight partials
    M m = make_model<M>(a, j);
    const double t1 = a.t0 + k * a.dt, t2 = a.t0 + (k + 1) * a.dt;
    const uint32_t step = (uint32_t)(k + 1);
    // the step's observation in registers: the model's terms that depend on it
    // alone (the log of a log-normal observation) leave the particle loop
    constexpr int NY = M::NOBS > 0 ? M::NOBS : 1;
    double y[NY];
#pragma unroll
    for (int o = 0; o < NY; ++o) y[o] = o < M::NOBS ? a.y[(size_t)k * a.nobs + o] : 0.0;
    const T* Xin = static_cast<const T*>(a.Xin);
    T* Xout = static_cast<T*>(a.Xout);
    double tm = -__builtin_inf(), ts1 = 0.0, ts2 = 0.0;  // running (max, sum w, sum w^2)
    for (int q = 0; q < N; q += BLOCK) {
      const int i = q + tid;
      if (i >= N) continue;
      const int src = res ? a.as[off + i] : i;
      if (a.anc_rec != nullptr) a.anc_rec[(size_t)k * a.ld + off + i] = src;
      T x[DL];
#pragma unroll
      for (int c = 0; c < DL; ++c) x[c] = Xin[(size_t)c * a.ld + off + src];
      M mm = m;
      if constexpr (is_generated<M>::value) {
        mm.template transition<INTEG>(grp, x, t1, t2, a.m, a.seed, batch, (uint32_t)i, step);
      } else {
        if constexpr (M::NOISE_FIRST) noise_first(mm, a.seed, batch, i, step);
        if constexpr (INTEG == BIHIP_RK4) {
          integrate_rk4(mm, grp, x, t1, t2, a.m.h0);
        } else {
          integrate_dopri5(mm, grp, x, t1, t2, a.m);
        }
        if constexpr (!M::NOISE_FIRST) noise_after(mm, a.seed, batch, i, 0, step, x);
      }
#pragma unroll
      for (int c = 0; c < DL; ++c) Xout[(size_t)c * a.ld + off + i] = x[c];
      const double lw = (res ? F.logLikelihood : a.lws[off + i]) + obs_of(mm, x, y, 0);
      a.lws[off + i] = lw;
      if (!isnan(lw)) {
        if (lw > tm) {
          const double f = nanexp(tm - lw);
          ts1 = ts1 * f + 1.0;
          ts2 = ts2 * (f * f) + 1.0;
          tm = lw;
        } else {
          const double z = nanexp(lw - tm);
          ts1 += z;
          ts2 += z * z;
        }
      }
    }
    // block combination in a fixed order
    const double Mx = block_max<BLOCK>(ts1 > 0.0 ? tm : __builtin_nan(""), ld);
    double c1 = 0.0, c2 = 0.0;
    if (ts1 > 0.0) {
      const double f = exp(tm - Mx);
      c1 = ts1 * f;
      c2 = ts2 * (f * f);
    }
    const double S1 = block_sum<BLOCK>(c1, ld);
    const double S2 = block_sum<BLOCK>(c2, ld);
    if (tid == 0) {
      const double ess = S1 * S1 / S2;
      const double lW = Mx + log(S1) - log((double)N);
      const size_t h = (size_t)j * a.T + k;
      a.hist_li[h] = lW - F.logLikelihood;
      a.hist_ess[h] = ess;
      a.hist_res[h] = res ? 1 : 0;
      BatchFilter G = F;
      G.status = a.f[j].status;
      G.logLikelihood = lW;
      G.ess = ess;
      G.mx = Mx;
      G.lW = lW;
      G.resample = ess < a.essRel * (double)N ? 1 : 0;
      G.k = k + 1;
      a.f[j] = G;
    }
    __syncthreads();  // LDS and s_job reused by the next filter
  }
}

// Initial particles of every filter (filter of a particle by a search over offsets).
template <class M>
__global__ void k_batch_init(ModelArgs ma, const double* theta, int ntheta, const long* off, int B,
                             long total, uint64_t seed, uint32_t stream_base, void* Xv, long ld) {
  using T = typename M::real_t;
  constexpr int DL = M::DL;
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  int lo = 0, hi = B;  // last filter with off <= g
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= g) lo = mid;
    else hi = mid;
  }
  const int j = lo;
  const int p = (int)(g - off[j]);
  T* X = static_cast<T*>(Xv);
  if constexpr (is_generated<M>::value) {
    ModelArgs mj = ma;
    for (int i = 0; i < ntheta && i < 16; ++i) mj.par[i] = theta[(size_t)j * ntheta + i];
    double xd[DL];
    M(mj).init(xd, seed, stream_base + (uint32_t)j, (uint32_t)p, 0);
    for (int c = 0; c < DL; ++c) X[(size_t)c * ld + g] = (T)xd[c];
    return;
  }
  T z[DL];
  slice_normals<T, DL>(seed, (uint32_t)p, 0, TAG_INIT, stream_base + (uint32_t)j, 0, z);
  for (int c = 0; c < DL; ++c) {
    double par[16];
    for (int i = 0; i < 16; ++i) par[i] = i < ntheta ? theta[(size_t)j * ntheta + i] : ma.par[i];
    double v;
    if (ma.model == BIHIP_MODEL_PZ) {
      v = c == 0 ? exp(par[7] + par[8] * (double)z[0]) : exp(par[9] + par[10] * (double)z[1]);
    } else {
      v = par[3] + par[4] * (double)z[c];
    }
    X[(size_t)c * ld + g] = (T)v;
  }
}

__global__ void k_batch_reset(uint32_t* counter) { *counter = 0u; }

// Whole-filter copies: one lane per destination particle, the filter found by
// a search over the destination offsets (strictly increasing, N_j >= 1).
template <class T>
__global__ void k_batch_copy(const long* dst_off, const long* src_off, const int* len,
                             const int* mask, int B, long total, int d, const T* Xs, long lds,
                             const double* lws_s, T* Xd, long ldd, double* lws_d) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  int lo = 0, hi = B;  // last filter with dst_off <= g
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (dst_off[mid] <= g) lo = mid;
    else hi = mid;
  }
  const int j = lo;
  const long p = g - dst_off[j];
  if (p >= len[j] || (mask != nullptr && mask[j] == 0)) return;
  const long src = src_off[j] + p;
  for (int c = 0; c < d; ++c) Xd[(size_t)c * ldd + g] = Xs[(size_t)c * lds + src];
  lws_d[g] = lws_s[src];
}

int launch_batch_copy(const long* dst_off, const long* src_off, const int* len, const int* mask,
                      int B, long total, int d, int elem, const void* Xs, long lds,
                      const double* lws_s, void* Xd, long ldd, double* lws_d, hipStream_t s) {
  if (total <= 0) return BIHIP_OK;
  const dim3 grid(ceil_div(total, 256)), block(256);
  if (elem == 8) {
    hipLaunchKernelGGL(k_batch_copy<double>, grid, block, 0, s, dst_off, src_off, len, mask, B,
                       total, d, static_cast<const double*>(Xs), lds, lws_s,
                       static_cast<double*>(Xd), ldd, lws_d);
  } else {
    hipLaunchKernelGGL(k_batch_copy<float>, grid, block, 0, s, dst_off, src_off, len, mask, B,
                       total, d, static_cast<const float*>(Xs), lds, lws_s,
                       static_cast<float*>(Xd), ldd, lws_d);
  }
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

// ---------------------------------------------------------------- dispatch
// Tile widths compiled for a model and integrator (batch_items_for).
template <class M, int INTEG>
constexpr bool many_widths() {
  return std::is_same<M, PZ<double>>::value && INTEG == BIHIP_RK4;
}

template <int BLOCK>
struct BatchRunner {
  BatchArgs a;
  int grid, integ, items;
  hipStream_t s;
  template <class M, int INTEG>
  int launch() {
    if constexpr (many_widths<M, INTEG>()) {
      switch (items) {
        case 2:
          hipLaunchKernelGGL((k_batch_step<M, INTEG, BLOCK, 2>), dim3(grid), dim3(BLOCK), 0, s, a);
          BIHIP_LAUNCH_CHECK();
          return BIHIP_OK;
        case 8:
          hipLaunchKernelGGL((k_batch_step<M, INTEG, BLOCK, 8>), dim3(grid), dim3(BLOCK), 0, s, a);
          BIHIP_LAUNCH_CHECK();
          return BIHIP_OK;
        default:
          break;
      }
    }
    if (items != BITEMS_DEFAULT)
      return set_error(BIHIP_EINVAL, "tile width %d not compiled for this model", items);
    hipLaunchKernelGGL((k_batch_step<M, INTEG, BLOCK>), dim3(grid), dim3(BLOCK), 0, s, a);
    BIHIP_LAUNCH_CHECK();
    return BIHIP_OK;
  }
  template <class M>
  int run() {
    if constexpr (M::LPP != 1) {
      return set_error(BIHIP_EINVAL, "batched filters take models of one lane per particle");
    } else if (integ == BIHIP_RK4) {
      return launch<M, BIHIP_RK4>();
    } else {
      return launch<M, BIHIP_DOPRI5>();
    }
  }
  template <class M>
  int occupancy() {
    int n = 0;
    if constexpr (M::LPP != 1) {
      return 1;
    } else if (integ == BIHIP_RK4) {
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
              &n, reinterpret_cast<const void*>(&k_batch_step<M, BIHIP_RK4, BLOCK>), BLOCK, 0) !=
          hipSuccess)
        n = 1;
    } else {
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
              &n, reinterpret_cast<const void*>(&k_batch_step<M, BIHIP_DOPRI5, BLOCK>), BLOCK,
              0) != hipSuccess)
        n = 1;
    }
    return n;
  }
};

template <class F>
static int with_small_model(const bihip_model* m, F& f) {
  const bool f64 = m->precision == BIHIP_F64;
  if (m->model >= BIHIP_MODEL_GENERATED) return with_generated(m, f);
  if (m->model == BIHIP_MODEL_PZ)
    return f64 ? f.template run<PZ<double>>() : f.template run<PZ<float>>();
  return f64 ? f.template run<LG<double>>() : f.template run<LG<float>>();
}

template <class F>
struct OccupancyOf {
  F& f;
  template <class M>
  int run() {
    return f.template occupancy<M>();
  }
};

template <class F>
static int occupancy_of(const bihip_model* m, F& f) {
  const bool f64 = m->precision == BIHIP_F64;
  if (m->model >= BIHIP_MODEL_GENERATED) {
    OccupancyOf<F> o{f};
    return with_generated(m, o);
  }
  if (m->model == BIHIP_MODEL_PZ)
    return f64 ? f.template occupancy<PZ<double>>() : f.template occupancy<PZ<float>>();
  return f64 ? f.template occupancy<LG<double>>() : f.template occupancy<LG<float>>();
}

bool batch_model_supported(const bihip_model* m) {
  const GenInfo* g = generated_info(m->model);
  return model_supported(m) && (m->model == BIHIP_MODEL_PZ || m->model == BIHIP_MODEL_LG ||
                                (g != nullptr && g->lpp == 1));
}

int batch_nmax() { return BATCH_NMAX; }

template <int BLOCK>
static int launch_block(const BatchLaunch& L, hipStream_t s, int* grid_out, int items,
                        int oversub) {
  BatchRunner<BLOCK> r;
  r.items = items;
  r.a = BatchArgs{to_args(L.model), L.theta, L.ntheta, L.order, L.B, L.f, L.Xin, L.Xout, L.ld,
                  L.lws, L.as, L.anc_rec, L.hist_li, L.hist_ess, L.hist_res, L.y, L.nobs,
                  L.T, L.t0, L.dt, L.seed, L.essRel, L.resampler, L.counter};
  r.integ = L.model->integrator;
  r.s = s;
  int cus = 0, dev = 0;
  BIHIP_HIP(hipGetDevice(&dev));
  BIHIP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  // Workgroups beyond what is resident wait in the dispatcher and then take
  // whatever the queue still holds (or leave at once); no workgroup waits on
  // another, so oversubscription is safe and lets the hardware fill each CU.
  const int per_cu = std::max(8, occupancy_of(L.model, r)) * std::max(1, oversub);
  r.grid = std::max(1, std::min(L.B, cus * per_cu));
  if (grid_out) *grid_out = r.grid;
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(1), 0, s, L.counter);
  BIHIP_LAUNCH_CHECK();
  return with_small_model(L.model, r);
}

int launch_batch_step(const BatchLaunch& L, int block, hipStream_t s, int* grid_out, int items,
                      int oversub) {
  switch (block) {
    case 256: return launch_block<256>(L, s, grid_out, items, oversub);
    case 512: return launch_block<512>(L, s, grid_out, items, oversub);
    case 1024: return launch_block<1024>(L, s, grid_out, items, oversub);
  }
  return set_error(BIHIP_EINVAL, "batch block size must be 256, 512 or 1024");
}

// The tile widths compiled for a model (the autotuner's candidates).
int batch_items_for(const bihip_model* m, int* items) {
  if (m->model == BIHIP_MODEL_PZ && m->precision == BIHIP_F64 && m->integrator == BIHIP_RK4) {
    items[0] = 2;
    items[1] = 4;
    items[2] = 8;
    return 3;
  }
  items[0] = BITEMS_DEFAULT;
  return 1;
}

struct BatchInitRunner {
  ModelArgs ma;
  const double* theta;
  int ntheta;
  const long* off;
  int B;
  long total;
  uint64_t seed;
  uint32_t stream_base;
  void* X;
  long ld;
  hipStream_t s;
  template <class M>
  int run() {
    if constexpr (M::LPP != 1) {
      return set_error(BIHIP_EINVAL, "batched filters take models of one lane per particle");
    } else {
      hipLaunchKernelGGL((k_batch_init<M>), dim3(ceil_div(total, 256)), dim3(256), 0, s, ma,
                         theta, ntheta, off, B, total, seed, stream_base, X, ld);
      BIHIP_LAUNCH_CHECK();
      return BIHIP_OK;
    }
  }
  template <class M>
  int occupancy() {
    return 1;
  }
};

int launch_batch_init(const bihip_model* m, const double* theta, int ntheta, const long* off,
                      int B, long total, uint64_t seed, uint32_t stream_base, void* X, long ld,
                      hipStream_t s) {
  if (total <= 0) return BIHIP_OK;
  BatchInitRunner r{to_args(m), theta, ntheta, off, B, total, seed, stream_base, X, ld, s};
  return with_small_model(m, r);
}

}  // namespace bihip
