// Fused predict + correct kernel: (ancestor gather) -> noise -> ODE flow ->
// observation log-density -> per-block weight partials.  One particle per
// lane group; state in registers; each state column is read and written
// once, coalesced across particles (structure-of-arrays).
//
// Replaces, for the ON_DEVICE location:
//   Model::transitionSamples       share/tt/cpp/model.hpp.tt:566-573
//   StaticSamplerHost::samples     share/src/bi/host/updater/StaticSamplerHost.hpp:43-65
//   RK4/DOPRI5IntegratorHost       share/src/bi/host/ode/{RK4,DOPRI5}IntegratorHost.hpp
//   observationLogDensities        share/src/bi/host/updater/SparseStaticLogDensityHost.hpp:46-68
//   gather_rows (fused read)       share/src/bi/host/primitive/matrix_primitive.hpp:65-75
#include <cstring>

#include "model_ops.hpp"
#include "blockreduce.hpp"
#include "internal.hpp"
#include "gen_model.hpp"

namespace bihip {

struct TransArgs {
  ModelArgs m;
  const void* Xin;
  void* Xout;
  int ld, P;
  const int* as;
  double* lws;
  const double* y;
  double t1, t2;
  uint64_t seed;
  uint32_t step, batch;
  int* nsteps_out;
  double* partials;
  const PFDevice* dev;
  double t0, dt;
  const double* y_base;
  int nobs;
  int flow_only;        // 1: deterministic flow, no noise (integrator update)
  const double* aux;    // flow_only: per-particle noise variable (PZ alpha) or null
  const double* times;  // with dev: schedule times (nullable: t0 + k dt)
  const int* mask;      // with dev: observed flags (nullable: every step observed)
  const double* inputs; // with dev: 16 model parameters per step (nullable)
  const int* imp_r;     // with dev: imposed resampling flags per step (nullable)
  BankStrides bk;       // BANK: strides of the filters (filter f = blockIdx.y)
  const double* bank_par;  // BANK: BANK_ROW parameters per filter
  void* noise_hist;     // SCHED, with dev (nullable): the step's noise variables into
                        // slot k + 1 of a (T + 1) x noise_width x ld history
  int noise_width;
};

// ---------------------------------------------------------------- the kernel
// A value the whole wave shares, moved to a scalar register so that the
// compiler keeps what derives from it out of vector registers.
__device__ __forceinline__ double uniform(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)b);
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// SCHED: the filter step has a schedule or inputs (times, observed flags,
// per-step parameters); without them the kernel is compiled without that code,
// which costs the plain step about 7% when present but unused.
// Waves per SIMD the compiler must leave room for: a generated functor names
// its own (MIN_WAVES), the hand-written ones fit their budget unaided.
template <class M>
constexpr int trans_min_waves() {
  if constexpr (is_generated<M>::value) return M::MIN_WAVES;
  else return 1;
}

// ONE: a generated model's interval is one sub-interval of its delta that
// starts at a multiple of it (its transition runs without the sub-interval loop).
// BANK: filter blockIdx.y of a bank, with its own state, vectors, PFDevice,
// Philox batch id and parameter row (then the step is exactly a lone
// filter's with those parameters and that batch id).
template <class M, int INTEG, bool SCHED = false, bool ONE = false, bool BANK = false>
__global__ __launch_bounds__(TRANS_BLOCK) __attribute__((amdgpu_waves_per_eu(trans_min_waves<M>())))
void k_transition(TransArgs a) {
  using T = typename M::real_t;
  constexpr int LPP = M::LPP, DL = M::DL;
  if constexpr (BANK) {
    const long f = blockIdx.y;
    a.Xin = bank_shift(a.Xin, f * a.bk.x * (long)sizeof(T));
    a.Xout = bank_shift(a.Xout, f * a.bk.x * (long)sizeof(T));
    a.as = bank_shift(a.as, f * a.bk.p * (long)sizeof(int));
    a.lws = bank_shift(a.lws, f * a.bk.p * (long)sizeof(double));
    a.partials = bank_shift(a.partials, f * (long)gridDim.x * 3 * (long)sizeof(double));
    a.dev = a.dev + f;
    a.batch += (uint32_t)f;
    const double* row = a.bank_par + BANK_ROW * f;
#pragma unroll
    for (int i = 0; i < 16; ++i) a.m.par[i] = uniform(row[i]);
    a.m.obs_inv = uniform(row[16]);
    a.m.obs_c = uniform(row[17]);
  }
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int tid = blk * TRANS_BLOCK + threadIdx.x;
  const int p = tid / LPP;
  const Group<LPP> grp;
  const int lo = grp.g * DL;
  const bool active = p < a.P;

  bool gather = a.as != nullptr;
  bool lws_const = false;
  double lws_base = 0.0;
  double t1 = a.t1, t2 = a.t2;
  uint32_t step = a.step;
  const double* y = a.y;
  // the inputs in force over [t_k, t_{k+1}) (Forcer): the step's parameters
  const double* in = SCHED && a.dev != nullptr && a.inputs != nullptr
                         ? a.inputs + (size_t)16 * a.dev->k
                         : nullptr;
  if (a.dev != nullptr) {
    const int k = a.dev->k;
    gather = a.imp_r != nullptr ? a.imp_r[k] != 0
                                : a.dev->resample != 0 && a.dev->status == BIHIP_OK;
    lws_const = gather;
    lws_base = a.dev->logLikelihood;
    t1 = SCHED && a.times != nullptr ? a.times[k] : a.t0 + k * a.dt;
    t2 = SCHED && a.times != nullptr ? a.times[k + 1] : a.t0 + (k + 1) * a.dt;
    step = (uint32_t)(k + 1);
    // BootstrapPF::correct applies only at observed elements
    // (share/src/bi/filter/BootstrapPF.hpp:171-181); null y_base: flow only
    const bool observed = !SCHED || a.mask == nullptr || a.mask[k] != 0;
    y = a.y_base != nullptr && observed ? a.y_base + (size_t)k * a.nobs : nullptr;
  }

  double lw = __builtin_nan("");
  if (active) {
    const int src = gather ? a.as[p] : p;
    const T* Xin = static_cast<const T*>(a.Xin);
    T x[DL];
#pragma unroll
    for (int i = 0; i < DL; ++i) x[i] = Xin[(size_t)(lo + i) * a.ld + src];

    auto model = [&]() {
      if (!SCHED || in == nullptr) return M(a.m);
      ModelArgs ma = a.m;
#pragma unroll
      for (int i = 0; i < 16; ++i) ma.par[i] = uniform(in[i]);
      return M(ma);
    };
    M m = model();
    int n;
    T* zout = nullptr;  // the noise variables of this step (recorded histories only)
    if constexpr (SCHED) {
      if (a.noise_hist != nullptr && a.dev != nullptr && !a.flow_only)
        zout = static_cast<T*>(a.noise_hist) +
               (size_t)(a.dev->k + 1) * (size_t)a.noise_width * (size_t)a.ld;
    }
    if constexpr (is_generated<M>::value) {
      // the model's transition block: discrete actions and its ode, in order
      n = m.template transition<INTEG, ONE>(grp, x, t1, t2, a.m, a.seed, a.batch, (uint32_t)p,
                                            step);
      if constexpr (SCHED) {
        if (zout != nullptr) m.noise_out(grp, zout, a.ld, (uint32_t)p);
      }
    } else {
      if constexpr (M::NOISE_FIRST) {
        if (!a.flow_only) {
          noise_first(m, a.seed, a.batch, p, step);
          if (SCHED && zout != nullptr && grp.g == 0) zout[p] = m.alpha;  // PZ's alpha
        } else {
          m.alpha = a.aux != nullptr ? (T)a.aux[p] : (T)m.mu;
        }
      }
      if constexpr (INTEG == BIHIP_RK4) {
        n = integrate_rk4(m, grp, x, t1, t2, a.m.h0);
      } else {
        n = integrate_dopri5(m, grp, x, t1, t2, a.m);
      }
      if constexpr (!M::NOISE_FIRST) {
        if (!a.flow_only) noise_after(m, a.seed, a.batch, p, lo, step, x, zout, a.ld);
      }
    }

    T* Xout = static_cast<T*>(a.Xout);
#pragma unroll
    for (int i = 0; i < DL; ++i) Xout[(size_t)(lo + i) * a.ld + p] = x[i];
    if (a.nsteps_out != nullptr && grp.g == 0) a.nsteps_out[p] = n;

    if (y != nullptr) {
      // observation constants from the host, or from the step's inputs
      const double part =
          grp.sum(obs_of(m, x, y, lo, !SCHED || in == nullptr ? &a.m : nullptr));
      if (grp.g == 0) {
        const double base = lws_const ? lws_base : a.lws[p];
        lw = base + part;
        a.lws[p] = lw;
      }
    } else if (SCHED && lws_const && a.lws != nullptr && grp.g == 0) {
      a.lws[p] = lws_base;  // resampled, then no observation: lws = logLikelihood
    }
  }
  // only group leaders (lanes g == 0) hold a log-weight
  if (a.partials != nullptr)
    block_weight_partial<TRANS_BLOCK, LPP>(lw, a.partials + 3 * blk);
}

// ---------------------------------------------------------------- initial states
// bank_par (nullable): filter blockIdx.y of a bank, parameter row
// bank_par + BANK_ROW y, batch id batch + y, state at Xv + y * bx elements.
template <class M>
__global__ __launch_bounds__(TRANS_BLOCK) void k_init(ModelArgs ma, uint64_t seed,
                                                      uint32_t batch, int P, void* Xv, int ld,
                                                      const double* mean0,
                                                      const double* bank_par, long bx) {
  using T = typename M::real_t;
  constexpr int LPP = M::LPP, DL = M::DL;
  if (bank_par != nullptr) {
    const long f = blockIdx.y;
    const double* row = bank_par + BANK_ROW * f;
    for (int i = 0; i < 16; ++i) ma.par[i] = row[i];
    ma.obs_inv = row[16];
    ma.obs_c = row[17];
    batch += (uint32_t)f;
    Xv = bank_shift(Xv, f * bx * (long)sizeof(T));
  }
  const int tid = blockIdx.x * TRANS_BLOCK + threadIdx.x;
  const int p = tid / LPP;
  const Group<LPP> grp;
  const int lo = grp.g * DL;
  if (p >= P) return;
  T* X = static_cast<T*>(Xv);
  if constexpr (is_generated<M>::value) {
    double xd[DL];
    M(ma).init(xd, seed, batch, (uint32_t)p, lo);
#pragma unroll
    for (int i = 0; i < DL; ++i) X[(size_t)(lo + i) * ld + p] = (T)xd[i];
    return;
  }
  T z[DL];
  slice_normals<T, DL>(seed, (uint32_t)p, 0, TAG_INIT, batch, lo, z);
#pragma unroll
  for (int i = 0; i < DL; ++i) {
    double v;
    if (ma.model == BIHIP_MODEL_PZ) {
      v = i == 0 ? exp(ma.par[7] + ma.par[8] * (double)z[0])
                 : exp(ma.par[9] + ma.par[10] * (double)z[1]);
    } else if (ma.model == BIHIP_MODEL_L96) {
      v = mean0[lo + i] + ma.par[3] * (double)z[i];
    } else {
      v = ma.par[3] + ma.par[4] * (double)z[i];
    }
    X[(size_t)(lo + i) * ld + p] = (T)v;
  }
}

// ---------------------------------------------------------------- standalone correction
template <class M>
__global__ __launch_bounds__(TRANS_BLOCK) void k_obs_logdensity(ModelArgs ma, const double* y,
                                                                int P, const void* Xv, int ld,
                                                                double* lws) {
  using T = typename M::real_t;
  constexpr int LPP = M::LPP, DL = M::DL;
  const int tid = blockIdx.x * TRANS_BLOCK + threadIdx.x;
  const int p = tid / LPP;
  const Group<LPP> grp;
  const int lo = grp.g * DL;
  if (p >= P) return;
  const T* X = static_cast<const T*>(Xv);
  T x[DL];
#pragma unroll
  for (int i = 0; i < DL; ++i) x[i] = X[(size_t)(lo + i) * ld + p];
  M m(ma);
  const double part = grp.sum(obs_of(m, x, y, lo));
  if (grp.g == 0) lws[p] = lws[p] + part;
}

// Observation samples (Model::observationSamples), y is P x nobs.
template <class M>
__global__ __launch_bounds__(TRANS_BLOCK) void k_obs_sample(ModelArgs ma, uint64_t seed,
                                                            uint32_t batch, uint32_t step, int P,
                                                            const void* Xv, int ld, double* y) {
  using T = typename M::real_t;
  const int p = blockIdx.x * TRANS_BLOCK + threadIdx.x;
  if (p >= P) return;
  const T* X = static_cast<const T*>(Xv);
  constexpr int NOBS = M::NOBS;
  if constexpr (is_generated<M>::value) {  // the whole state of the particle
    T x[M::D];
#pragma unroll
    for (int i = 0; i < M::D; ++i) x[i] = X[(size_t)i * ld + p];
    double yo[NOBS > 0 ? NOBS : 1];
    M(ma).obs_sample(x, seed, batch, (uint32_t)p, step, yo);
    for (int o = 0; o < NOBS; ++o) y[(size_t)p * NOBS + o] = yo[o];
    return;
  }
  for (int b = 0; 4 * b < NOBS; ++b) {
    T z[4];
    normals4<T>(seed, (uint32_t)p, step, TAG_OBS, batch, (uint32_t)b, z);
    for (int i = 0; i < 4 && 4 * b + i < NOBS; ++i) {
      const int o = 4 * b + i;
      double v;
      if (ma.model == BIHIP_MODEL_PZ) {
        v = exp(log((double)X[p]) + ma.par[6] * (double)z[i]);
      } else if (ma.model == BIHIP_MODEL_L96) {
        v = (double)X[(size_t)(2 * o) * ld + p] + ma.par[2] * (double)z[i];
      } else {
        v = (double)X[p] + ma.par[2] * (double)z[i];
      }
      y[(size_t)p * NOBS + o] = v;
    }
  }
}

// ---------------------------------------------------------------- random-number utilities
__global__ void k_philox(const uint32_t* ctr, uint32_t k0, uint32_t k1, int n, uint32_t* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const u32x4 r = philox4x32_10(ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3], k0, k1);
  out[4 * i] = r.x;
  out[4 * i + 1] = r.y;
  out[4 * i + 2] = r.z;
  out[4 * i + 3] = r.w;
}

__global__ void k_uniforms(uint64_t seed, uint32_t step, uint32_t tag, uint32_t batch, int n,
                           double* u) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const u32x4 r = philox_block(seed, (uint32_t)j, step, tag, batch, 0);
  u[j] = u53(r.x, r.y);
}

template <class T>
__global__ void k_normals(uint64_t seed, uint32_t index, uint32_t step, uint32_t tag,
                          uint32_t batch, int n, T* z) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (4 * b >= n) return;
  T q[4];
  normals4<T>(seed, index, step, tag, batch, (uint32_t)b, q);
  for (int i = 0; i < 4 && 4 * b + i < n; ++i) z[4 * b + i] = q[i];
}

// ---------------------------------------------------------------- dispatch
const GenInfo* generated_info(int model) {
  const int i = model - BIHIP_MODEL_GENERATED;
  return i >= 0 && i < GEN_COUNT ? &GEN_MODELS[i] : nullptr;
}

bool model_supported(const bihip_model* m) {
  if (m->integrator != BIHIP_RK4 && m->integrator != BIHIP_DOPRI5) return false;
  if (m->precision != BIHIP_F32 && m->precision != BIHIP_F64) return false;
  if (const GenInfo* g = generated_info(m->model)) return m->d == g->d;
  switch (m->model) {
    case BIHIP_MODEL_PZ: return m->d == 2;
    case BIHIP_MODEL_L96: return m->d == 8 || m->d == 40 || m->d == 200;
    case BIHIP_MODEL_LG: return m->d == 1;
  }
  return false;
}

// Lanes per particle of the Lorenz-96 d=40 model.  Four: ten dimensions per
// lane, 112 VGPRs.  Two lanes (twenty dimensions each, ten Philox blocks per
// particle instead of twelve) need 216 VGPRs, two waves per SIMD, and
// measured 234 us per N = 2^20 FP32 transition against 207-210 us.
#ifndef L96_D40_LPP
#define L96_D40_LPP 4
#endif

// Calls F::template run<M>() for the model type selected by m.
template <class F>
static int with_model(const bihip_model* m, F& f) {
  const bool f64 = m->precision == BIHIP_F64;
  if (m->model >= BIHIP_MODEL_GENERATED) return with_generated(m, f);
  switch (m->model) {
    case BIHIP_MODEL_PZ:
      return f64 ? f.template run<PZ<double>>() : f.template run<PZ<float>>();
    case BIHIP_MODEL_L96:
      if (m->d == 40)
        return f64 ? f.template run<L96<double, 40, L96_D40_LPP>>()
                   : f.template run<L96<float, 40, L96_D40_LPP>>();
      if (m->d == 200)
        return f64 ? f.template run<L96<double, 200, 8>>() : f.template run<L96<float, 200, 8>>();
      return f64 ? f.template run<L96<double, 8, 2>>() : f.template run<L96<float, 8, 2>>();
    default:
      return f64 ? f.template run<LG<double>>() : f.template run<LG<float>>();
  }
}

int transition_blocks(const bihip_model* m, int P) {
  int lpp = 1;
  if (m->model == BIHIP_MODEL_L96) lpp = m->d == 200 ? 8 : (m->d == 40 ? L96_D40_LPP : 2);
  if (const GenInfo* g = generated_info(m->model)) lpp = g->lpp;
  return ceil_div((long)P * lpp, TRANS_BLOCK);
}

// [t1, t2] is one sub-interval of delta starting at a multiple of it, with a
// margin (1e-6 of delta) far inside the device's rounding window (1e-4,
// gen_scaled), so that the device's own rounding of t1 and t2 decides alike.
static bool one_interval(double t1, double t2, double delta) {
  const double s1 = t1 / delta, s2 = t2 / delta;
  const double r1 = std::round(s1), r2 = std::round(s2);
  return std::fabs(s1 - r1) < 1e-6 && std::fabs(s2 - r2) < 1e-6 && r2 == r1 + 1.0;
}

bool gen_one_delta(const bihip_model* m, double t0, double dt, int steps) {
  const GenInfo* g = generated_info(m->model);
  if (g == nullptr || steps <= 0) return false;
  for (int k = 0; k < steps; ++k)
    if (!one_interval(t0 + k * dt, t0 + (k + 1) * dt, g->delta)) return false;
  return true;
}

struct TransitionRunner {
  TransArgs a;
  int blocks;
  hipStream_t s;
  int integ;
  bool one;
  int banks;
  template <class M>
  int run() {
    const bool sched = a.dev != nullptr && (a.times != nullptr || a.mask != nullptr ||
                                            a.inputs != nullptr || a.noise_hist != nullptr);
    if (banks > 1 || a.bank_par != nullptr) {  // a bank, also of one filter: its rows
      // a generated model runs its general interval loop over delta (the
      // one-delta form gives the same values)
      if (sched || a.dev == nullptr || a.bank_par == nullptr || a.imp_r != nullptr)
        return set_error(BIHIP_EINVAL, "a filter bank step takes no schedule, inputs or "
                                       "imposed history");
      const dim3 grid(blocks, banks > 1 ? banks : 1);
      if (integ == BIHIP_RK4)
        hipLaunchKernelGGL((k_transition<M, BIHIP_RK4, false, false, true>), grid,
                           dim3(TRANS_BLOCK), 0, s, a);
      else
        hipLaunchKernelGGL((k_transition<M, BIHIP_DOPRI5, false, false, true>), grid,
                           dim3(TRANS_BLOCK), 0, s, a);
      BIHIP_LAUNCH_CHECK();
      return BIHIP_OK;
    }
    if constexpr (is_generated<M>::value) {
      if (one && !sched) {
        if (integ == BIHIP_RK4)
          hipLaunchKernelGGL((k_transition<M, BIHIP_RK4, false, true>), dim3(blocks),
                             dim3(TRANS_BLOCK), 0, s, a);
        else
          hipLaunchKernelGGL((k_transition<M, BIHIP_DOPRI5, false, true>), dim3(blocks),
                             dim3(TRANS_BLOCK), 0, s, a);
        BIHIP_LAUNCH_CHECK();
        return BIHIP_OK;
      }
    }
    if (integ == BIHIP_RK4) {
      if (sched)
        hipLaunchKernelGGL((k_transition<M, BIHIP_RK4, true>), dim3(blocks), dim3(TRANS_BLOCK), 0,
                           s, a);
      else
        hipLaunchKernelGGL((k_transition<M, BIHIP_RK4>), dim3(blocks), dim3(TRANS_BLOCK), 0, s, a);
    } else {
      if (sched)
        hipLaunchKernelGGL((k_transition<M, BIHIP_DOPRI5, true>), dim3(blocks), dim3(TRANS_BLOCK),
                           0, s, a);
      else
        hipLaunchKernelGGL((k_transition<M, BIHIP_DOPRI5>), dim3(blocks), dim3(TRANS_BLOCK), 0, s,
                           a);
    }
    BIHIP_LAUNCH_CHECK();
    return BIHIP_OK;
  }
};

int launch_transition(const TransLaunch& L, hipStream_t s) {
  if (L.P <= 0) return BIHIP_OK;
  TransitionRunner r;
  r.a.m = to_args(L.model);
  r.a.Xin = L.Xin;
  r.a.Xout = L.Xout;
  r.a.ld = L.ld;
  r.a.P = L.P;
  r.a.as = L.as;
  r.a.lws = L.lws;
  r.a.y = L.y;
  r.a.t1 = L.t1;
  r.a.t2 = L.t2;
  r.a.seed = L.seed;
  r.a.step = L.step;
  r.a.batch = L.batch;
  r.a.nsteps_out = L.nsteps_out;
  r.a.partials = L.partials;
  r.a.dev = L.dev;
  r.a.t0 = L.t0;
  r.a.dt = L.dt;
  r.a.times = L.times;
  r.a.mask = L.mask;
  r.a.inputs = L.inputs;
  r.a.y_base = L.y_base;
  r.a.nobs = L.nobs;
  r.a.flow_only = L.flow_only;
  r.a.aux = L.aux;
  r.a.imp_r = L.imp_r;
  r.a.bk = L.bk;
  r.a.bank_par = L.bank_par;
  r.a.noise_hist = L.noise_hist;
  r.a.noise_width = L.noise_width;
  r.banks = L.banks;
  r.blocks = transition_blocks(L.model, L.P);
  r.s = s;
  r.integ = L.model->integrator;
  if (const GenInfo* g = generated_info(L.model->model))
    r.one = L.dev != nullptr ? L.one_delta != 0 : one_interval(L.t1, L.t2, g->delta);
  else
    r.one = false;
  return with_model(L.model, r);
}

struct InitRunner {
  ModelArgs ma;
  uint64_t seed;
  uint32_t batch;
  int P, ld, blocks;
  void* X;
  const double* mean0;
  hipStream_t s;
  int banks;
  const double* bank_par;
  long bx;
  template <class M>
  int run() {
    hipLaunchKernelGGL((k_init<M>), dim3(blocks, banks), dim3(TRANS_BLOCK), 0, s, ma, seed, batch,
                       P, X, ld, mean0, bank_par, bx);
    BIHIP_LAUNCH_CHECK();
    return BIHIP_OK;
  }
};

int launch_init_samples(const bihip_model* m, uint64_t seed, uint32_t batch, int P, void* X,
                        int ld, const double* mean0_dev, hipStream_t s) {
  if (P <= 0) return BIHIP_OK;
  InitRunner r{to_args(m), seed, batch, P, ld, transition_blocks(m, P), X, mean0_dev, s,
               1, nullptr, 0};
  return with_model(m, r);
}

int launch_init_bank(const bihip_model* m, uint64_t seed, uint32_t batch, int P, void* X, int ld,
                     const double* mean0_dev, int banks, long x, const double* bank_par,
                     hipStream_t s) {
  if (P <= 0 || banks <= 0) return BIHIP_OK;
  InitRunner r{to_args(m), seed, batch, P, ld, transition_blocks(m, P), X, mean0_dev, s,
               banks, bank_par, x};
  return with_model(m, r);
}

int noise_width_of(const bihip_model* m) {
  if (const GenInfo* g = generated_info(m->model)) {  // the model file's noise variables
    int w = 0;
    for (int i = 0; g->names[BIHIP_VAR_NOISE][i] != nullptr; ++i) w += g->sizes[BIHIP_VAR_NOISE][i];
    return w;
  }
  switch (m->model) {
    case BIHIP_MODEL_PZ: return 1;    // alpha
    case BIHIP_MODEL_L96: return m->d;  // w[n]
    default: return 1;                // w
  }
}

void bank_row(const bihip_model* m, double* row) {
  const ModelArgs a = to_args(m);
  for (int i = 0; i < 16; ++i) row[i] = a.par[i];
  row[16] = a.obs_inv;
  row[17] = a.obs_c;
}

struct ObsRunner {
  ModelArgs ma;
  const double* y;
  int P, ld, blocks;
  const void* X;
  double* lws;
  hipStream_t s;
  template <class M>
  int run() {
    hipLaunchKernelGGL((k_obs_logdensity<M>), dim3(blocks), dim3(TRANS_BLOCK), 0, s, ma, y, P, X,
                       ld, lws);
    BIHIP_LAUNCH_CHECK();
    return BIHIP_OK;
  }
};

int launch_obs_logdensity(const bihip_model* m, const double* y, int P, const void* X, int ld,
                          double* lws, hipStream_t s) {
  if (P <= 0) return BIHIP_OK;
  ObsRunner r{to_args(m), y, P, ld, transition_blocks(m, P), X, lws, s};
  return with_model(m, r);
}

struct ObsSampleRunner {
  ModelArgs ma;
  uint64_t seed;
  uint32_t batch, step;
  int P, ld;
  const void* X;
  double* y;
  hipStream_t s;
  template <class M>
  int run() {
    hipLaunchKernelGGL((k_obs_sample<M>), dim3(ceil_div(P, TRANS_BLOCK)), dim3(TRANS_BLOCK), 0, s,
                       ma, seed, batch, step, P, X, ld, y);
    BIHIP_LAUNCH_CHECK();
    return BIHIP_OK;
  }
};

int launch_obs_sample(const bihip_model* m, uint64_t seed, uint32_t batch, uint32_t step, int P,
                      const void* X, int ld, double* y, hipStream_t s) {
  if (P <= 0) return BIHIP_OK;
  ObsSampleRunner r{to_args(m), seed, batch, step, P, ld, X, y, s};
  return with_model(m, r);
}

int launch_philox(const uint32_t* ctr, uint32_t k0, uint32_t k1, int n, uint32_t* out,
                  hipStream_t s) {
  if (n <= 0) return BIHIP_OK;
  hipLaunchKernelGGL(k_philox, dim3(ceil_div(n, 256)), dim3(256), 0, s, ctr, k0, k1, n, out);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

int launch_uniforms(uint64_t seed, uint32_t step, uint32_t tag, uint32_t batch, int n, double* u,
                    hipStream_t s) {
  if (n <= 0) return BIHIP_OK;
  hipLaunchKernelGGL(k_uniforms, dim3(ceil_div(n, 256)), dim3(256), 0, s, seed, step, tag, batch,
                     n, u);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

int launch_normals(uint64_t seed, uint32_t index, uint32_t step, uint32_t tag, uint32_t batch,
                   int n, int precision, void* z, hipStream_t s) {
  if (n <= 0) return BIHIP_OK;
  const int nb = ceil_div(n, 4);
  if (precision == BIHIP_F64) {
    hipLaunchKernelGGL(k_normals<double>, dim3(ceil_div(nb, 256)), dim3(256), 0, s, seed, index,
                       step, tag, batch, n, static_cast<double*>(z));
  } else {
    hipLaunchKernelGGL(k_normals<float>, dim3(ceil_div(nb, 256)), dim3(256), 0, s, seed, index,
                       step, tag, batch, n, static_cast<float*>(z));
  }
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

}  // namespace bihip

using namespace bihip;

int bihip_model_generated(const char* name, int precision, bihip_model* m) {
  if (name == nullptr || m == nullptr) return set_error(BIHIP_EINVAL, "NULL argument");
  if (precision != BIHIP_F32 && precision != BIHIP_F64)
    return set_error(BIHIP_EINVAL, "precision must be BIHIP_F32 or BIHIP_F64");
  for (int i = 0; i < GEN_COUNT; ++i) {
    const GenInfo& g = GEN_MODELS[i];
    if (std::strcmp(g.name, name) != 0) continue;
    *m = bihip_model{};
    m->model = BIHIP_MODEL_GENERATED + i;
    m->d = g.d;
    m->integrator = g.integrator;
    m->nsteps = 1000;
    m->host_quirk = 0;
    m->precision = precision;
    m->h0 = g.h;
    m->atol = g.atol;
    m->rtol = g.rtol;
    return BIHIP_OK;
  }
  return set_error(BIHIP_EINVAL, "no generated model named '%s'", name);
}

const char* bihip_model_var_name(const bihip_model* m, int kind, int i) {
  const GenInfo* g = m ? generated_info(m->model) : nullptr;
  if (g == nullptr || i < 0 || kind < 0 || kind > 4) return nullptr;
  const char* const* names = g->names[kind];
  for (int k = 0; k <= i; ++k)
    if (names[k] == nullptr) return nullptr;
  return names[i];
}

int bihip_model_var_size(const bihip_model* m, int kind, int i) {
  return bihip_model_var_name(m, kind, i) != nullptr ? generated_info(m->model)->sizes[kind][i] : 0;
}

const char* bihip_model_name(const bihip_model* m) {
  const GenInfo* g = m ? generated_info(m->model) : nullptr;
  return g ? g->name : nullptr;
}
