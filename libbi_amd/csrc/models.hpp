// Device-side models and ODE integrators.
//
// One particle is owned by a group of LPP consecutive lanes of a wavefront;
// lane g of the group holds state dimensions [g*DL, (g+1)*DL) in registers
// (DL = D / LPP).  The Lorenz-96 stencil fetches its three halo values from
// the neighbouring lanes of the group; error norms are group butterfly sums,
// so every lane of a group computes bit-identical step-control decisions and
// the group never diverges internally.
//
// Semantics follow the reference host integrators:
//   RK4    share/src/bi/host/ode/RK4IntegratorHost.hpp:63-97,  share/src/bi/ode/RK4Stage.hpp:26-76
//   DOPRI5 share/src/bi/host/ode/DOPRI5IntegratorHost.hpp:65-139, share/src/bi/ode/DOPRI5Stage.hpp:26-141,
//          controller constants share/src/bi/host/ode/IntegratorConstants.cpp:77-87.
// Step control (t, h, the controller) runs in double; stage arithmetic runs
// in the state type T.
#pragma once

#include "common.hpp"

namespace bihip {

constexpr double HALF_LOG_TWO_PI = 0.91893853320467274178;

// Arguments shared by the model kernels (uniform across the launch).
struct ModelArgs {
  int model, d, integrator, nsteps, host_quirk, precision;
  double h0, atol, rtol;
  double par[16];
  // Gaussian observation constants 1/obs_sd and -log(2 pi)/2 - log(obs_sd),
  // computed once on the host when obs_pre != 0 (kernels read them from their
  // arguments at the point of use, so they occupy no vector registers)
  int obs_pre;
  double obs_inv, obs_c;
};

// ---------------------------------------------------------------- lane groups
// Cross-lane moves inside a group.  Groups of four use DPP quad permutations
// (a VALU operand modifier, no LDS traffic); other sizes use ds_bpermute.
// A quad permutation reads only lanes of the same group, which are active
// together, so bound_ctrl (zero for an invalid source lane) never applies; it
// is set because it lets the compiler fold the move into the consuming
// instruction's DPP operand.
constexpr int QP_PREV = 3 | (0 << 2) | (1 << 4) | (2 << 6);  // lane g reads g-1 (mod 4)
constexpr int QP_NEXT = 1 | (2 << 2) | (3 << 4) | (0 << 6);  // lane g reads g+1 (mod 4)
constexpr int QP_XOR1 = 1 | (0 << 2) | (3 << 4) | (2 << 6);
constexpr int QP_XOR2 = 2 | (3 << 2) | (0 << 4) | (1 << 6);

template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <int LPP>
struct Group {
  int lane;  // lane in the wavefront
  int g;     // lane in the group
  __device__ __forceinline__ Group() {
    lane = (int)(threadIdx.x & (WAVE - 1));
    g = lane & (LPP - 1);
  }
  template <class T>
  __device__ __forceinline__ T prev(T v) const {
    if constexpr (LPP == 1) {
      return v;
    } else if constexpr (LPP == 2) {
      return dpp<QP_XOR1>(v);
    } else if constexpr (LPP == 4) {
      return dpp<QP_PREV>(v);
    } else {
      const int src = (lane & ~(LPP - 1)) | ((g + LPP - 1) & (LPP - 1));
      return __shfl(v, src, WAVE);
    }
  }
  template <class T>
  __device__ __forceinline__ T next(T v) const {
    if constexpr (LPP == 1) {
      return v;
    } else if constexpr (LPP == 2) {
      return dpp<QP_XOR1>(v);
    } else if constexpr (LPP == 4) {
      return dpp<QP_NEXT>(v);
    } else {
      const int src = (lane & ~(LPP - 1)) | ((g + 1) & (LPP - 1));
      return __shfl(v, src, WAVE);
    }
  }
  // butterfly sum: every lane of the group gets the same bits
  template <class T>
  __device__ __forceinline__ T sum(T v) const {
    if constexpr (LPP == 2) {
      return v + dpp<QP_XOR1>(v);
    } else if constexpr (LPP == 4) {
      v += dpp<QP_XOR1>(v);
      v += dpp<QP_XOR2>(v);
      return v;
    } else {
#pragma unroll
      for (int o = 1; o < LPP; o <<= 1) v += __shfl_xor(v, o, WAVE);
      return v;
    }
  }
};

// ---------------------------------------------------------------- models
// PZ: docs/src/users.tex:36-85.  State (P, Z); alpha ~ N(mu, sigma) per interval.
template <class T>
struct ModelPZ {
  static constexpr int D = 2, LPP = 1, DL = 2, NOBS = 1;
  static constexpr bool NOISE_FIRST = true;
  static constexpr bool HAS_FORCING = false;
  T c, e, ml, mq, alpha;
  double mu, sigma, obs_sd;
  __device__ __forceinline__ explicit ModelPZ(const ModelArgs& a) {
    mu = a.par[0];
    sigma = a.par[1];
    c = (T)a.par[2];
    e = (T)a.par[3];
    ml = (T)a.par[4];
    mq = (T)a.par[5];
    obs_sd = a.par[6];
    alpha = T(0);
  }
  template <class G>
  __device__ __forceinline__ void rhs(const G&, const T (&x)[DL], T (&k)[DL]) const {
    const T P = x[0], Z = x[1];
    k[0] = alpha * P - c * P * Z;
    k[1] = e * c * P * Z - ml * Z - mq * Z * Z;
  }
};

// Lorenz-96: dx_i/dt = (x_{i+1} - x_{i-2}) x_{i-1} - x_i + F, cyclic.
template <class T, int D_, int LPP_>
struct ModelL96 {
  static constexpr int D = D_, LPP = LPP_, DL = D_ / LPP_, NOBS = D_ / 2;
  static constexpr bool NOISE_FIRST = false;
  // the right-hand side is g(x) + F with a constant forcing F (rhs_unforced: g)
  static constexpr bool HAS_FORCING = true;
  static_assert(D_ % LPP_ == 0 && DL >= 2 && D_ % 2 == 0, "even dimension, lane slice >= 2");
  T F;
  double noise_sd, obs_sd;
  __device__ __forceinline__ explicit ModelL96(const ModelArgs& a) {
    F = (T)a.par[0];
    noise_sd = a.par[1];
    obs_sd = a.par[2];
  }
  template <bool FORCED, class G>
  __device__ __forceinline__ void eval(const G& grp, const T (&x)[DL], T (&k)[DL]) const {
    T xm2, xm1, xp1;  // x[lo-2], x[lo-1], x[hi]
    if constexpr (LPP == 1) {
      xm2 = x[DL - 2];
      xm1 = x[DL - 1];
      xp1 = x[0];
    } else {
      xm2 = grp.prev(x[DL - 2]);
      xm1 = grp.prev(x[DL - 1]);
      xp1 = grp.next(x[0]);
    }
#pragma unroll
    for (int i = 0; i < DL; ++i) {
      const T a = (i + 1 < DL) ? x[i + 1 < DL ? i + 1 : 0] : xp1;
      const T b = (i >= 2) ? x[i >= 2 ? i - 2 : 0] : (i == 1 ? xm1 : xm2);
      const T c = (i >= 1) ? x[i >= 1 ? i - 1 : 0] : xm1;
      if constexpr (FORCED) k[i] = (a - b) * c - x[i] + F;
      else k[i] = (a - b) * c - x[i];
    }
  }
  template <class G>
  __device__ __forceinline__ void rhs(const G& grp, const T (&x)[DL], T (&k)[DL]) const {
    eval<true>(grp, x, k);
  }
  template <class G>
  __device__ __forceinline__ void rhs_unforced(const G& grp, const T (&x)[DL], T (&k)[DL]) const {
    eval<false>(grp, x, k);
  }
  __device__ __forceinline__ T forcing() const { return F; }
};

// Scalar linear-Gaussian: dx/dt = -lambda x.
template <class T>
struct ModelLG {
  static constexpr int D = 1, LPP = 1, DL = 1, NOBS = 1;
  static constexpr bool NOISE_FIRST = false;
  static constexpr bool HAS_FORCING = false;
  T lambda;
  double noise_sd, obs_sd;
  __device__ __forceinline__ explicit ModelLG(const ModelArgs& a) {
    lambda = (T)a.par[0];
    noise_sd = a.par[1];
    obs_sd = a.par[2];
  }
  template <class G>
  __device__ __forceinline__ void rhs(const G&, const T (&x)[DL], T (&k)[DL]) const {
    k[0] = -lambda * x[0];
  }
};

// ---------------------------------------------------------------- integrators
// RK4IntegratorHost::update for one particle; returns the number of steps.
template <class M, class T, class G>
__device__ __forceinline__ int integrate_rk4(const M& m, const G& grp, T (&x0)[M::DL],
                                             double t1, double t2, double h0) {
  constexpr int DL = M::DL;
  double t = t1, h = h0;
  int n = 0;
  while (t < t2 && n < RK4_MAX_STEPS) {
    if (t + 1.01 * h - t2 > 0.0) {
      h = t2 - t;
      if (h <= 0.0) break;
    }
    const T hT = (T)h;
    T k[DL], xs[DL], x4[DL];
    m.rhs(grp, x0, k);
#pragma unroll
    for (int i = 0; i < DL; ++i) {
      xs[i] = hT * (T(0.5) * k[i]) + x0[i];
      x4[i] = T(1.0 / 6.0) * k[i];
    }
    m.rhs(grp, xs, k);
#pragma unroll
    for (int i = 0; i < DL; ++i) {
      x4[i] += T(1.0 / 3.0) * k[i];
      xs[i] = hT * (T(0.5) * k[i]) + x0[i];
    }
    m.rhs(grp, xs, k);
#pragma unroll
    for (int i = 0; i < DL; ++i) {
      x4[i] += T(1.0 / 3.0) * k[i];
      xs[i] = hT * k[i] + x0[i];
    }
    m.rhs(grp, xs, k);
#pragma unroll
    for (int i = 0; i < DL; ++i) {
      x4[i] += T(1.0 / 6.0) * k[i];
      x0[i] = hT * x4[i] + x0[i];
    }
    t += h;
    ++n;
  }
  return n;
}

// Scaled error of one component: err h / (atol + rtol max(|a|, |b|)).
// double: as written.  float: err / den with v_rcp_f32 (1 ulp; the norm only
// steers h), the factor h applied once to the whole norm by the caller, and
// the larger magnitude by one v_max_f32 with absolute-value source modifiers
// (a NaN state already makes err, hence the norm, NaN).
__device__ __forceinline__ float max_abs(float a, float b) {
  // one instruction with the operands' absolute values (the compiler's fmaxf
  // first canonicalizes operands it cannot prove quiet, two more instructions)
  float r;
  asm("v_max_f32_e64 %0, |%1|, |%2|" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

template <class T>
__device__ __forceinline__ T err_scale(T err, T h, T atol, T rtol, T a, T b) {
  if constexpr (sizeof(T) == 4) {
    const T den = atol + rtol * max_abs(a, b);
    return err * __builtin_amdgcn_rcpf(den);
  } else {
    const T den = atol + rtol * fmax(fabs(a), fabs(b));
    return err * h / den;
  }
}

// DOPRI5IntegratorHost::update for one particle; returns the number of steps.
// The step-size controller (IntegratorConstants.cpp:45-61, 77-87):
// accept iff e2 <= 1; rejected: h *= max(facl, safe e2^-expo); accepted:
// h *= clamp(safe e2^-expo errold^beta, facl, facr), errold = sqrt(max(e2, 1e-8)).
// Its logarithms run in double for double states and in float (native log2 /
// exp2) for float states.
// Stage derivative: the model's right-hand side, or without its constant
// forcing when the integrator carries the forcing itself (FOLD).
template <bool FOLD, class M, class G, class T, int DL>
__device__ __forceinline__ void stage_rhs(const M& m, const G& grp, const T (&x)[DL], T (&k)[DL]) {
  if constexpr (FOLD) m.rhs_unforced(grp, x, k);
  else m.rhs(grp, x, k);
}

// The controller's logarithms: natural in double; base 2 in float, where
// v_log_f32 and v_exp_f32 compute log2 and exp2 natively (every logarithm of
// the controller is then in base 2: the same factors, with the constants
// log2 0.9 and log2 1e-4).
__device__ __forceinline__ double ctl_log(double x) { return log(x); }
__device__ __forceinline__ double ctl_exp(double x) { return exp(x); }
__device__ __forceinline__ float ctl_log(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float ctl_exp(float x) { return __builtin_amdgcn_exp2f(x); }

template <bool QUIRK, class M, class T, class G>
__device__ __forceinline__ int dopri5_loop(const M& m, const G& grp, T (&x0)[M::DL],
                                           double t1, double t2, const ModelArgs& a) {
  constexpr int DL = M::DL;
  constexpr bool F32 = sizeof(T) == 4;
  const T facl = T(0.2), facr = T(10.0), beta = T(0.04);
  // log(0.9), in base 2 for float (see ctl_log)
  const T logsafe = F32 ? T(-0.15200309344504995) : T(-0.10536051565782630123);
  const T expo = T(0.5 * (0.2 - 0.04 * 0.75));
  const T atol = (T)a.atol, rtol = (T)a.rtol;
  const T invD = T(1.0 / (double)M::D);
  // FP32 states of a model with constant forcing F: the stage derivatives
  // leave F out (k' = k - F) and each stage adds c_s F once, where c_s, the
  // sum of the stage's coefficients, is its node (1/5, 3/10, 4/5, 8/9, 1;
  // 1 for the solution weights; 0 for the error weights).  Algebraically the
  // same step; it saves one add per component and stage.  FP64 states keep
  // the reference's arithmetic order.
  constexpr bool FOLD = M::HAS_FORCING && sizeof(T) == 4;
  T fc = T(0);
  if constexpr (FOLD) fc = m.forcing();
  const T c2F = T(0.2) * fc, c3F = T(0.3) * fc, c4F = T(0.8) * fc, c5F = T(8.0 / 9.0) * fc;

  T k1[DL];
  double t = t1, h = a.h0;
  T logfacold = F32 ? T(-13.287712379549449) : T(-9.2103403719761836);  // log(1e-4)
  bool k1in = false;
  int n = 0;
  while (t < t2 && n < a.nsteps) {
    if (t + 1.01 * h - t2 > 0.0) {
      h = t2 - t;
      if (h <= 0.0) break;
    }
    const T hT = (T)h;
    if (!k1in) stage_rhs<FOLD>(m, grp, x0, k1);
    k1in = true;
    T xs[DL], a3[DL], a4[DL], a5[DL], a6[DL], y6[DL], er[DL], k[DL];
    // stage 1
    if constexpr (FOLD) {
#pragma unroll
      for (int i = 0; i < DL; ++i) {
        xs[i] = hT * (T(0.2) * k1[i] + c2F) + x0[i];
        a3[i] = T(3.0 / 40.0) * k1[i] + c3F;
        a4[i] = T(44.0 / 45.0) * k1[i] + c4F;
        a5[i] = T(19372.0 / 6561.0) * k1[i] + c5F;
        a6[i] = T(9017.0 / 3168.0) * k1[i] + fc;
        y6[i] = T(35.0 / 384.0) * k1[i] + fc;
        er[i] = T(71.0 / 57600.0) * k1[i];
      }
    } else {
#pragma unroll
      for (int i = 0; i < DL; ++i) {
        xs[i] = hT * (T(0.2) * k1[i]) + x0[i];
        a3[i] = T(3.0 / 40.0) * k1[i];
        a4[i] = T(44.0 / 45.0) * k1[i];
        a5[i] = T(19372.0 / 6561.0) * k1[i];
        a6[i] = T(9017.0 / 3168.0) * k1[i];
        y6[i] = T(35.0 / 384.0) * k1[i];
        er[i] = T(71.0 / 57600.0) * k1[i];
      }
    }
    // stage 2
    stage_rhs<FOLD>(m, grp, xs, k);
#pragma unroll
    for (int i = 0; i < DL; ++i) {
      a3[i] += T(9.0 / 40.0) * k[i];
      a4[i] += T(-56.0 / 15.0) * k[i];
      a5[i] += T(-25360.0 / 2187.0) * k[i];
      a6[i] += T(-355.0 / 33.0) * k[i];
      xs[i] = hT * a3[i] + x0[i];
    }
    // stage 3
    stage_rhs<FOLD>(m, grp, xs, k);
#pragma unroll
    for (int i = 0; i < DL; ++i) {
      a4[i] += T(32.0 / 9.0) * k[i];
      a5[i] += T(64448.0 / 6561.0) * k[i];
      a6[i] += T(46732.0 / 5247.0) * k[i];
      y6[i] += T(500.0 / 1113.0) * k[i];
      er[i] += T(-71.0 / 16695.0) * k[i];
      xs[i] = hT * a4[i] + x0[i];
    }
    // stage 4
    stage_rhs<FOLD>(m, grp, xs, k);
#pragma unroll
    for (int i = 0; i < DL; ++i) {
      a5[i] += T(-212.0 / 729.0) * k[i];
      a6[i] += T(49.0 / 176.0) * k[i];
      y6[i] += T(125.0 / 192.0) * k[i];
      er[i] += T(71.0 / 1920.0) * k[i];
      xs[i] = hT * a5[i] + x0[i];
    }
    // stage 5
    stage_rhs<FOLD>(m, grp, xs, k);
#pragma unroll
    for (int i = 0; i < DL; ++i) {
      a6[i] += T(-5103.0 / 18656.0) * k[i];
      y6[i] += T(-2187.0 / 6784.0) * k[i];
      er[i] += T(-17253.0 / 339200.0) * k[i];
      xs[i] = hT * a6[i] + x0[i];
    }
    // stage 6: xs now holds x5
    stage_rhs<FOLD>(m, grp, xs, k);
#pragma unroll
    for (int i = 0; i < DL; ++i) {
      y6[i] += T(11.0 / 84.0) * k[i];
      er[i] += T(22.0 / 525.0) * k[i];
      y6[i] = hT * y6[i] + x0[i];
    }
    // stage 7 (FSAL): k7 at x6 into k; the reference host evaluates it at
    // x5, which is k6 = f(x5) already in k
    if constexpr (!QUIRK) stage_rhs<FOLD>(m, grp, y6, k);
    T e2l = T(0);
#pragma unroll
    for (int i = 0; i < DL; ++i) {
      er[i] += T(-1.0 / 40.0) * k[i];
      const T e = err_scale(er[i], hT, atol, rtol, x0[i], y6[i]);
      e2l += e * e;
    }
    T e2 = grp.sum(e2l) * invD;
    if constexpr (sizeof(T) == 4) e2 *= hT * hT;  // err_scale leaves h out for float
    if (e2 <= T(1)) {
      t += h;
#pragma unroll
      for (int i = 0; i < DL; ++i) {
        x0[i] = y6[i];
        k1[i] = k[i];
      }
    }
    if (t < t2) {
      const T logfac11 = expo * ctl_log(e2);
      if (e2 > T(1)) {
        h *= (double)fmax(facl, ctl_exp(logsafe - logfac11));
      } else {
        T fac = ctl_exp(beta * logfacold + logsafe - logfac11);
        fac = fmin(facr, fmax(facl, fac));
        h *= (double)fac;
        logfacold = T(0.5) * ctl_log(fmax(e2, T(1.0e-8)));
      }
    }
    ++n;
  }
  return n;
}

// The reference host's k7 convention is a launch-wide setting: the loop is
// compiled for each, so neither carries the other's register moves.
template <class M, class T, class G>
__device__ __forceinline__ int integrate_dopri5(const M& m, const G& grp, T (&x0)[M::DL],
                                                double t1, double t2, const ModelArgs& a) {
  if (a.host_quirk != 0) return dopri5_loop<true>(m, grp, x0, t1, t2, a);
  return dopri5_loop<false>(m, grp, x0, t1, t2, a);
}

// Normals for the lane's slice [lo, lo+DL) of a D-vector drawn with the
// layout "normal j = word pair (j%4)/2 of Philox block j/4".
template <class T, int DL>
__device__ __forceinline__ void slice_normals(uint64_t seed, uint32_t index, uint32_t step,
                                              uint32_t tag, uint32_t batch, int lo,
                                              T (&z)[DL]) {
  if constexpr (DL % 2 == 0) {
    // Slices of an even length start at even elements (lo = g DL), so the
    // slice is DL/2 whole Box-Muller pairs: global pairs q0 .. q0 + DL/2 - 1,
    // q0 = lo/2, pair q being word pair q%2 of block q/2.  Each pair is
    // transformed once, from words chosen by the parity of q0.
    constexpr int NP = DL / 2;
    constexpr int NB = NP / 2 + 1;  // blocks an odd q0 touches
    const int q0 = lo >> 1, b0 = q0 >> 1;
    const bool odd = (q0 & 1) != 0;
    uint32_t w[4 * NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      // an even q0 with an even number of pairs does not reach the last block
      if (NP % 2 == 1 || b < NB - 1 || odd) {
        const u32x4 r = philox_block(seed, index, step, tag, batch, (uint32_t)(b0 + b));
        w[4 * b] = r.x;
        w[4 * b + 1] = r.y;
        w[4 * b + 2] = r.z;
        w[4 * b + 3] = r.w;
      } else {
        w[4 * b] = w[4 * b + 1] = w[4 * b + 2] = w[4 * b + 3] = 0u;
      }
    }
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      const uint32_t r0 = odd ? w[2 * i + 2] : w[2 * i];
      const uint32_t r1 = odd ? w[2 * i + 3] : w[2 * i + 1];
      Normal<T>::pair(r0, r1, z[2 * i], z[2 * i + 1]);
    }
    return;
  }
  constexpr int NB = (DL + 3) / 4 + 1;
  T Z[4 * NB];
  const int b0 = lo >> 2, s = lo & 3;
  const bool need_last = (s + DL) > 4 * (NB - 1);
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    T q[4];
    if (b < NB - 1 || need_last) {
      normals4<T>(seed, index, step, tag, batch, (uint32_t)(b0 + b), q);
    } else {
      q[0] = q[1] = q[2] = q[3] = T(0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) Z[4 * b + i] = q[i];
  }
#pragma unroll
  for (int i = 0; i < DL; ++i) {
    const T v0 = Z[i], v1 = Z[i + 1], v2 = Z[i + 2], v3 = Z[i + 3];
    z[i] = s == 0 ? v0 : (s == 1 ? v1 : (s == 2 ? v2 : v3));
  }
}

}  // namespace bihip
