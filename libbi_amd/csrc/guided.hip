// Learned guided proposal for the Lorenz-96 d=200 filter, with the importance
// weight corrected exactly.
//
// The filter step runs the exact flow Phi(x) first (k_transition without
// noise or correction, gathered through the ancestors), then this kernel
// proposes, in place over Phi,
//   x' = Phi + q (m + e^{s/2} eps),   eps ~ N(0, I) (the transition-noise stream),
// where (m, s) = MLP(u) per dimension, u = ((Phi - shift) / scale,
// (y - H Phi) / obs_sd, (F - F0) / F_scale), and adds to the log-weight
//   log p(y | x') + log N(x'; Phi, q^2 I) - log N(x'; Phi + q m, q^2 e^s)
//   = log p(y | x') + sum_i [ -(m_i + e^{s_i/2} eps_i)^2 / 2 + eps_i^2 / 2 + s_i / 2 ].
// With m = 0 and s = 0 it is the bootstrap filter's step, draw for draw.
// The reference's nearest pattern is the auxiliary-weight correction of
// BridgePF / LookaheadPF (share/src/bi/filter/BridgePF.hpp:114-144,
// share/src/bi/filter/LookaheadPF.hpp:91-124); the learned proposal itself
// has no counterpart there.
//
// Network 301-256-256-256-400 (outputs: m for 200 dimensions, then s), BF16
// MFMA with FP32 accumulation, laid out as the surrogate kernel's (32
// particles per wave, weights staged in LDS in row chunks, activations in
// registers).  The inference weights of the output layer interleave the rows
// per 16 dimensions (m of dimensions 16 t .. 16 t + 15, then their s), so one
// 32-row output tile carries both numbers of the dimensions it covers.
#include "blockreduce.hpp"
#include "internal.hpp"
#include "mlp.hpp"
#include "model_ops.hpp"

namespace bihip {

struct GuidedArgs {
  const uint16_t* W[MLP_MAX_LAYERS];
  const float* b[MLP_MAX_LAYERS];
  float x_shift, x_inv_scale, innov_inv, par_u;
  float* X;          // in: Phi, out: x' (column-major, ld)
  int ld, P;
  const double* y;   // observations of this step (filter mode: from dev)
  uint64_t seed;
  uint32_t batch, step;
  float q;           // transition noise sd
  double obs_sd;
  // filter mode
  const PFDevice* dev;
  const double* y_base;
  int nobs;
  double* lws;
  double* partials;
  // standalone mode: log-weight increment per particle
  double* inc;
  const int* imp_r;  // filter mode: imposed resampling flags per step (nullable)
  BankStrides bk;    // filter mode: filter blockIdx.y of a bank at these strides
  const float* par_u_bank;  // bank: the parameter input per filter (nullable)
};

namespace {

constexpr int GP_THREADS = 512, GP_TILE = 256;
constexpr int GD = 200, GNOBS = 100, GH = 256;
constexpr int GK0 = 320;                 // 301 inputs padded
constexpr int GOUT_TILES = (GD + 15) / 16;  // 13 tiles of 16 dimensions (m and s)
constexpr int LDSK_H = GH + 8, LDSK_0 = GK0 + 8;

template <bool FILTER>
__global__ __launch_bounds__(GP_THREADS) void k_guided(GuidedArgs a) {
  // layer 0 is staged in two halves of 128 rows (168 KB would not fit)
  constexpr int LDS_ELEMS = GH * LDSK_H > 128 * LDSK_0 ? GH * LDSK_H : 128 * LDSK_0;
  __shared__ __attribute__((aligned(16))) uint16_t Ws[LDS_ELEMS];

  if constexpr (FILTER) {
    const long f = blockIdx.y;
    a.X = bank_shift(a.X, f * a.bk.x * (long)sizeof(float));
    a.lws = bank_shift(a.lws, f * a.bk.p * (long)sizeof(double));
    a.partials = bank_shift(a.partials, f * (long)gridDim.x * 3 * (long)sizeof(double));
    a.dev = a.dev + f;
    a.batch += (uint32_t)f;
    if (a.par_u_bank != nullptr) a.par_u = a.par_u_bank[f];
  }
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, h = lane >> 5;
  const int p = blockIdx.x * GP_TILE + w * 32 + (lane & 31);
  const bool active = p < a.P;
  const int pc = active ? p : 0;

  uint32_t step = a.step;
  const double* y = a.y;
  bool lws_const = false;
  double lws_base = 0.0;
  if constexpr (FILTER) {
    lws_const = a.imp_r != nullptr ? a.imp_r[a.dev->k] != 0
                                   : a.dev->resample != 0 && a.dev->status == BIHIP_OK;
    lws_base = a.dev->logLikelihood;
    step = (uint32_t)(a.dev->k + 1);
    y = a.y_base + (size_t)a.dev->k * a.nobs;
  }

  // layer-0 operand: inputs 16 c + 8 h + j
  uint32_t in0[GK0 / 16][4];
#pragma unroll
  for (int c = 0; c < GK0 / 16; ++c) {
    float u[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int f = 16 * c + 8 * h + j;
      float v = 0.0f;
      if (f < GD) {
        v = (a.X[(size_t)f * a.ld + pc] - a.x_shift) * a.x_inv_scale;
      } else if (f < GD + GNOBS) {
        const int o = f - GD;
        v = (float)(y[o] - (double)a.X[(size_t)(2 * o) * a.ld + pc]) * a.innov_inv;
      } else if (f == GD + GNOBS) {
        v = a.par_u;
      }
      u[j] = v;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) in0[c][j] = pack_bf2(u[2 * j], u[2 * j + 1]);
  }

  uint32_t cur[GH / 16][4], nxt[GH / 16][4];
  stage_rows<128, GK0, LDSK_0, GP_THREADS>(a.W[0], 0, Ws);
  __syncthreads();
  gelu_tiles<GK0 / 16, GH / 16, LDSK_0, 0, 4>(Ws, a.b[0], in0, nxt, lane);
  __syncthreads();
  stage_rows<128, GK0, LDSK_0, GP_THREADS>(a.W[0], 128, Ws);
  __syncthreads();
  gelu_tiles<GK0 / 16, GH / 16, LDSK_0, 4, 8>(Ws, a.b[0], in0, nxt, lane);
  __syncthreads();
  stage_rows<GH, GH, LDSK_H, GP_THREADS>(a.W[1], 0, Ws);
  __syncthreads();
  gelu_tiles<GH / 16, GH / 16, LDSK_H, 0, 8>(Ws, a.b[1], nxt, cur, lane);
  __syncthreads();
  stage_rows<GH, GH, LDSK_H, GP_THREADS>(a.W[2], 0, Ws);
  __syncthreads();
  gelu_tiles<GH / 16, GH / 16, LDSK_H, 0, 8>(Ws, a.b[2], cur, nxt, lane);

  // output layer in chunks of four tiles; each tile: 16 dimensions (m, s)
  const double inv_obs = 1.0 / a.obs_sd;
  const double c_obs = -HALF_LOG_TWO_PI - log(a.obs_sd);
  double acc_w = 0.0;
#pragma unroll
  for (int t0 = 0; t0 < GOUT_TILES; t0 += 4) {
    constexpr int CH = 4;
    const int nt = GOUT_TILES - t0 < CH ? GOUT_TILES - t0 : CH;  // tiles in this chunk
    __syncthreads();
    stage_rows_n<GH, LDSK_H, GP_THREADS>(a.W[3], 32 * t0, 32 * nt, Ws);
    __syncthreads();
#pragma unroll
    for (int dt = 0; dt < CH; ++dt) {
      const int t = t0 + dt;
      if (dt >= nt) break;
      const f32x16 o = linear_tile<GH / 16, LDSK_H>(Ws, t, t0, a.b[3], nxt, lane);
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int f0 = 16 * t + 8 * q + 4 * h;  // dimensions f0 .. f0 + 3
        if (f0 >= GD || !active) continue;
        float z[4];
        normals4<float>(a.seed, (uint32_t)p, step, TAG_NOISE, a.batch, (uint32_t)(f0 >> 2), z);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int f = f0 + i;
          const float m = o[4 * q + i], s = o[8 + 4 * q + i];
          const float d = m + __expf(0.5f * s) * z[i];  // (x' - Phi) / q
          float* xp = &a.X[(size_t)f * a.ld + p];
          const float x = *xp + a.q * d;
          *xp = x;
          acc_w += (double)(-0.5f * d * d + 0.5f * z[i] * z[i] + 0.5f * s);
          if ((f & 1) == 0) {
            const double e = (y[f >> 1] - (double)x) * inv_obs;
            acc_w += -0.5 * (e * e) + c_obs;
          }
        }
      }
    }
  }
  acc_w += __shfl_xor(acc_w, 32, WAVE);  // the particle's other half-wave
  if constexpr (FILTER) {
    double lw = __builtin_nan("");
    if (active && h == 0) {
      lw = (lws_const ? lws_base : a.lws[p]) + acc_w;
      a.lws[p] = lw;
    }
    block_weight_partial<GP_THREADS>(lw, a.partials + 3 * blockIdx.x);
  } else {
    if (active && h == 0) a.inc[p] = acc_w;
  }
}

}  // namespace

int guided_blocks(int P) { return ceil_div(P, GP_TILE); }

int launch_guided(const GuidedLaunch& L, hipStream_t s) {
  const MlpView& n = L.net;
  if (n.L != 4 || n.dims[0] != GD + GNOBS + 1 || n.dims[1] != GH || n.dims[2] != GH ||
      n.dims[3] != GH || n.dims[4] != 2 * GD)
    return set_error(BIHIP_EINVAL, "the guided proposal is built for a 301-256-256-256-400 network");
  if (L.P <= 0) return BIHIP_OK;
  if (L.ld < L.P) return set_error(BIHIP_EINVAL, "ld < P");
  GuidedArgs a{};
  for (int l = 0; l < n.L; ++l) {
    a.W[l] = L.Winf[l];
    a.b[l] = L.binf[l];
  }
  a.x_shift = L.x_shift;
  a.x_inv_scale = 1.0f / L.x_scale;
  a.innov_inv = (float)(1.0 / L.innov_scale);
  a.par_u = (float)((L.par - L.par_shift) / L.par_scale);
  a.X = static_cast<float*>(L.X);
  a.ld = L.ld;
  a.P = L.P;
  a.y = L.y;
  a.seed = L.seed;
  a.batch = L.batch;
  a.step = L.step;
  a.q = L.noise_sd;
  a.obs_sd = L.obs_sd;
  a.dev = L.dev;
  a.y_base = L.y_base;
  a.nobs = L.nobs;
  a.lws = L.lws;
  a.partials = L.partials;
  a.inc = L.inc;
  a.imp_r = L.imp_r;
  a.bk = L.bk;
  a.par_u_bank = L.par_u_bank;
  const int banks = L.banks > 1 ? L.banks : 1;
  if (banks > 1 && L.dev == nullptr)
    return set_error(BIHIP_EINVAL, "a bank runs the guided proposal as a filter step");
  const dim3 grid(guided_blocks(L.P), banks), block(GP_THREADS);
  if (L.dev != nullptr) {
    if (L.lws == nullptr || L.partials == nullptr || L.y_base == nullptr)
      return set_error(BIHIP_EINVAL, "the guided filter step needs the filter's state");
    hipLaunchKernelGGL((k_guided<true>), grid, block, 0, s, a);
  } else {
    if (L.y == nullptr || L.inc == nullptr)
      return set_error(BIHIP_EINVAL, "the guided proposal needs observations and an output");
    hipLaunchKernelGGL((k_guided<false>), grid, block, 0, s, a);
  }
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

}  // namespace bihip
