// Learned surrogate of the deterministic flow over one observation interval,
// fused with the rest of the bootstrap filter's predict + correct:
//   (ancestor gather) -> MLP flow x' = x + s_dx MLP(u(x), u(par)) on the
//   matrix cores -> additive Philox noise -> observation log-density ->
//   per-block weight partials.
// It stands in for the DOPRI5 flow of k_transition on the Lorenz-96 model
// (the deterministic part of Model::transitionSamples,
// share/tt/cpp/model.hpp.tt:566-573); the noise, observation density and
// weight bookkeeping are the same as the exact kernel's.  The network has no
// counterpart in the reference.
//
// Layout: a workgroup of 8 waves takes 256 particles, 32 per wave.  Each
// layer is computed transposed, out^T = W in^T, with the weights as the MFMA
// A operand (staged in LDS, one layer at a time) and the activations as the B
// operand, held in registers: the accumulator of an output tile lands in the
// lane that owns the particle, so it becomes the next layer's B operand
// without any data movement.  The accumulator's row order inside a 16-wide
// k block ([0-3, 8-11, 4-7, 12-15] across the two half-waves) is absorbed by
// storing the weights of layers l >= 1 with their input columns permuted the
// same way (mlp_capi.cpp, inference weights).
#include "blockreduce.hpp"
#include "internal.hpp"
#include "mlp.hpp"
#include "model_ops.hpp"

namespace bihip {

struct SurrArgs {
  const uint16_t* W[MLP_MAX_LAYERS];  // inference weights [pad_{l+1}][pad_l]
  const float* b[MLP_MAX_LAYERS];
  float x_shift, x_inv_scale, par_u, dx_scale;
  const float* Xin;
  float* Xout;
  int ld, P;
  // filter step (STEP = true)
  const int* as;
  double* lws;
  double* partials;
  const PFDevice* dev;
  uint64_t seed;
  uint32_t batch;
  double t0, dt;
  const double* y_base;
  int nobs;
  float noise_sd;
  double obs_sd;
  const int* imp_r;  // imposed resampling flags per step (nullable)
  BankStrides bk;    // STEP: filter blockIdx.y of a bank at these strides
  const float* par_u_bank;  // bank: the parameter input per filter (nullable)
};

namespace {

constexpr int SG_THREADS = 512, SG_TILE = 256;

template <int D, int H, int NL, bool STEP>
__global__ __launch_bounds__(SG_THREADS) void k_surrogate(SurrArgs a) {
  constexpr int K0 = (D + 1 + 63) / 64 * 64;  // padded input width (state + parameter)
  constexpr int DO = (D + 63) / 64 * 64;      // padded output width
  constexpr int LDSK = H + 8;                 // LDS row stride (bank spread)
  static_assert(NL >= 2 && H % 32 == 0 && K0 <= H && DO <= H, "surrogate shape");
  __shared__ __attribute__((aligned(16))) uint16_t Ws[H * LDSK];

  if constexpr (STEP) {
    const long f = blockIdx.y;
    a.Xin = bank_shift(a.Xin, f * a.bk.x * (long)sizeof(float));
    a.Xout = bank_shift(a.Xout, f * a.bk.x * (long)sizeof(float));
    a.as = bank_shift(a.as, f * a.bk.p * (long)sizeof(int));
    a.lws = bank_shift(a.lws, f * a.bk.p * (long)sizeof(double));
    a.partials = bank_shift(a.partials, f * (long)gridDim.x * 3 * (long)sizeof(double));
    a.dev = a.dev + f;
    a.batch += (uint32_t)f;
    if (a.par_u_bank != nullptr) a.par_u = a.par_u_bank[f];
  }
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, h = lane >> 5;
  const int p = blockIdx.x * SG_TILE + w * 32 + (lane & 31);
  const bool active = p < a.P;

  bool gather = false, lws_const = false;
  double lws_base = 0.0;
  uint32_t step = 0;
  const double* y = nullptr;
  if constexpr (STEP) {
    gather = a.imp_r != nullptr ? a.imp_r[a.dev->k] != 0
                                : a.dev->resample != 0 && a.dev->status == BIHIP_OK;
    lws_const = gather;
    lws_base = a.dev->logLikelihood;
    const int k = a.dev->k;
    step = (uint32_t)(k + 1);
    y = a.y_base + (size_t)k * a.nobs;
  }
  const int src = active ? (gather ? a.as[p] : p) : 0;

  // layer-0 operand: lane holds input features 16 c + 8 h + j of its particle
  uint32_t in0[K0 / 16][4], cur[H / 16][4], nxt[H / 16][4];
#pragma unroll
  for (int c = 0; c < K0 / 16; ++c) {
    float u[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int f = 16 * c + 8 * h + j;
      u[j] = 0.0f;
      if (f < D) {
        if (active) u[j] = (a.Xin[(size_t)f * a.ld + src] - a.x_shift) * a.x_inv_scale;
      } else if (f == D) {
        u[j] = a.par_u;
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) in0[c][j] = pack_bf2(u[2 * j], u[2 * j + 1]);
  }

  // hidden layers
  stage_rows<H, K0, LDSK, SG_THREADS>(a.W[0], 0, Ws);
  __syncthreads();
  gelu_tiles<K0 / 16, H / 16, LDSK, 0, H / 32>(Ws, a.b[0], in0, nxt, lane);
#pragma unroll
  for (int l = 1; l < NL - 1; ++l) {
    __syncthreads();  // every wave is done with the previous layer's weights
    stage_rows<H, H, LDSK, SG_THREADS>(a.W[l], 0, Ws);
    __syncthreads();
    if (l & 1) gelu_tiles<H / 16, H / 16, LDSK, 0, H / 32>(Ws, a.b[l], nxt, cur, lane);
    else gelu_tiles<H / 16, H / 16, LDSK, 0, H / 32>(Ws, a.b[l], cur, nxt, lane);
  }
  const uint32_t(&last)[H / 16][4] = ((NL - 2) & 1) ? cur : nxt;

  // output layer: y for features 32 t + acc_row(r, lane)
  __syncthreads();
  stage_rows<DO, H, LDSK, SG_THREADS>(a.W[NL - 1], 0, Ws);
  __syncthreads();
  float yv[DO / 32][16];
#pragma unroll
  for (int t = 0; t < DO / 32; ++t) {
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
#pragma unroll
    for (int c = 0; c < H / 16; ++c) {
      const u16x8 wa = *reinterpret_cast<const u16x8*>(
          &Ws[(32 * t + (lane & 31)) * LDSK + 16 * c + 8 * h]);
      const u32v4 bv = {last[c][0], last[c][1], last[c][2], last[c][3]};
      acc = mfma32(wa, bv, acc);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) yv[t][r] = acc[r] + a.b[NL - 1][32 * t + acc_row(r, lane)];
  }

  // x' = x + s_dx y (+ noise), observation density over the even dimensions
  double obs = 0.0;
  const double inv_obs = STEP ? 1.0 / a.obs_sd : 0.0;
  const double c_obs = STEP ? -HALF_LOG_TWO_PI - log(a.obs_sd) : 0.0;
#pragma unroll
  for (int t = 0; t < DO / 32; ++t) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int f0 = 32 * t + 8 * q + 4 * h;  // four consecutive features f0 .. f0 + 3
      if (f0 >= D) continue;
      float z[4] = {0.0f, 0.0f, 0.0f, 0.0f};
      if (STEP && active) normals4<float>(a.seed, (uint32_t)p, step, TAG_NOISE, a.batch,
                                          (uint32_t)(f0 >> 2), z);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int f = f0 + i;
        if (f >= D || !active) continue;
        float x = a.Xin[(size_t)f * a.ld + src] + a.dx_scale * yv[t][4 * q + i];
        if (STEP) x = x + a.noise_sd * z[i];
        a.Xout[(size_t)f * a.ld + p] = x;
        if (STEP && (f & 1) == 0) {
          const double e = (y[f >> 1] - (double)x) * inv_obs;
          obs += -0.5 * (e * e) + c_obs;
        }
      }
    }
  }
  if constexpr (STEP) {
    obs += __shfl_xor(obs, 32, WAVE);  // the particle's other half-wave
    double lw = __builtin_nan("");
    if (active && h == 0) {
      lw = (lws_const ? lws_base : a.lws[p]) + obs;
      a.lws[p] = lw;
    }
    block_weight_partial<SG_THREADS>(lw, a.partials + 3 * blockIdx.x);
  }
}

}  // namespace

int surrogate_blocks(int P) { return ceil_div(P, SG_TILE); }

int launch_surrogate(const SurrogateLaunch& L, hipStream_t s) {
  const MlpView& n = L.net;
  if (n.L != 4 || n.dims[0] != 41 || n.dims[1] != 256 || n.dims[2] != 256 || n.dims[3] != 256 ||
      n.dims[4] != 40)
    return set_error(BIHIP_EINVAL, "the fused surrogate is built for a 41-256-256-256-40 network");
  if (L.P <= 0) return BIHIP_OK;
  if (L.ld < L.P) return set_error(BIHIP_EINVAL, "ld < P");
  SurrArgs a{};
  for (int l = 0; l < n.L; ++l) {
    a.W[l] = L.Winf[l];
    a.b[l] = n.b[l];
  }
  a.x_shift = L.x_shift;
  a.x_inv_scale = 1.0f / L.x_scale;
  a.par_u = (float)((L.par - L.par_shift) / L.par_scale);
  a.dx_scale = L.dx_scale;
  a.Xin = static_cast<const float*>(L.Xin);
  a.Xout = static_cast<float*>(L.Xout);
  a.ld = L.ld;
  a.P = L.P;
  a.as = L.as;
  a.lws = L.lws;
  a.partials = L.partials;
  a.dev = L.dev;
  a.seed = L.seed;
  a.batch = L.batch;
  a.t0 = L.t0;
  a.dt = L.dt;
  a.y_base = L.y_base;
  a.nobs = L.nobs;
  a.noise_sd = L.noise_sd;
  a.obs_sd = L.obs_sd;
  a.imp_r = L.imp_r;
  a.bk = L.bk;
  a.par_u_bank = L.par_u_bank;
  const int banks = L.banks > 1 ? L.banks : 1;
  if (banks > 1 && !L.step)
    return set_error(BIHIP_EINVAL, "a bank runs the surrogate as a filter step");
  const dim3 grid(surrogate_blocks(L.P), banks), block(SG_THREADS);
  if (L.step) {
    if (L.dev == nullptr || L.lws == nullptr || L.partials == nullptr || L.y_base == nullptr)
      return set_error(BIHIP_EINVAL, "the surrogate filter step needs the filter's state");
    hipLaunchKernelGGL((k_surrogate<40, 256, 4, true>), grid, block, 0, s, a);
  } else {
    hipLaunchKernelGGL((k_surrogate<40, 256, 4, false>), grid, block, 0, s, a);
  }
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

}  // namespace bihip
