// Internal (host-side) interfaces between the kernel files and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "../../include/bihip.h"

namespace bihip {

// ---------------------------------------------------------------- errors
// Sets the calling thread's last-error string; returns the status code.
int set_error(int code, const char* fmt, ...);

// Host-side description of a model generated from a LibBi model file (the
// name lists end with nullptr); generated_info gives it for a model id, or
// nullptr for the hand-written models and unknown ids.
struct GenInfo {
  const char* name;
  int d, nobs, integrator, lpp;  // lpp: lanes per particle
  double h, atol, rtol, delta;
  const char* const* names[5];  // by BIHIP_VAR_* kind
  const int* sizes[5];          // elements of each variable
  const char* const* state_dims;  // dimension of each state variable ("" for a scalar)
  const char* const* noise_dims;  // dimension of each noise variable ("" for a scalar)
};
const GenInfo* generated_info(int model);
// Multiples of a generated model's delta one interval may span (each is a
// loop iteration of its transition).
constexpr int GEN_MAX_DELTAS = 1 << 16;
// EINVAL unless [t1, t2] is an interval the model's kernels take in bounded time.
int check_interval(const bihip_model* m, double t1, double t2);
int check_model(const bihip_model* m);
int nobs_of(const bihip_model* m);

#define BIHIP_HIP(call)                                                          \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess)                                                        \
      return ::bihip::set_error(BIHIP_EHIP, "%s failed: %s (%s:%d)", #call,      \
                                hipGetErrorString(e_), __FILE__, __LINE__);      \
  } while (0)

#define BIHIP_LAUNCH_CHECK()                                                     \
  do {                                                                           \
    hipError_t e_ = hipGetLastError();                                           \
    if (e_ != hipSuccess)                                                        \
      return ::bihip::set_error(BIHIP_EHIP, "kernel launch failed: %s (%s:%d)",  \
                                hipGetErrorString(e_), __FILE__, __LINE__);      \
  } while (0)

struct ModelArgs;

// Filter scalars kept on the device so that a step never waits for the host.
struct PFDevice {
  double logLikelihood;  // running log-likelihood (FilterState::logLikelihood)
  double ess;            // ESS after the last correct
  double mx;             // max log-weight after the last correct
  double lW;             // last incremental estimate
  double W;              // sum of weights found by the last resampling scan
  int resample;          // 1: resample at the current time (ess < essRel * P)
  int status;            // BIHIP_OK or BIHIP_DEGENERATE
  int k;                 // steps completed
  int pad;
};

// Scalars of one inner filter of a batch.
struct BatchFilter {
  double logLikelihood, ess, mx, lW;
  int resample, status, k, N;
  long off;         // first particle in the flattened state
  uint32_t stream;  // Philox batch id of the filter's draws
  int pad;
};

// Per-step histories on the device (T entries each; ancestors T x P when recorded).
struct PFHistory {
  double* logIncrements;
  double* ess;
  int* resampled;
};

// Scan bookkeeping: the single-pass scan's ticket, epoch and error flag, and
// the number of free slots of the tiled resampling chain's last permutation.
struct ScanState {
  uint32_t ticket;
  uint32_t epoch;
  uint32_t error;
  uint32_t free_slots;
};

struct ScanWorkspace {
  ScanState* state;  // device
  uint64_t* words;   // device: SCAN_WORDS words per tile (see SCAN_WORDS)
  int max_tiles;
};
// Words per tile: words [0, 4 max_tiles) hold the single-pass scan's
// epoch-tagged tile aggregates and [4 max_tiles, 8 max_tiles) its inclusive
// prefixes; [8 max_tiles, 12 max_tiles) the tiled resampling chain's plain
// tile sums, kept apart so that no tile sum can pass for a published value.
constexpr int SCAN_WORDS = 12;

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;
constexpr int RED_BLOCK = 256;
constexpr int RED_ITEMS = 8;
constexpr int RED_TILE = RED_BLOCK * RED_ITEMS;
constexpr int TRANS_BLOCK = 256;
constexpr int COMB_BLOCK = 256;
// RK4 steps per interval are bounded (t + h must advance t).
constexpr int RK4_MAX_STEPS = 1 << 20;

__host__ __device__ inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

// ---- transition.hip
// Filters of a bank advanced by shared launches, filter f = blockIdx.y:
// its state at f * x elements, its per-particle vectors (as, lws, Os, EF
// words) at f * p elements, its resampling tile words at f * words bytes,
// its scan state at f * state bytes, its PFDevice at dev + f, its partials
// after the f gridDim.x triples of the filters before it, and its Philox
// batch id batch + f.  A lone filter: one y-block and zero strides.
struct BankStrides {
  long x = 0, p = 0, words = 0, state = 0;
};

// ptr advanced by `bytes`, null kept null.
template <class T>
__host__ __device__ __forceinline__ T* bank_shift(T* ptr, long bytes) {
  return ptr != nullptr ? reinterpret_cast<T*>(reinterpret_cast<
                              typename std::conditional<std::is_const<T>::value, const char,
                                                        char>::type*>(ptr) + bytes)
                        : ptr;
}

// Parameter row of one filter of a bank: the model's 16 parameters, then the
// observation constants that the host computes from them (1 / obs_sd and
// -log(2 pi)/2 - log obs_sd), so that a bank filter computes exactly what a
// lone filter with those parameters computes.
constexpr int BANK_ROW = 18;

struct TransLaunch {
  const bihip_model* model;
  const void* Xin;
  void* Xout;
  int ld, P;
  const int* as;            // gather rows (nullable)
  double* lws;              // nullable
  const double* y;          // nullable: no correction
  double t1, t2;
  uint64_t seed;
  uint32_t step, batch;
  int* nsteps_out;          // nullable
  double* partials;         // nullable: 3 doubles per block
  const PFDevice* dev;      // nullable: device-side resample decision
  // with dev: the step k is read from dev->k; t1 = t0 + k dt, t2 = t0 + (k+1) dt,
  // Philox step k+1 and observations y_base + k * nobs
  double t0, dt;
  const double* y_base;
  int nobs;
  int flow_only;            // 1: no noise (RK4Integrator / DOPRI5Integrator::update)
  const double* aux;        // flow_only: per-particle noise variable, e.g. PZ alpha
  // with dev, a schedule (both nullable): times[k], times[k+1] in place of
  // t0 + k dt, and mask[k] == 0 for a step without observation (no correction)
  const double* times;
  const int* mask;
  // with dev: the model's 16 parameters for each step (nullable), LibBi's
  // inputs as set by a Forcer over [t_k, t_{k+1})
  const double* inputs;
  // with dev and no schedule: every step's [t1, t2] is one sub-interval of a
  // generated model's delta starting at a multiple of it (gen_one_delta);
  // without dev, launch_transition decides from t1 and t2 itself
  int one_delta;
  // with dev (nullable): imposed resampling flags, one per step; the step
  // gathers through `as` iff imp_r[k] != 0, in place of dev->resample
  const int* imp_r;
  // a bank of `banks` filters (0 or 1: a lone filter) with parameter rows
  // bank_par (BANK_ROW doubles per filter; with dev only)
  int banks;
  BankStrides bk;
  const double* bank_par;
  // with dev (nullable): a (T + 1) x noise_width x ld history of the noise
  // variables (in the model's precision) that the step fills at slot k + 1
  void* noise_hist;
  int noise_width;
};
// Noise variables per particle of a built-in model (PZ: alpha; Lorenz-96:
// w[d]; LG: w), 0 for a generated one.
int noise_width_of(const bihip_model* m);
int transition_blocks(const bihip_model* m, int P);
// A generated model whose steps [t0 + k dt, t0 + (k+1) dt], k < steps, are each
// one sub-interval of its delta starting at a multiple of it.
bool gen_one_delta(const bihip_model* m, double t0, double dt, int steps);
int launch_transition(const TransLaunch& L, hipStream_t s);
int launch_init_samples(const bihip_model* m, uint64_t seed, uint32_t batch, int P, void* X,
                        int ld, const double* mean0_dev, hipStream_t s);
// A bank's initial samples: filter f with parameter row bank_par + BANK_ROW f,
// batch id batch + f, state at X + f * x elements.
int launch_init_bank(const bihip_model* m, uint64_t seed, uint32_t batch, int P, void* X, int ld,
                     const double* mean0_dev, int banks, long x, const double* bank_par,
                     hipStream_t s);
// The parameter row of a model (its par and observation constants).
void bank_row(const bihip_model* m, double* row);
int launch_obs_logdensity(const bihip_model* m, const double* y, int P, const void* X, int ld,
                          double* lws, hipStream_t s);
int launch_obs_sample(const bihip_model* m, uint64_t seed, uint32_t batch, uint32_t step, int P,
                      const void* X, int ld, double* y, hipStream_t s);
int launch_philox(const uint32_t* ctr, uint32_t k0, uint32_t k1, int n, uint32_t* out,
                  hipStream_t s);
int launch_uniforms(uint64_t seed, uint32_t step, uint32_t tag, uint32_t batch, int n, double* u,
                    hipStream_t s);
int launch_normals(uint64_t seed, uint32_t index, uint32_t step, uint32_t tag, uint32_t batch,
                   int n, int precision, void* z, hipStream_t s);
bool model_supported(const bihip_model* m);
int launch_fill_f64(double* x, int n, double v, hipStream_t s);
int launch_iota(int* x, int n, hipStream_t s);

// ---- reduce.hip
// Level-2 partials and the last-block ticket of the two-level reduction.
struct ReduceWork {
  uint64_t* words;    // 6 epoch-tagged words (3 doubles) per combining block, then
                      // 3 plain doubles per block (the two-launch combination)
  ScanState* state;   // ticket (zero between launches), epoch, error
  int max_blocks;
};
int reduce_blocks(int nb);
// Per-tile (max, sum w, sum w^2) partials of log-weights.
int launch_lws_partials(const double* lws, int P, double* partials, hipStream_t s);
// Combine nb partials: out = {mx, S1, S2, ess, lW} (5 doubles).
int launch_finalize(const double* partials, int nb, int P, double* out, ReduceWork w,
                    hipStream_t s);
// Combine partials into the filter scalars and histories.
// The correct step of each filter of a bank: filter f's nb partials at
// partials + 3 nb f, its PFDevice at dev + f, its histories at f * T.
int launch_correct_bank(const double* partials, int nb, int P, double essRel, PFDevice* dev,
                        PFHistory hist, int T, int banks, hipStream_t s);
// With tile_sums (and correct_stores_sums(nb)), the last launch also stores
// the fixed-point weight sum of each tile of SCAN_TILE log-weights against the
// step's maximum, which the next step's resampling chain then reads instead of
// computing them (ResampleTilesLaunch::sums_ready): whenever the step leaves
// the resampling gate open, and under a mask at every step.
int launch_pf_correct_finalize(const double* partials, int nb, int P, double essRel,
                               PFDevice* dev, PFHistory hist, ReduceWork w, hipStream_t s,
                               const int* mask = nullptr, const double* lws = nullptr,
                               uint64_t* tile_sums = nullptr);
bool correct_stores_sums(int nb);

// ---- scan.hip
int launch_sumexp_scan(const double* lws, int P, const double* mx_dev, const PFDevice* gate,
                       double* Ws, ScanWorkspace ws, hipStream_t s);
int launch_count_scan(const int* os, int P, int* Os, ScanWorkspace ws, hipStream_t s);
// The same inclusive scan of nanexp(lws - *mx) by the tiled chain's kernels.
int launch_sumexp_scan_tiles(const double* lws, int P, const double* mx, double* Ws,
                             ScanWorkspace ws, hipStream_t s);
// Tiles that are certainly resident together (two per CU).
int scan_resident_tiles();
int scan_tiles(int P);

// ---- resample.hip
// mode: 0 systematic (a from uniform a_host or Philox when gate != nullptr),
//       1 stratified (alphas array or Philox), with Ws/W on device.
struct OffspringLaunch {
  int mode;                 // BIHIP_SYSTEMATIC / BIHIP_STRATIFIED
  const double* Ws;
  int P, n;
  double a;                 // systematic offset (when alphas == nullptr and !philox)
  const double* alphas;     // stratified alphas (nullable)
  bool philox;              // draw a / alphas from Philox(seed, step, tag, batch)
  uint64_t seed;
  uint32_t step, batch;
  const PFDevice* gate;     // nullable: skip unless gate->resample; step = gate->k
  PFDevice* status;         // nullable: where to flag W <= 0
  int* Os;
};
int launch_cumulative_offspring(const OffspringLaunch& L, hipStream_t s);
// Permutation: counts-or-cumulative input -> ancestors in the host layout.
int launch_ancestors_permute(const int* in, bool cumulative, int P, int* as, uint64_t* EF,
                             ScanWorkspace ws, const PFDevice* gate, hipStream_t s);
int launch_cumulative_to_ancestors(const int* Os, int P, int* as, hipStream_t s);
// The filter's whole resampling chain in tiles, from the log-weights to the
// permuted ancestors: systematic / stratified (cumulative offspring into Os)
// or multinomial (Ws, then the offspring counts into os).  Gated by gate.
struct McWork;
struct ResampleTilesLaunch {
  int mode;                 // BIHIP_SYSTEMATIC / BIHIP_STRATIFIED / BIHIP_MULTINOMIAL
  const double* lws;
  const double* mx;         // device: the maximum log-weight
  int P, n;
  uint64_t seed;
  uint32_t batch;
  const PFDevice* gate;     // the step number is gate->k
  PFDevice* status;         // where to flag W <= 0
  double* Ws;               // multinomial only
  int* Os;                  // systematic / stratified
  int* os;                  // multinomial
  int* as;
  uint64_t* EF;             // P words of scratch
  ScanWorkspace ws;
  int banks;                // 0 or 1: a lone filter; otherwise gate, status and mx of
  BankStrides bk;           // filter f at + f PFDevice
  const McWork* mc;         // multinomial (nullable): the chunked search's scratch
  int sums_ready;           // the tile weight sums are stored (a lone filter whose
                            // correct step stores them): no k_weight_sums
};
// Where the tiled chain keeps its tile weight sums (2 words per tile).
uint64_t* tile_sum_words(const ScanWorkspace& ws);
int launch_resample_tiles(const ResampleTilesLaunch& L, hipStream_t s);
struct MultinomialLaunch {
  const double* Ws;
  int P, n;
  const double* U;          // nullable: draw from Philox
  uint64_t seed;
  uint32_t step, batch;
  const PFDevice* gate;     // nullable: skip unless gate->resample; step = gate->k
  PFDevice* status;
  int* as;                  // nullable: ancestors out
  int* os;                  // nullable: offspring counts (atomic histogram; must be zeroed)
};
int launch_multinomial(const MultinomialLaunch& L, hipStream_t s);
// Scratch of the chunked multinomial: per-block chunk counts, chunk offsets,
// the segment -> chunk map and the draws sorted by chunk (targets and, for
// ancestors out, draw indices).
struct McWork {
  uint64_t* cstart = nullptr;
  double* st = nullptr;
  uint32_t* H = nullptr;
  int* segmap = nullptr;
  int* sj = nullptr;
  static size_t bytes(int P, int n, bool draws);
  void carve(void* base, int P, int n, bool draws);
};
bool multinomial_chunks_fit(int P);
// The same draws and results as launch_multinomial, sorted by chunk of the
// cumulative weights and searched in LDS (L.os: counts, zeroed here).
int launch_multinomial_chunks(const MultinomialLaunch& L, const McWork& w, hipStream_t s);
int launch_zero_ints(int* p, int n, const PFDevice* gate, hipStream_t s);
int launch_ancestors_to_offspring(const int* as, int n, int P, int* os, hipStream_t s);
int launch_gather_rows(const int* as, int P, int ncols, void* X, int ld, int elem,
                       hipStream_t s);
int launch_record_state(const void* X, int ld, int D, int P, int elem, const PFDevice* dev,
                        void* hist, int slot0, hipStream_t s);
int launch_paths(const int* anc, const void* hist, int ld, int D, int P, int K, int elem,
                 const int* idx, int n, double* out, hipStream_t s);
// Whole-filter copies: filter j of dst (bytes per filter) takes filter
// from[j] of src (from null: j) where mask is null or mask[j] != 0.
int launch_filter_copy(const void* src, void* dst, long bytes, int banks, const int* from,
                       const int* mask, hipStream_t s);
// banks > 1: filter f's ancestors at as + f bp, its history at out + f hs.
int launch_record_ancestors(const int* as, int P, const PFDevice* gate, int* out,
                            hipStream_t s, int banks = 1, long bp = 0, long hs = 0);

// ---- batch.hip
struct BatchLaunch {
  const bihip_model* model;
  const double* theta;
  int ntheta;
  const int* order;
  int B;
  BatchFilter* f;
  const void* Xin;
  void* Xout;
  long ld;
  double* lws;
  int* as;
  int* anc_rec;
  double* hist_li;
  double* hist_ess;
  int* hist_res;
  const double* y;
  int nobs, T;
  double t0, dt;
  uint64_t seed;
  double essRel;
  int resampler;
  uint32_t* counter;
};
bool batch_model_supported(const bihip_model* m);
int batch_nmax();
// One fused step of every filter: workgroups of `block` threads, `items`
// consecutive particles per thread in the resampling passes (a width that
// batch_items_for lists for the model), and `oversub` times the resident
// workgroups per CU in the persistent grid.
int launch_batch_step(const BatchLaunch& L, int block, hipStream_t s, int* grid_out, int items,
                      int oversub);
int batch_items_for(const bihip_model* m, int* items);  // returns how many (at most 3)
int launch_batch_init(const bihip_model* m, const double* theta, int ntheta, const long* off,
                      int B, long total, uint64_t seed, uint32_t stream_base, void* X, long ld,
                      hipStream_t s);
// Whole-filter copies between layouts (destination particle g of filter j
// takes source particle src_off[j] + g - dst_off[j]; mask[j] == 0 skips j).
int launch_batch_copy(const long* dst_off, const long* src_off, const int* len, const int* mask,
                      int B, long total, int d, int elem, const void* Xs, long lds,
                      const double* lws_s, void* Xd, long ldd, double* lws_d, hipStream_t s);

// ---- BF16 MFMA multilayer perceptron (mlp.hip)
enum { EPI_STORE = 0, EPI_BIAS_GELU = 1, EPI_BIAS = 2, EPI_GELU_GRAD = 3 };
struct GemmArgs {
  const uint16_t* A;  // A(m,k) = TA ? A[k lda + m] : A[m lda + k]  (bf16)
  long lda;
  const uint16_t* B;  // B(k,n) = TB ? B[n ldb + k] : B[k ldb + n]  (bf16)
  long ldb;
  int M, N, K;        // multiples of 64, 64, 32
  int ksplit;         // k range per split (multiple of 32); splits write Cf + z M ldc
  const float* bias;  // [N] (EPI_BIAS_GELU, EPI_BIAS)
  float* Cf;          // f32 out: product, pre-activation or linear output
  long ldc;
  uint16_t* Cb;       // bf16 out: GELU activation or back-propagated error
  long ldcb;
  const float* Zprev; // EPI_GELU_GRAD: pre-activation whose GELU derivative scales the product
  long ldz;
};
int launch_gemm(const GemmArgs& g, bool ta, bool tb, int epi, hipStream_t s);
int launch_rows_to_bf16(const float* X, long ldx, const long* idx, int n, int rows, int d,
                        uint16_t* Y, int ldy, hipStream_t s);
int launch_mse_grad(const float* Y, int ldy, const float* T, long ldt, const long* idx, int n,
                    int rows, int d, float scale, uint16_t* dY, float* partial, int* nblocks,
                    hipStream_t s);
int launch_gauss_grad(const float* Y, int ldy, const float* T, long ldt, const long* idx, int n,
                      int rows, int m, float scale, uint16_t* dY, float* partial, int* nblocks,
                      hipStream_t s);
int launch_colsum_bf16(const uint16_t* dZ, int ld, int rows, int cols, float* part, float* out,
                       hipStream_t s);
int launch_sum_splits(const float* part, int S, long count, float* out, hipStream_t s);
int launch_sum_all(const float* x, int n, float* out, hipStream_t s);
int launch_adam(float* w, float* m1, float* m2, const float* grad, long n, float lr, float b1,
                float b2, float eps, float bc1, float bc2, uint16_t* wbf, hipStream_t s);
int launch_to_bf16(const float* x, long n, uint16_t* y, hipStream_t s);

// Device parameters of an MLP (bf16 weights [pad_{l+1}][pad_l], FP32 biases),
// for the fused inference kernels.
constexpr int MLP_MAX_LAYERS = 6;
struct MlpView {
  int L;
  int dims[MLP_MAX_LAYERS + 1], pad[MLP_MAX_LAYERS + 1];
  const uint16_t* W[MLP_MAX_LAYERS];
  const float* b[MLP_MAX_LAYERS];
};
MlpView mlp_view(const bihip_mlp* m);
// bf16 weights with the input columns of layers l >= 1 permuted for the fused
// kernels (rebuilt from the FP32 master copy); per-layer pointers into W.
// out_interleave > 0: the output layer's rows are reordered in blocks of
// out_interleave outputs, [m of outputs 16 t .., then s of the same outputs],
// for networks whose outputs are a mean (first half) and a log variance;
// b (nullable) then gets biases in the same row order.
int mlp_inference_weights(bihip_mlp* m, const uint16_t* W[MLP_MAX_LAYERS],
                          const float* b[MLP_MAX_LAYERS] = nullptr, int out_interleave = 0);

// ---- learned surrogate flow (surrogate.hip)
struct SurrogateLaunch {
  MlpView net;
  const uint16_t* Winf[MLP_MAX_LAYERS];
  float x_shift, x_scale, dx_scale;  // u = (x - x_shift) / x_scale; x' = x + dx_scale y
  double par, par_shift, par_scale;  // u_d = (par - par_shift) / par_scale
  const void* Xin;
  void* Xout;
  int ld, P;
  bool step;                         // filter step: gather, noise, observation, partials
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
  const int* imp_r;                  // step: imposed resampling flags (nullable)
  int banks;                         // step: a bank of filters (0 or 1: a lone filter)
  BankStrides bk;
  const float* par_u_bank;           // bank: the normalised parameter input per filter
};
int surrogate_blocks(int P);
int launch_surrogate(const SurrogateLaunch& L, hipStream_t s);

// ---- learned guided proposal (guided.hip)
struct GuidedLaunch {
  MlpView net;
  const uint16_t* Winf[MLP_MAX_LAYERS];
  const float* binf[MLP_MAX_LAYERS];
  float x_shift, x_scale;
  double innov_scale, par, par_shift, par_scale;
  void* X;            // Phi in, x' out (FP32)
  int ld, P;
  const double* y;    // standalone
  uint64_t seed;
  uint32_t batch, step;
  float noise_sd;
  double obs_sd;
  const PFDevice* dev;  // filter mode when non-null
  const double* y_base;
  int nobs;
  double* lws;
  double* partials;
  double* inc;        // standalone: log-weight increments
  const int* imp_r;   // filter mode: imposed resampling flags (nullable)
  int banks;          // filter mode: a bank of filters (0 or 1: a lone filter)
  BankStrides bk;
  const float* par_u_bank;  // bank: the normalised parameter input per filter
};
int guided_blocks(int P);
int launch_guided(const GuidedLaunch& L, hipStream_t s);

// ---- host Philox4x32-10 block {index, step, tag | batch << 8, block}
void philox_host_block(uint64_t seed, uint32_t index, uint32_t step, uint32_t tag,
                       uint32_t batch, uint32_t block, uint32_t out[4]);

// ---- priors of a generated model's parameter block (prior.cpp), for the
// first n param elements: prior_check returns BIHIP_OK when every one has a
// prior with positive support (SMC^2 and PMMH move in log space);
// prior_sample draws them with the keyed stream {index, 0, THETA, 0};
// prior_logdensity is their log-density (-inf outside the support).
struct GenPrior {
  int dist;  // 0 gaussian, 1 log-gaussian, 2 uniform, 3 exponential, 4 gamma,
             // 5 inverse gamma, 6 beta; -1 none
  double a, b;
};
int prior_check(const bihip_model* m, int n, bool positive);
void prior_sample(const bihip_model* m, int n, uint64_t seed, uint32_t index, double* theta);
double prior_logdensity(const bihip_model* m, int n, const double* theta);

// ---- context access for the other host files (capi.cpp)
int ctx_enter_device(bihip_ctx* ctx);     // makes the context's device current
hipStream_t ctx_stream_of(bihip_ctx* ctx);
// ---- noalloc.hip: looks up the current device's copy of the static scratch
int noalloc_scratch_init(int device);

// ---- ancestry.hip: the pruned ancestry tree of a filter (record = 3)
struct AncestryCache {
  int P = 0, D = 0, elem = 4, cur = 0;
  long cap = 0;                       // node slots
  void* Xs = nullptr;                 // node states, column-major, leading dimension cap
  int* node_as = nullptr;             // parent slot (-1: first generation)
  int* node_os = nullptr;             // surviving children (0: free slot)
  int* ls[2] = {nullptr, nullptr};    // slot of each current particle (double-buffered)
  int* cnt = nullptr;                 // children of each leaf in the last step
  int* slots = nullptr;               // free slots taken by the new generation
  int* tile_free = nullptr;           // compaction: free slots per tile, then tile offsets
  long* nfree = nullptr;              // free slots found by the last write
  int* live = nullptr;                // live nodes (device)
  long live_ub = 0;                   // host: upper bound of *live since it was last read
};
int ancestry_alloc(AncestryCache* c, int P, int D, int elem, long cap, hipStream_t s);
void ancestry_free(AncestryCache* c);
int ancestry_enlarge(AncestryCache* c, long new_cap, hipStream_t s);
int ancestry_init(AncestryCache* c, const void* X, int ld, hipStream_t s);
int ancestry_write(AncestryCache* c, const void* X, int ld, const int* as, const PFDevice* dev,
                   const int* imp_r, hipStream_t s);
int ancestry_paths(const AncestryCache* c, int K, const int* idx, int n, double* out,
                   hipStream_t s);

// ---- device memory helpers
template <class T>
inline int dalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  BIHIP_HIP(hipMalloc((void**)p, n * sizeof(T)));
  return BIHIP_OK;
}

inline void dfree(void* p) {
  if (p) (void)hipFree(p);
}

}  // namespace bihip
