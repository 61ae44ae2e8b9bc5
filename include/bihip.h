/*
 * bihip.h -- C ABI of the MI355X (gfx950) particle-filter library.
 *
 * LibBi has no FFI: its operator API is a set of C++ template concepts
 * resolved at compile time by Location (ON_HOST / ON_DEVICE).  This header
 * is the boundary that the ON_DEVICE side of those concepts calls; the C++
 * shims in include/bi/amd/ (Filter, BootstrapPF, Resampler, integrators)
 * keep the reference's call sites.  Every entry point below names the
 * reference function it replaces (paths relative to the reference root).
 *
 * Conventions (as the reference):
 *  - state X is column-major, P rows (particles) x d columns, leading
 *    dimension ld >= P, so each variable is contiguous across particles
 *    (share/src/bi/state/State.hpp:469-479);
 *  - log-weights are double, ancestors are 0-based int;
 *  - lws carries the running log-likelihood: after a resample
 *    lws[:] = logLikelihood (share/src/bi/resampler/Resampler.hpp:177).
 *
 * Pointers named *_dev are device pointers; *_host are host pointers.
 * All work is ordered on the context's stream.  Calls that return host
 * scalars synchronise that stream.  Every function returns BIHIP_OK or an
 * error code; bihip_last_error() gives the calling thread's message.  No
 * C++ exception crosses this boundary: the C++ shims turn
 * BIHIP_DEGENERATE into bi::ParticleFilterDegeneratedException
 * (share/src/bi/misc/exception.hpp:47-49).
 */
#ifndef BIHIP_H
#define BIHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define BIHIP_OK 0
#define BIHIP_DEGENERATE 1 /* sum of weights <= 0 (ParticleFilterDegeneratedException) */
#define BIHIP_EINVAL 2     /* invalid argument */
#define BIHIP_EHIP 3       /* HIP runtime error */
#define BIHIP_ESTATE 4     /* object used out of order */
#define BIHIP_EIO 5        /* file input/output error */

/* ---- model, integrator, resampler and precision identifiers ---- */
#define BIHIP_MODEL_PZ 0  /* predator-prey, docs/src/users.tex:36-85 */
#define BIHIP_MODEL_L96 1 /* Lorenz-96 with additive noise */
#define BIHIP_MODEL_LG 2  /* scalar linear-Gaussian (Kalman-checkable) */
/* first id of the models generated from LibBi model files (libbi_amd/bi);
 * look them up by name with bihip_model_generated */
#define BIHIP_MODEL_GENERATED 64

#define BIHIP_RK4 0    /* share/src/bi/ode/RK4Integrator.hpp */
#define BIHIP_DOPRI5 1 /* share/src/bi/ode/DOPRI5Integrator.hpp */

#define BIHIP_SYSTEMATIC 0  /* share/src/bi/resampler/SystematicResampler.hpp */
#define BIHIP_STRATIFIED 1  /* share/src/bi/resampler/StratifiedResampler.hpp */
#define BIHIP_MULTINOMIAL 2 /* share/src/bi/resampler/MultinomialResampler.hpp */

#define BIHIP_F32 0 /* state and transition arithmetic in float */
#define BIHIP_F64 1 /* state and transition arithmetic in double */

/* random-stream tags of the Philox counter {index, step, tag|(batch<<8), block} */
#define BIHIP_TAG_INIT 1
#define BIHIP_TAG_NOISE 2
#define BIHIP_TAG_SYSTEMATIC 3
#define BIHIP_TAG_STRATIFIED 4
#define BIHIP_TAG_MULTINOMIAL 5
#define BIHIP_TAG_OBS 6
#define BIHIP_TAG_THETA 7     /* SMC^2 prior draws of theta */
#define BIHIP_TAG_PROPOSAL 8  /* SMC^2 random-walk proposals */
#define BIHIP_TAG_ACCEPT 9    /* SMC^2 acceptance uniforms */
#define BIHIP_TAG_OUTER 10    /* SMC^2 outer systematic offset */
#define BIHIP_TAG_PATH 13     /* path sampling (BootstrapPF::samplePath) */

/*
 * Model description: the runtime counterpart of a generated Model<name>
 * (share/tt/cpp/model.hpp.tt) plus the ode block's integrator settings
 * (share/tt/cpp/block/ode.hpp.tt:39-54).
 * par[] meaning per model:
 *  PZ : mu, sigma, c, e, m_l, m_q, obs_sd, P0_logmean, P0_logsd, Z0_logmean, Z0_logsd
 *  L96: F, noise_sd, obs_sd, init_sd      (d in {8, 40})
 *  LG : lambda, noise_sd, obs_sd, m0, s0  (d = 1)
 */
typedef struct {
  int model;      /* BIHIP_MODEL_* */
  int d;          /* state dimension */
  int integrator; /* BIHIP_RK4 / BIHIP_DOPRI5 */
  int nsteps;     /* DOPRI5 step limit per interval (IntegratorConstants.cpp:86) */
  int host_quirk; /* DOPRI5: 1 = k7 evaluated at x5 as the reference host does */
  int precision;  /* BIHIP_F32 / BIHIP_F64 */
  double h0, atol, rtol;
  double par[16];
} bihip_model;

/*
 * Models generated from LibBi model files (the reference's `libbi
 * --model-file`, lib/Bi/Parser.pm over share/bi.yp, code from
 * share/tt/cpp/model.hpp.tt and the block and action templates).  Every
 * file in libbi_amd/bi is compiled into the library when it is built.
 * bihip_model_generated fills *m for the model named `name` (its `model`
 * name in the file): id, state dimension, the ode block's algorithm ('RK4'
 * or 'RK5(4)'), h, atoler and rtoler; par[] holds the model's `param` then
 * `input` variables in declaration order, all zero.  The model's constants
 * are compiled in.
 */
#define BIHIP_VAR_PARAM 0
#define BIHIP_VAR_INPUT 1 /* par slots after the params; per step via bihip_pf_set_inputs */
#define BIHIP_VAR_STATE 2
#define BIHIP_VAR_OBS 3   /* order of the values of one observation */
#define BIHIP_VAR_NOISE 4
int bihip_model_generated(const char* name, int precision, bihip_model* m);
/* Name of variable i of a kind (BIHIP_VAR_*) of a generated model; NULL past
 * the end or for a hand-written model. */
const char* bihip_model_var_name(const bihip_model* m, int kind, int i);
/* Its number of elements (the size of its dimension, 1 for a scalar; 0 past
 * the end).  Vector variables take consecutive par slots, state rows and
 * observation values. */
int bihip_model_var_size(const bihip_model* m, int kind, int i);
/* The prior of the first n param elements of a generated model, from its
 * parameter block (Model::parameterSamples / parameterLogDensities,
 * share/tt/cpp/model.hpp.tt): a draw with the keyed stream {index, 0, THETA, 0}
 * (host), and the log-density of theta_host (-inf outside the support).
 * Gaussian, log-gaussian, uniform, exponential, gamma, inverse gamma and
 * beta priors with constant arguments. */
int bihip_model_prior_sample(const bihip_model* m, uint64_t seed, uint32_t index, int n,
                             double* theta_host);
int bihip_model_prior_logdensity(const bihip_model* m, int n, const double* theta_host,
                                 double* lp);
/* The model's name for a generated model, NULL otherwise. */
const char* bihip_model_name(const bihip_model* m);

typedef struct bihip_ctx bihip_ctx;
typedef struct bihip_pf bihip_pf;

/* ================================================================ context */
/* bi_init (share/src/bi/init.hpp:33-49): choose the device, create the context. */
int bihip_ctx_create(int device, bihip_ctx** ctx);
int bihip_ctx_destroy(bihip_ctx* ctx);
/* Order all work on this hipStream_t (NULL = the null stream). */
int bihip_ctx_set_stream(bihip_ctx* ctx, void* stream);
int bihip_ctx_synchronize(bihip_ctx* ctx);
const char* bihip_last_error(void);
/* Number of observed variables of a model (Mask size per observation). */
int bihip_model_nobs(const bihip_model* m);
/* Bytes per state element of a model (4 or 8). */
int bihip_model_elem_size(const bihip_model* m);

/* ================================================================ memory */
/* Device storage for the C++ shims (containers of State<B,ON_DEVICE>). */
int bihip_malloc(bihip_ctx* ctx, size_t bytes, void** ptr_dev);
int bihip_free(bihip_ctx* ctx, void* ptr_dev);
/* Synchronous copies host <-> device; asynchronous device -> device. */
int bihip_memcpy_to_device(bihip_ctx* ctx, void* dst_dev, const void* src_host, size_t bytes);
int bihip_memcpy_to_host(bihip_ctx* ctx, void* dst_host, const void* src_dev, size_t bytes);
int bihip_memcpy_device(bihip_ctx* ctx, void* dst_dev, const void* src_dev, size_t bytes);
/* set_elements(x, value) (share/src/bi/primitive/vector_primitive.hpp). */
int bihip_fill(bihip_ctx* ctx, double* x_dev, int n, double value);
/* seq_elements(x, 0): x[i] = i. */
int bihip_iota(bihip_ctx* ctx, int* x_dev, int n);

/* ================================================================ random numbers */
/* One uniform in [0,1) computed on the host from Philox counter
 * {index, step, tag|(batch<<8), 0} (the host draws of Random::uniform,
 * share/src/bi/random/Random.hpp, e.g. the systematic offset). */
double bihip_uniform_host(uint64_t seed, uint32_t index, uint32_t step, uint32_t tag,
                          uint32_t batch);
/* Philox4x32-10 blocks of n counters (ctr_dev: n x 4 words) under one key;
 * replaces the cuRAND XORWOW state of share/src/bi/cuda/random/RngGPU.cuh:19-206. */
int bihip_philox(bihip_ctx* ctx, const uint32_t* ctr_dev, uint32_t key0, uint32_t key1, int n,
                 uint32_t* out_dev);
/* u_j in [0,1), j < n, from counter {j, step, tag|(batch<<8), 0}
 * (Random::uniforms, share/src/bi/random/Random.hpp:44-380). */
int bihip_uniforms(bihip_ctx* ctx, uint64_t seed, uint32_t step, uint32_t tag, uint32_t batch,
                   int n, double* u_dev);
/* n standard normals for one index (normal j from block j/4), in the given
 * precision (float normals use 24-bit uniforms). */
int bihip_normals(bihip_ctx* ctx, uint64_t seed, uint32_t index, uint32_t step, uint32_t tag,
                  uint32_t batch, int n, int precision, void* z_dev);

/* ================================================================ model */
/* Initial-state sampling (Model::initialSamples, share/tt/cpp/model.hpp.tt).
 * mean0_host: d initial means (L96) or NULL. */
int bihip_init_samples(bihip_ctx* ctx, const bihip_model* m, uint64_t seed, uint32_t batch,
                       int P, void* X_dev, int ld, const double* mean0_host);
/*
 * Model::transitionSamples(rng, t1, t2, onDelta, s) (share/tt/cpp/model.hpp.tt:566-573):
 * noise pdf_ block then the ode block, one interval [t1,t2], for P particles.
 * Replaces RK4IntegratorHost/DOPRI5IntegratorHost::update
 * (share/src/bi/host/ode/RK4IntegratorHost.hpp:42-100,
 *  share/src/bi/host/ode/DOPRI5IntegratorHost.hpp:42-142) and the GPU kernels
 * share/src/bi/cuda/ode/{RK4,DOPRI5}KernelGPU.cuh.
 * Reads row as_dev[p] of Xin (as_dev NULL = row p), writes row p of Xout;
 * Xin == Xout is allowed when as_dev is NULL.  nsteps_dev (nullable) gets the
 * integrator step count per particle.
 */
int bihip_transition(bihip_ctx* ctx, const bihip_model* m, uint64_t seed, uint32_t batch,
                     uint32_t step, double t1, double t2, int P, const void* Xin_dev,
                     void* Xout_dev, int ld, const int* as_dev, int* nsteps_dev);
/* The ode block alone: RK4Integrator / DOPRI5Integrator::update(t1, t2, s)
 * (share/src/bi/ode/RK4Integrator.hpp:43-75, share/src/bi/ode/DOPRI5Integrator.hpp:44-76),
 * in place, without noise.  aux_dev: the per-particle noise variable the
 * right-hand side reads (PZ alpha), or NULL (PZ then uses alpha = mu). */
int bihip_flow(bihip_ctx* ctx, const bihip_model* m, double t1, double t2, int P, void* X_dev,
               int ld, const double* aux_dev, int* nsteps_dev);
/* Model::observationLogDensities(s, mask, lp) (share/tt/cpp/model.hpp.tt:757-765):
 * lws[p] += log p(y | x_p); y_dev holds nobs values. */
int bihip_obs_logdensity(bihip_ctx* ctx, const bihip_model* m, const double* y_dev, int P,
                         const void* X_dev, int ld, double* lws_dev);
/* Model::observationSamples for P particles: y_dev is P x nobs (row-major). */
int bihip_obs_sample(bihip_ctx* ctx, const bihip_model* m, uint64_t seed, uint32_t batch,
                     uint32_t step, int P, const void* X_dev, int ld, double* y_dev);

/* ================================================================ reductions */
/* ess_reduce (share/src/bi/primitive/vector_primitive.hpp:959-974):
 * returns ESS = (sum w)^2 / sum w^2 and lW = max + log sum w - log P. */
int bihip_ess_reduce(bihip_ctx* ctx, const double* lws_dev, int P, double* ess_host,
                     double* lW_host);
/* logsumexp_reduce (vector_primitive.hpp:932-943). */
int bihip_logsumexp_reduce(bihip_ctx* ctx, const double* lws_dev, int P, double* out_host);
/* max_reduce with NaN below every number (vector_primitive.hpp:860-872). */
int bihip_max_reduce(bihip_ctx* ctx, const double* x_dev, int n, double* out_host);
/* sumexpu_inclusive_scan (vector_primitive.hpp:1075-1084): Ws = inclusive
 * scan of nanexp(lws - max); single-pass decoupled look-back. mx_host nullable. */
int bihip_sumexpu_inclusive_scan(bihip_ctx* ctx, const double* lws_dev, int P, double* Ws_dev,
                                 double* mx_host);

/* ================================================================ resamplers */
/* SystematicResampler::cumulativeOffspring (SystematicResampler.hpp:110-130):
 * O_i = min(n, (int)(Ws_i / W * n + a)), W = Ws[P-1]. BIHIP_DEGENERATE if W <= 0. */
int bihip_systematic_cumulative_offspring(bihip_ctx* ctx, const double* Ws_dev, int P, int n,
                                          double a, int* Os_dev);
/* StratifiedResampler::cumulativeOffspring / StratifiedResamplerHost::op
 * (share/src/bi/host/resampler/StratifiedResamplerHost.hpp:25-74); alphas_dev has n values. */
int bihip_stratified_cumulative_offspring(bihip_ctx* ctx, const double* Ws_dev, int P, int n,
                                          const double* alphas_dev, int* Os_dev);
/* MultinomialResampler::ancestors (MultinomialResampler.hpp:82-92):
 * as_j = #{i : Ws_i <= W U_j} for n uniforms U_dev in [0,1). */
int bihip_multinomial_ancestors(bihip_ctx* ctx, const double* Ws_dev, int P, int n,
                                const double* U_dev, int* as_dev);
/* ancestorsToOffspring (share/src/bi/resampler/misc.hpp:100-108). */
int bihip_ancestors_to_offspring(bihip_ctx* ctx, const int* as_dev, int n, int P, int* os_dev);
/* inclusive scan of offspring counts into cumulative offspring. */
int bihip_offspring_to_cumulative(bihip_ctx* ctx, const int* os_dev, int P, int* Os_dev);
/* cumulativeOffspringToAncestors (misc.hpp:130-138; ResamplerHost.hpp:109-131). */
int bihip_cumulative_offspring_to_ancestors(bihip_ctx* ctx, const int* Os_dev, int P,
                                            int* as_dev);
/* cumulativeOffspringToAncestorsPermute (misc.hpp:140-149; host layout of
 * ResamplerHost.hpp:133-168): as[i]=i for particles with offspring, extra
 * copies fill the other slots in ascending order. Bit-exact with the host. */
int bihip_cumulative_offspring_to_ancestors_permute(bihip_ctx* ctx, const int* Os_dev, int P,
                                                    int* as_dev);
/* The same permutation without any allocation: no counterpart in the
 * reference.  It is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * Its scratch is a static device array of the library (about 8 MiB per GPU,
 * held until the process ends), so P may not exceed BIHIP_NOALLOC_MAX_P
 * (BIHIP_EINVAL beyond it) and a context must have been created on the
 * device.  Every call on one GPU shares that array: calls that overlap on
 * one GPU (other streams, threads or contexts) race and must be serialised
 * by the caller.  Results equal bihip_cumulative_offspring_to_ancestors_permute's. */
#define BIHIP_NOALLOC_MAX_P (1 << 20)
int bihip_cumulative_offspring_to_ancestors_permute_noalloc(bihip_ctx* ctx, const int* Os_dev,
                                                            int P, int* as_dev);
/* offspringToAncestorsPermute (misc.hpp:120-128; ResamplerHost.hpp:86-107). */
int bihip_offspring_to_ancestors_permute(bihip_ctx* ctx, const int* os_dev, int P,
                                         int* as_dev);
/* gather_rows in place, X[i,:] = X[as[i],:] (State::gather, State.hpp:905-909;
 * share/src/bi/cuda/primitive/matrix_primitive.cuh:67-78). Requires the
 * permuted layout (as[as[i]] == as[i]). elem_size is 4 or 8. */
int bihip_gather_rows(bihip_ctx* ctx, const int* as_dev, int P, int ncols, void* X_dev, int ld,
                      int elem_size);

/* ================================================================ filter */
/*
 * Device-resident bootstrap particle filter: Filter<BootstrapPF>::filter
 * (share/src/bi/filter/Filter.hpp:56-71, share/src/bi/filter/BootstrapPF.hpp:156-196)
 * with the ESS test, logLikelihood and logIncrements kept on the device, so a
 * step never waits for the host.  Resampling is triggered when the
 * observation at the current time has been assimilated and
 * ess < essRel * P (Resampler.hpp:168).  Schedule: t_k = t0 + k dt,
 * observed for k = 1..T.
 */
typedef struct {
  int P;             /* particles */
  int T;             /* observations */
  int resampler;     /* BIHIP_SYSTEMATIC / BIHIP_STRATIFIED / BIHIP_MULTINOMIAL */
  double essRel;     /* relative ESS threshold */
  double t0, dt;     /* schedule */
  uint64_t seed;     /* Philox key */
  uint32_t batch;    /* Philox batch id (theta index under SMC^2) */
  int record;        /* 1 = keep ancestors of every step on the device;
                        2 = also the state at every time (paths);
                        3 = the pruned ancestry tree only (paths, no full history) */
  int use_graph;     /* 1 = replay each step as a captured hipGraph */
} bihip_pf_config;

int bihip_pf_create(bihip_ctx* ctx, const bihip_model* m, const bihip_pf_config* cfg,
                    bihip_pf** pf);
int bihip_pf_destroy(bihip_pf* pf);
/* Observations, T x nobs row-major host array (Observer::update, Observer.hpp:115-126). */
int bihip_pf_set_observations(bihip_pf* pf, const double* y_host);
/* Filter::init + output0: sample initial particles, lws = 0, logLikelihood = 0. */
int bihip_pf_init(bihip_pf* pf, const double* mean0_host);
/* BootstrapPF::step: resample (device decision) + predict + correct for the
 * next n observation intervals. Asynchronous. */
int bihip_pf_step(bihip_pf* pf, int n);
/* BootstrapPF::term: logZ = logsumexp(lws) - log P; synchronises. */
int bihip_pf_term(bihip_pf* pf, double* logZ_host);
/* Status after the steps so far: BIHIP_OK or BIHIP_DEGENERATE; synchronises. */
int bihip_pf_status(bihip_pf* pf);
/* Copies of the filter's outputs (host pointers; any may be NULL):
 * logIncrements[T], ess[T], resampled[T], logLikelihood (running). */
int bihip_pf_outputs(bihip_pf* pf, double* logIncrements_host, double* ess_host,
                     int* resampled_host, double* logLikelihood_host);
/* Ancestors of step k (record=1), or of the last resample when k < 0. */
int bihip_pf_ancestors(bihip_pf* pf, int k, int* as_host);
/*
 * Paths (record = 2; AncestryCache::readPath, share/src/bi/cache/AncestryCache.hpp:328-352):
 * for each particle idx_host[q] at the current time K, its lineage's state at
 * times 0..K, paths_host[q][k][c] (n x (K + 1) x d doubles).  All states stay
 * on the device (no pruning: (T + 1) x d x P values, e.g. 84 GB for
 * d = 40, N = 2^20, T = 500 in FP32, within one GPU's HBM).
 */
int bihip_pf_paths(bihip_pf* pf, const int* idx_host, int n, double* paths_host);
/* record = 3: AncestryCache<ON_DEVICE> (share/src/bi/cache/AncestryCache.hpp:58-470,
 * prune kernel share/src/bi/cuda/cache/AncestryCacheKernel.cuh:29-52): after
 * every step the particles become a generation of nodes of a tree, leaves
 * without children are freed with every ancestor left without surviving
 * children, and only lineages that reach the current particles stay, so
 * bihip_pf_paths and bihip_pf_sample_path work on histories whose full record
 * would not fit in HBM.  The node store starts at 4 P nodes and doubles when a
 * step could overflow it (the only host waits of such a filter).  Not with
 * use_graph.  bihip_pf_cache_nodes: the live nodes and the node capacity
 * (AncestryCache::report, AncestryCache.hpp:438-444); synchronises. */
int bihip_pf_cache_nodes(bihip_pf* pf, long* live, long* capacity);
/* BootstrapPF::samplePath (share/src/bi/filter/BootstrapPF.hpp:146-154): one
 * particle drawn in proportion to its weight (keyed uniform {0, K, PATH,
 * batch}, the first with cumulative weight above u W), and its path
 * ((K + 1) x d); index_host (nullable) gets the particle. */
int bihip_pf_sample_path(bihip_pf* pf, double* path_host, int* index_host);
/* Impose another filter's resampling history: step k resamples iff
 * resampled_host[k] != 0, gathering through the P ancestors
 * ancestors_host[k P .. k P + P - 1] (checked to lie in [0, P)), in place of
 * the filter's own ESS test and draws; outputs report the imposed flags.
 * Both arrays are copied (T flags; T x P ancestors kept on the device).
 * NULL resampled_host removes the history.  No counterpart in the reference:
 * it runs two filters on one genealogy (an FP32 filter against the FP64
 * oracle's, a learned flow against the exact one).  Not with use_graph. */
int bihip_pf_impose_resampling(bihip_pf* pf, const int* resampled_host,
                               const int* ancestors_host);
/* Current state buffer (device pointer, leading dimension) and log-weights:
 * read-only views between steps (the next step resamples with the weight totals
 * its correct step computed). */
int bihip_pf_state(bihip_pf* pf, void** X_dev, int* ld, double** lws_dev);
/* Steps completed so far. */
int bihip_pf_position(bihip_pf* pf);
/*
 * Schedule (share/src/bi/state/Schedule.hpp:214-337, restricted to observation and
 * output times): step k runs from times[k] to times[k+1] (T + 1 increasing times, in
 * place of t0 + k dt), and observed[k] == 0 (nullable: every step observed) makes it a
 * prediction without correction, which BootstrapPF::correct
 * (share/src/bi/filter/BootstrapPF.hpp:171-181) and Resampler::resample
 * (share/src/bi/resampler/Resampler.hpp:165-182) skip: no weight update, no
 * log-likelihood increment, no resampling at the next step.  Observation row k is read
 * only when step k is observed.  Not combined with the learned flow or proposal.
 */
int bihip_pf_set_schedule(bihip_pf* pf, const double* times_host, const int* observed_host);
/*
 * Inputs (Forcer, share/src/bi/simulator/Forcer.hpp, for parameters the model reads as
 * inputs): inputs_host holds the model's 16 parameters for each step, T x 16; row k is
 * in force over [t_k, t_{k+1}), in the transition and the observation density (the
 * initial sampling keeps the model's parameters).  NULL removes them.  Not combined with
 * the learned flow or proposal.
 */
int bihip_pf_set_inputs(bihip_pf* pf, const double* inputs_host);
/*
 * Kernel timing with HIP events on the filter's stream: enable, then query
 * the summed milliseconds and launch count of one kernel class
 * (0 = transition, 1 = resample chain, 2 = weight reduction, 3 = the guided
 * proposal kernel alone, which class 0 also covers).  mode: 0 off,
 * BIHIP_TIMING_ALL every class, BIHIP_TIMING_TRANSITION classes 0 and 3 only
 * (each recorded event costs several microseconds of stream time, so the
 * narrower mode disturbs the timed steps less).
 */
#define BIHIP_TIMING_ALL 1
#define BIHIP_TIMING_TRANSITION 2
int bihip_pf_set_timing(bihip_pf* pf, int mode);
int bihip_pf_kernel_time(bihip_pf* pf, int kernel, double* ms_host, long* count_host);

/* ================================================================ NetCDF output */
/*
 * ParticleFilterNetCDFBuffer (share/src/bi/netcdf/ParticleFilterNetCDFBuffer.cpp:19-38,
 * with SimulatorNetCDFBuffer::create, share/src/bi/netcdf/SimulatorNetCDFBuffer.cpp:23-83):
 * LibBi's DEFAULT particle-filter schema -- dimensions nr, the model's (L96: n), np;
 * variables time(nr), the noise variables (a generated model's from its file; PZ: alpha;
 * L96: w(nr, n, np); LG: w) and the state variables (PZ: P, Z; L96: x(nr, n, np); LG: x),
 * ancestor(nr, np), logweight(nr, np), loglikelihood, clock; attributes libbi_schema =
 * "ParticleFilter", libbi_schema_version = 1 -- written without libnetcdf as a NetCDF
 * classic file with 64-bit offsets (CDF-2), or in the 64-bit data format (CDF-5) when a
 * variable before the last (the state variables) exceeds 4 GiB; nc_open reads both like
 * the reference's NetCDF-4 files.  clock is NC_DOUBLE microseconds.  All buffers are host
 * memory.
 */
typedef struct bihip_pfout bihip_pfout;
int bihip_pfout_create(const char* path, const bihip_model* m, int nr, int np,
                       bihip_pfout** out);
/* The same with the format chosen: 0 as bihip_pfout_create, 2 (CDF-2; an error when a
 * variable before the last exceeds 4 GiB) or 5 (CDF-5).  No counterpart in the reference,
 * which writes NetCDF-4 through libnetcdf. */
int bihip_pfout_create_format(const char* path, const bihip_model* m, int nr, int np,
                              int format, bihip_pfout** out);
/* SimulatorNetCDFBuffer::writeTime (SimulatorNetCDFBuffer.cpp:264) */
int bihip_pfout_write_time(bihip_pfout* o, int k, double t);
/* ParticleFilterNetCDFBuffer::writeState(k, X, as) (ParticleFilterNetCDFBuffer.hpp:43-44):
 * X is d x ld column-major in the model's precision. */
int bihip_pfout_write_state(bihip_pfout* o, int k, const void* X_host, int ld,
                            const int* as_host);
/* SimulatorNetCDFBuffer::writeState's noise part (R_VAR, SimulatorNetCDFBuffer.hpp:341-352):
 * the noise variables drawn in the step to output k (PZ: alpha; Lorenz-96: w[n]; LG: w),
 * noise_width x ld column-major in the model's precision; a generated model's noise
 * variables in their declaration order. */
int bihip_pfout_write_noise(bihip_pfout* o, int k, const void* Z_host, int ld);
/* ParticleFilterNetCDFBuffer::writeLogWeights (ParticleFilterNetCDFBuffer.hpp:54-55) */
int bihip_pfout_write_log_weights(bihip_pfout* o, int k, const double* lws_host);
/* ParticleFilterNetCDFBuffer::writeLogLikelihood (ParticleFilterNetCDFBuffer.cpp:87-89) */
int bihip_pfout_write_log_likelihood(bihip_pfout* o, double ll);
/* SimulatorNetCDFBuffer::writeClock (SimulatorNetCDFBuffer.cpp:277) */
int bihip_pfout_write_clock(bihip_pfout* o, double clock_us);
int bihip_pfout_close(bihip_pfout* o);
/*
 * The whole recorded history of a filter created with record = 2, as Filter::filter's
 * outputs write it (share/src/bi/filter/Filter.hpp:56-71 with ParticleFilterBuffer::write):
 * output 0 at t0 (the initial samples, identity ancestors, zero log-weights), output k at
 * t0 + k dt after step k, for the steps taken so far; then the log-likelihood.
 */
int bihip_pf_write_netcdf(bihip_pf* pf, const char* path);

/* ================================================================ batched inner filters */
/*
 * B bootstrap particle filters with their own particle counts N_j
 * (1 <= N_j <= 16384) and parameters theta_j, advanced together: the inner
 * loop of MarginalSIR::step (share/src/bi/sampler/MarginalSIR.hpp:346-368),
 * which the reference runs serially over theta with equal N
 * (share/src/bi/state/MarginalSIRState.hpp:198-207).  One persistent fused
 * launch per observation interval: workgroups take whole filters from a work
 * queue (largest first) and run BootstrapPF::step for each.  Filter j draws
 * its Philox streams with batch id stream_base + j (fixed at
 * bihip_batch_init and kept when the filter's content is replaced), so it
 * equals bihip_pf (and the reference filter semantics) run with that batch
 * id.  Models: PZ and LG.  theta_host is B x ntheta; column i replaces
 * bihip_model::par[i].
 */
typedef struct bihip_batch bihip_batch;
typedef struct {
  int B;          /* number of inner filters */
  int T;          /* observations */
  int resampler;  /* BIHIP_SYSTEMATIC or BIHIP_STRATIFIED */
  double essRel;
  double t0, dt;
  uint64_t seed;
  int record;     /* 1 = keep every step's ancestors (T x total ints) */
  int use_graph;  /* 1 = replay each step as a captured hipGraph */
  int block;      /* workgroup size 256/512/1024, or 0 = autotune at the first step */
  long capacity;  /* particles the layout may grow to (0 = the initial total) */
  uint32_t stream_base; /* Philox batch id of filter 0; ids stay below 2^24 */
  int items;      /* with block != 0: particles per thread in the resampling passes
                     (0 = 4; 2 and 8 exist for PZ FP64 RK4 only) */
  int oversub;    /* with block != 0: persistent grid of this many times the resident
                     workgroups per CU (0 = 1) */
} bihip_batch_config;

int bihip_batch_create(bihip_ctx* ctx, const bihip_model* m, const bihip_batch_config* cfg,
                       const int* N_host, const double* theta_host, int ntheta,
                       bihip_batch** b);
int bihip_batch_destroy(bihip_batch* b);
int bihip_batch_set_observations(bihip_batch* b, const double* y_host);
/* Initial particles of every filter; lws = 0; logLikelihood = 0. */
int bihip_batch_init(bihip_batch* b);
/* n observation intervals of every filter (asynchronous). */
int bihip_batch_step(bihip_batch* b, int n);
/* Host copies (any may be NULL): logLikelihood[B], logIncrements[B x T],
 * ess[B x T], resampled[B x T], status[B]. */
int bihip_batch_outputs(bihip_batch* b, double* logLikelihood_host, double* logIncrements_host,
                        double* ess_host, int* resampled_host, int* status_host);
/* Ancestors of filter j at step k (record = 1): N_j ints. */
int bihip_batch_ancestors(bihip_batch* b, int j, int k, int* as_host);
/* The autotuner's choice (workgroup size, persistent grid) and its timings
 * of one step at 256 / 512 / 1024 threads (ms). */
int bihip_batch_tuning(bihip_batch* b, int* block, int* grid, double* ms3_host);
/* The autotuner's other choices (resampling tile width, grid oversubscription)
 * and every configuration it timed: on entry *n is the capacity of cfg_host
 * (n x 3 ints: block, items, oversub) and ms_host (n doubles), on return the
 * number timed.  The reference has no autotuner (its kernels use fixed launch
 * shapes, share/src/bi/cuda/cuda.hpp). */
int bihip_batch_tuning_table(bihip_batch* b, int* items, int* oversub, int* n, int* cfg_host,
                             double* ms_host);
/* HIP-event timing of each fused step launch. */
int bihip_batch_set_timing(bihip_batch* b, int enable);
int bihip_batch_kernel_time(bihip_batch* b, double* ms_host, long* count_host);
/* Total particles over all inner filters. */
long bihip_batch_particles(bihip_batch* b);
/* Replace theta (B x ntheta) for the following steps or initialisation. */
int bihip_batch_set_theta(bihip_batch* b, const double* theta_host);
/* New particle counts (total within the capacity); bihip_batch_init must follow. */
int bihip_batch_set_sizes(bihip_batch* b, const int* N_host);
/* Philox batch id of filter 0 from the next bihip_batch_init on. */
int bihip_batch_set_stream_base(bihip_batch* b, uint32_t base);
/* logIncrements of step k of every filter (B doubles). */
int bihip_batch_increments(bihip_batch* b, int k, double* li_host);
/*
 * Resampling of whole filters, MarginalSIR's interact with the state gather
 * of MarginalSIRState (share/src/bi/sampler/MarginalSIR.hpp:370-393,
 * share/src/bi/state/MarginalSIRState.hpp): filter j takes the particles,
 * log-weights, log-likelihood, size and theta of filter anc_host[j] and keeps
 * its own random streams.  Recorded ancestors of earlier steps are dropped.
 */
int bihip_batch_resample_filters(bihip_batch* b, const int* anc_host);
/*
 * Accepted Metropolis-Hastings proposals (MarginalSIR.hpp:446-466, s1.swap(s2)):
 * filter j of dst takes the whole state of filter j of src where
 * accept_host[j] != 0.  Both batches must have equal sizes, position, model
 * and ntheta; dst's filters keep their own random streams.
 */
int bihip_batch_accept(bihip_batch* dst, bihip_batch* src, const int* accept_host);

/* ================================================================ SMC^2 */
/*
 * SMC^2 over theta (MarginalSIR::sample, share/src/bi/sampler/MarginalSIR.hpp:281-320)
 * with B theta-particles whose inner bootstrap filters (N_host[j] particles)
 * run as one bihip_batch.  theta_i = exp(lth_i) replaces bihip_model::par[i]
 * (i < ntheta); lth_i has the prior N(prior_mean[i], prior_sd[i]^2) and the
 * random-walk proposal lth_i + prop_sd[i] z.  A move (nmoves proposals per
 * slot, each a fresh filter over the observations so far) follows every
 * outer resampling, also after the last observation.  The outer resampler is
 * systematic with threshold essRel_outer * B; the inner filters resample
 * with `resampler` at essRel_inner * N_j.
 */
typedef struct bihip_smc2 bihip_smc2;
typedef struct {
  int B, T, ntheta, nmoves, resampler;
  double essRel_inner, essRel_outer, t0, dt;
  uint64_t seed;
  double prior_mean[8], prior_sd[8], prop_sd[8];
  int use_graph;  /* inner steps replayed as hipGraphs */
  int block;      /* inner workgroup size, 0 = autotune */
  /* 1: the prior of theta is the parameter block of the generated model's file
   * (each of its first ntheta params needs a prior with positive support;
   * prior_mean and prior_sd are unused); the random walk stays on log theta */
  int prior_from_model;
} bihip_smc2_config;

int bihip_smc2_create(bihip_ctx* ctx, const bihip_model* m, const bihip_smc2_config* cfg,
                      const int* N_host, bihip_smc2** s);
int bihip_smc2_destroy(bihip_smc2* s);
/* y_host: T x nobs observations at t0 + k dt, k = 1..T. */
int bihip_smc2_set_observations(bihip_smc2* s, const double* y_host);
/* The whole sampler: init, T x (move, step, interact), final move. */
int bihip_smc2_run(bihip_smc2* s);
/* Host copies (any may be NULL): log evidence, theta [B x ntheta], outer
 * log-weights [B], accepted moves [T + 1], outer ESS [T], resampled [T]. */
int bihip_smc2_outputs(bihip_smc2* s, double* logZ, double* theta_host, double* lw_host,
                       int* accepts_host, double* ess_host, int* resampled_host);

/*
 * PMMH (MarginalMH::sample, share/src/bi/sampler/MarginalMH.hpp:198-266) with
 * B independent chains of C iterations (the first is the prior draw), the B
 * proposals of an iteration filtered together in one bihip_batch.  Uses the
 * bihip_smc2_config fields B, T, ntheta, resampler, essRel_inner, t0, dt,
 * seed, prior_mean, prior_sd, prop_sd, use_graph, block.
 */
typedef struct bihip_pmmh bihip_pmmh;
int bihip_pmmh_create(bihip_ctx* ctx, const bihip_model* m, const bihip_smc2_config* cfg, int C,
                      const int* N_host, bihip_pmmh** s);
int bihip_pmmh_destroy(bihip_pmmh* s);
int bihip_pmmh_set_observations(bihip_pmmh* s, const double* y_host);
int bihip_pmmh_run(bihip_pmmh* s);
/* theta [C x B x ntheta], log-likelihood [C x B], accepted moves per chain [B]
 * (the initial draw counted), any may be NULL. */
int bihip_pmmh_outputs(bihip_pmmh* s, double* theta_host, double* ll_host, int* accepted_host);

/* ================================================================ multilayer perceptron */
/*
 * BF16 MFMA multilayer perceptron with FP32 accumulation: layers
 * l = 0..nlayers-1 map dims[l] -> dims[l+1] as z = W_l a + b_l, GELU in its
 * sigmoid form z sigma(1.702 z) on hidden layers, linear output.  No counterpart in the reference (LibBi has
 * no learned components); it is the learned surrogate flow and guided
 * proposal trained and run beside the filter.  Parameters are FP32 (Adam's
 * master copy) with bfloat16 copies of the weights; the flat host layout is,
 * per layer, W_l row-major [dims[l+1]][dims[l]] then b_l.  Inputs are rounded
 * to bfloat16, hidden activations are bfloat16, outputs FP32.
 * Initialisation: U(-1/sqrt(dims[l]), 1/sqrt(dims[l])) from keyed Philox draws.
 */
typedef struct bihip_mlp bihip_mlp;
#define BIHIP_LOSS_MSE 0    /* mean squared error, targets dims[L] */
#define BIHIP_LOSS_GAUSS 1  /* Gaussian NLL: outputs (mean, log variance), targets dims[L] / 2 */
int bihip_mlp_create(bihip_ctx* ctx, int nlayers, const int* dims, uint64_t seed,
                     bihip_mlp** mlp);
int bihip_mlp_destroy(bihip_mlp* mlp);
long bihip_mlp_nparams(const bihip_mlp* mlp);
int bihip_mlp_set_params(bihip_mlp* mlp, const float* flat_host);  /* resets Adam's moments */
int bihip_mlp_get_params(bihip_mlp* mlp, float* flat_host);
/* Training loss of bihip_mlp_loss_grad / bihip_mlp_train (default BIHIP_LOSS_MSE).
 * BIHIP_LOSS_GAUSS: sum over targets of [(t - mu)^2 e^{-s} + s] / 2, mean over samples
 * and targets, with mu the first dims[L]/2 outputs and s the last. */
int bihip_mlp_set_loss(bihip_mlp* mlp, int loss);
/* Y[n x dims[L]] (leading dimension ldy) = MLP(X[n x dims[0]], leading dimension ldx), device pointers. */
int bihip_mlp_forward(bihip_mlp* mlp, const float* X_dev, long ldx, long n, float* Y_dev,
                      long ldy);
/* The loss over n samples and its gradient (flat host layout). */
int bihip_mlp_loss_grad(bihip_mlp* mlp, const float* X_dev, long ldx, const float* T_dev,
                        long ldt, int n, double* loss_host, float* grad_host);
/* `steps` Adam steps (beta 0.9/0.999, eps 1e-8, learning rate lr) on the loss over
 * minibatches of `batch` rows (a multiple of 64), drawn without replacement from
 * a keyed permutation of the n rows per epoch; loss_host (nullable) gets each
 * step's minibatch loss. */
int bihip_mlp_train(bihip_mlp* mlp, const float* X_dev, long ldx, const float* T_dev, long ldt,
                    long n, int batch, int steps, float lr, uint64_t seed, float* loss_host);

/* ================================================================ learned surrogate flow */
/*
 * A 41-256-256-256-40 MLP standing in for the deterministic Lorenz-96 d=40
 * flow over one observation interval (the DOPRI5 part of
 * Model::transitionSamples, share/tt/cpp/model.hpp.tt:566-573):
 *   u_i = (x_i - x_shift) / x_scale (i < 40), u_40 = (par - par_shift) / par_scale,
 *   x'_i = x_i + dx_scale * MLP(u)_i.
 * One fused kernel per step: BF16 MFMA layers with FP32 accumulation, weights
 * staged in LDS, activations in registers.  No counterpart in the reference.
 */
typedef struct {
  float x_shift, x_scale;
  double par_shift, par_scale;
  float dx_scale;
} bihip_surrogate_norm;
/* Xout = surrogate flow of Xin (FP32, column-major, d = 40 columns of leading dimension ld). */
int bihip_surrogate_flow(bihip_ctx* ctx, bihip_mlp* mlp, const bihip_surrogate_norm* norm,
                         double par, const float* Xin_dev, float* Xout_dev, int ld, int P);
/*
 * Filter steps of an FP32 Lorenz-96 d=40 bihip_pf use the surrogate flow
 * (par = F) in place of the integrator; noise, observation density and
 * resampling are unchanged.  mlp = NULL restores the exact integrator.  The
 * weights are read when attached: attach again after further training.
 */
int bihip_pf_set_surrogate(bihip_pf* pf, bihip_mlp* mlp, const bihip_surrogate_norm* norm);

/* ================================================================ learned guided proposal */
/*
 * Lorenz-96 d=200 filter steps with a learned proposal and the exact
 * importance-weight correction: after the exact flow Phi(x) (DOPRI5, gathered
 * through the ancestors),
 *   x' = Phi + q (m + e^{s/2} eps),  (m, s) = MLP(u) per dimension,
 *   u = ((Phi - x_shift) / x_scale, (y - H Phi) / innov_scale, (par - par_shift) / par_scale),
 *   log w += log p(y | x') + log N(x'; Phi, q^2) - log N(x'; Phi + q m, q^2 e^s),
 * with eps the transition-noise stream, so m = s = 0 is the bootstrap step.
 * The network is 301-256-256-256-400 (m for the 200 dimensions, then s),
 * trained with BIHIP_LOSS_GAUSS on (input u, target (x' - Phi) / q).  The
 * reference's nearest pattern is the auxiliary-weight correction of BridgePF /
 * LookaheadPF (share/src/bi/filter/BridgePF.hpp:114-144); the learned
 * proposal has no counterpart there.
 */
typedef struct {
  float x_shift, x_scale;
  double innov_scale, par_shift, par_scale;
} bihip_guided_norm;
/* One proposal for P particles in place over Phi (X_dev, FP32, d = 200 columns
 * of leading dimension ld), observations y_dev (100 values), Philox step
 * `step`; inc_dev[p] gets the log-weight increment. */
int bihip_guided_propose(bihip_ctx* ctx, bihip_mlp* mlp, const bihip_guided_norm* norm, double par,
                         float* X_dev, int ld, int P, const double* y_dev, uint64_t seed,
                         uint32_t batch, uint32_t step, double noise_sd, double obs_sd,
                         double* inc_dev);
/* Filter steps of an FP32 Lorenz-96 d=200 bihip_pf use the guided proposal
 * (par = F); mlp = NULL restores the bootstrap step.  The weights are read when
 * attached. */
int bihip_pf_set_guided(bihip_pf* pf, bihip_mlp* mlp, const bihip_guided_norm* norm);

/* ================================================================ filter banks */
/*
 * B bootstrap particle filters of P particles each, one parameter vector
 * theta_j per filter, advanced by shared launches: every launch of a step
 * (the resampling chain, the transition or learned flow, the correct step)
 * covers all B filters, filter j on row j of the grid's workgroups.  This is
 * the inner loop of MarginalSIR::step (share/src/bi/sampler/MarginalSIR.hpp:
 * 346-368), which the reference runs serially over theta, for inner filters
 * larger than one workgroup (bihip_batch takes N_j <= 16384) and models of
 * any lane width (Lorenz-96 d = 200).  Filter j equals bihip_pf with Philox
 * batch id cfg->batch + j and bihip_model::par[i] = theta_host[j * ntheta + i]
 * (i < ntheta), step for step: the same ancestors and the same log-weights.
 * Built-in and generated models (a generated model's parameters and inputs
 * are its par slots; no schedule or per-step inputs); cfg->resampler
 * systematic or stratified; cfg->record 0 or 1 (ancestors); cfg->use_graph
 * replays a captured step.
 */
typedef struct bihip_pfbank bihip_pfbank;
int bihip_pfbank_create(bihip_ctx* ctx, const bihip_model* m, const bihip_pf_config* cfg, int B,
                        const double* theta_host, int ntheta, bihip_pfbank** bank);
int bihip_pfbank_destroy(bihip_pfbank* bank);
/* Observations shared by every filter, T x nobs row-major. */
int bihip_pfbank_set_observations(bihip_pfbank* bank, const double* y_host);
/* Initial particles of every filter (Lorenz-96: around mean0_host); lws = 0. */
int bihip_pfbank_init(bihip_pfbank* bank, const double* mean0_host);
/* n observation intervals of every filter (asynchronous). */
int bihip_pfbank_step(bihip_pfbank* bank, int n);
/* Host copies (any may be NULL): logLikelihood[B], logIncrements[B x T],
 * ess[B x T], resampled[B x T], status[B]. */
int bihip_pfbank_outputs(bihip_pfbank* bank, double* logLikelihood_host,
                         double* logIncrements_host, double* ess_host, int* resampled_host,
                         int* status_host);
/* Ancestors of filter j at step k (record = 1): P ints. */
int bihip_pfbank_ancestors(bihip_pfbank* bank, int j, int k, int* as_host);
/* Device state of filter j: d x P of leading dimension ld, and its log-weights. */
int bihip_pfbank_state(bihip_pfbank* bank, int j, void** X_dev, int* ld, double** lws_dev);
/* Steps completed so far. */
int bihip_pfbank_position(bihip_pfbank* bank);
/* The learned surrogate flow (bihip_pf_set_surrogate) or guided proposal
 * (bihip_pf_set_guided) for every filter, each with its own par[0] as the
 * network's parameter input; the filters may differ only in par[0]. */
int bihip_pfbank_set_surrogate(bihip_pfbank* bank, bihip_mlp* mlp,
                               const bihip_surrogate_norm* norm);
int bihip_pfbank_set_guided(bihip_pfbank* bank, bihip_mlp* mlp, const bihip_guided_norm* norm);
/* The outer loop of SMC^2 over a bank (MarginalSIR.hpp:281-488, as bihip_batch_set_theta,
 * _set_stream_base, _increments, _resample_filters and _accept do for bihip_batch):
 * new parameters (B x ntheta) for the following steps or initialisation; the Philox batch
 * id of filter 0 from the next step on; the log-likelihood increments of step k (B
 * doubles); filter j takes the particles, log-weights, running log-likelihood and
 * parameters of filter anc_host[j]; filter j of dst takes those of filter j of src where
 * accept_host[j] != 0 (banks of one shape at the same step). */
int bihip_pfbank_set_theta(bihip_pfbank* bank, const double* theta_host);
int bihip_pfbank_set_stream_base(bihip_pfbank* bank, uint32_t base);
int bihip_pfbank_increments(bihip_pfbank* bank, int k, double* li_host);
int bihip_pfbank_resample_filters(bihip_pfbank* bank, const int* anc_host);
int bihip_pfbank_accept(bihip_pfbank* dst, const bihip_pfbank* src, const int* accept_host);
/* SMC^2 (bihip_smc2_*) whose inner filters are filter banks of P particles each (the
 * theta-particles' filters and a move's proposals) in place of bihip_batch: any inner size,
 * lane-group models (Lorenz-96, Lorenz96d40.bi), the learned surrogate flow.  The
 * bihip_smc2_config field block is unused; prior_from_model as bihip_smc2_create's;
 * Lorenz-96 filters start around mean0_host.  Outputs and keyed draws as
 * bihip_smc2_create's. */
int bihip_smc2_create_bank(bihip_ctx* ctx, const bihip_model* m, const bihip_smc2_config* cfg,
                           int P, const double* mean0_host, bihip_smc2** s);
/* The learned surrogate flow in every inner filter (bihip_pfbank_set_surrogate). */
int bihip_smc2_set_surrogate(bihip_smc2* s, bihip_mlp* mlp, const bihip_surrogate_norm* norm);
/* The learned proposal in every inner filter (bihip_pfbank_set_guided; the auxiliary
 * weights of share/src/bi/filter/BridgePF.hpp:114-144 with a learned bridge). */
int bihip_smc2_set_guided(bihip_smc2* s, bihip_mlp* mlp, const bihip_guided_norm* norm);
/* MarginalMH::sample (share/src/bi/sampler/MarginalMH.hpp:198-266) with the B
 * chains' proposals filtered as one filter bank of B filters of P particles: lane-group
 * models (Lorenz-96, Lorenz96d40.bi) and inner filters of any size.  The
 * bihip_smc2_config field block is unused; prior_from_model as bihip_pmmh_create's;
 * Lorenz-96 filters start around mean0_host.  Run, outputs and keyed draws as
 * bihip_pmmh_create's. */
int bihip_pmmh_create_bank(bihip_ctx* ctx, const bihip_model* m, const bihip_smc2_config* cfg,
                           int C, int P, const double* mean0_host, bihip_pmmh** s);
/* The learned surrogate flow or proposal in the chains' filters (bank form only). */
int bihip_pmmh_set_surrogate(bihip_pmmh* s, bihip_mlp* mlp, const bihip_surrogate_norm* norm);
int bihip_pmmh_set_guided(bihip_pmmh* s, bihip_mlp* mlp, const bihip_guided_norm* norm);
/* Event timing of the bank's launches, as bihip_pf_set_timing / bihip_pf_kernel_time. */
int bihip_pfbank_set_timing(bihip_pfbank* bank, int mode);
int bihip_pfbank_kernel_time(bihip_pfbank* bank, int kernel, double* ms_host, long* count_host);

#ifdef __cplusplus
}
#endif

#endif /* BIHIP_H */
