/*
 * FP64 CPU oracle for LibBi's bootstrap particle-filter step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation in libbi_amd/ is compared against.  Only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it;
 * the product path never calls, links or executes anything in oracle/.
 *
 * It restates, in plain C99 + OpenMP, the host (OpenMP) semantics of the
 * reference (lawmurray/LibBi v1.4.5).  Every function cites the reference
 * file:line it follows.  Arithmetic is double precision throughout and is
 * compiled with -ffp-contract=off so that floating-point expressions are
 * evaluated exactly in the order written.
 *
 * Counter-based random numbers (Philox4x32-10, Salmon et al. 2011) replace
 * the reference's per-OpenMP-thread boost::mt19937 streams
 * (share/src/bi/host/random/RandomHost.cpp:12-27), so the oracle and the GPU
 * consume identical noise and uniforms.  The counter layout is:
 *   counter = { index, step, tag | (batch << 8), block },
 *   key     = { seed & 0xffffffff, seed >> 32 }.
 */
#ifndef LIBBI_AMD_ORACLE_H
#define LIBBI_AMD_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- random-stream tags (shared with the HIP library) ---- */
enum {
  ORC_TAG_INIT = 1,        /* initial-state draws */
  ORC_TAG_NOISE = 2,       /* transition noise */
  ORC_TAG_SYSTEMATIC = 3,  /* systematic resampler offset */
  ORC_TAG_STRATIFIED = 4,  /* stratified resampler alphas */
  ORC_TAG_MULTINOMIAL = 5, /* multinomial resampler uniforms */
  ORC_TAG_OBS = 6,         /* observation simulation */
  ORC_TAG_THETA = 7,       /* SMC^2 prior draws of theta */
  ORC_TAG_PROPOSAL = 8,    /* SMC^2 random-walk proposals */
  ORC_TAG_ACCEPT = 9,      /* SMC^2 acceptance uniforms */
  ORC_TAG_OUTER = 10       /* SMC^2 outer systematic offset */
};

/* ---- models ---- */
enum { ORC_MODEL_PZ = 0, ORC_MODEL_L96 = 1, ORC_MODEL_LG = 2 };
enum { ORC_RK4 = 0, ORC_DOPRI5 = 1 };
enum { ORC_SYSTEMATIC = 0, ORC_STRATIFIED = 1, ORC_MULTINOMIAL = 2 };
enum { ORC_OK = 0, ORC_DEGENERATE = 1, ORC_EINVAL = 2 };

/*
 * Model and integrator configuration.
 *
 * par[] meaning per model:
 *  PZ : mu, sigma, c, e, m_l, m_q, obs_sd, P0_logmean, P0_logsd, Z0_logmean, Z0_logsd
 *  L96: F, noise_sd, obs_sd, init_sd            (initial mean passed separately)
 *  LG : lambda, noise_sd, obs_sd, m0, s0
 */
typedef struct {
  int model;        /* ORC_MODEL_* */
  int d;            /* state dimension */
  int integrator;   /* ORC_RK4 or ORC_DOPRI5 */
  int nsteps;       /* DOPRI5 maximum number of steps per interval */
  int host_quirk;   /* DOPRI5: 1 = evaluate k7 at x5 as the reference host does */
  int noise_bits;   /* 32: normals from 32-bit uniforms; 24: from 24-bit (FP32 kernels) */
  double h0, atol, rtol;
  double par[16];
} orc_model;

/* ---- Philox and variates ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_philox_block(uint64_t seed, uint32_t index, uint32_t step, uint32_t tag,
    uint32_t batch, uint32_t block, uint32_t out[4]);
double orc_u53(uint32_t r0, uint32_t r1);
/* n uniforms in [0,1): u_j from counter {j, step, tag|(batch<<8), 0}, words 0,1 */
void orc_uniforms(uint64_t seed, uint32_t step, uint32_t tag, uint32_t batch, int n, double* u);
/* one uniform in [0,1) from block {index, step, tag|(batch<<8), 0} */
void orc_uniform_at(uint64_t seed, uint32_t index, uint32_t step, uint32_t tag, uint32_t batch,
    double* u);
/* n standard normals for one index: normal j from block j/4, word pair (j%4)/2 */
void orc_normals(uint64_t seed, uint32_t index, uint32_t step, uint32_t tag, uint32_t batch,
    int n, int bits, double* z);

/* ---- model functions (X is column-major, leading dimension ld >= P) ---- */
int orc_nobs(const orc_model* m);
void orc_init_samples(const orc_model* m, uint64_t seed, uint32_t batch, int P, double* X,
    int ld, const double* mean0);
void orc_transition(const orc_model* m, uint64_t seed, uint32_t batch, uint32_t step,
    double t1, double t2, int P, double* X, int ld, int* nsteps_out);
void orc_obs_logdensity(const orc_model* m, const double* y, int P, const double* X, int ld,
    double* lws);
void orc_obs_sample(const orc_model* m, uint64_t seed, uint32_t batch, uint32_t step,
    const double* x, double* y);
/* deterministic flow of one particle over [t1,t2] (no noise); returns steps taken */
int orc_flow(const orc_model* m, double alpha, double t1, double t2, double* x);
void orc_rhs(const orc_model* m, double alpha, double t, const double* x, double* dx);

/* ---- primitives (share/src/bi/primitive/vector_primitive.hpp) ---- */
double orc_max_reduce(const double* x, int n);
double orc_logsumexp_reduce(const double* x, int n);
double orc_ess_reduce(const double* lws, int n, double* lW);
double orc_sumexpu_inclusive_scan(const double* lws, int n, double* Ws);

/* ---- resamplers ---- */
int orc_systematic_cumulative_offspring(const double* Ws, int P, int n, double a, int* Os);
int orc_stratified_cumulative_offspring(const double* Ws, int P, int n, const double* alphas,
    int* Os);
int orc_multinomial_ancestors(const double* Ws, int P, int n, const double* U, int* as);
void orc_ancestors_to_offspring(const int* as, int n, int P, int* os);
void orc_offspring_to_cumulative(const int* os, int P, int* Os);
void orc_cumulative_offspring_to_ancestors(const int* Os, int P, int* as);
void orc_cumulative_offspring_to_ancestors_permute(const int* Os, int P, int* as);
void orc_offspring_to_ancestors_permute(const int* os, int P, int* as);
void orc_permute(int* as, int P);
void orc_gather_rows(const int* as, int P, int ncols, double* X, int ld);

/* ---- bootstrap particle filter ---- */
typedef struct {
  int P;            /* number of particles */
  int T;            /* number of observations, at t0 + k*dt, k = 1..T */
  int resampler;    /* ORC_SYSTEMATIC / ORC_STRATIFIED / ORC_MULTINOMIAL */
  double essRel;    /* resample when ess < essRel * P */
  double t0, dt;
  uint64_t seed;
  uint32_t batch;
  /* optional schedule (both nullable): times[0..T] replace t0 + k*dt, and
   * observed[k] = 0 makes step k a prediction without correction */
  const double* times;
  const int* observed;
  /* optional inputs (nullable): the model's 16 parameters for each step,
   * T x 16, in force over [t_k, t_{k+1}) */
  const double* inputs;
  /* optional imposed resampling (nullable; both or neither): at step k the
   * filter resamples iff imposed_resampled[k] != 0, and then takes its
   * ancestors from imposed_ancestors[k P .. k P + P - 1] in place of the ones
   * its own weights and uniforms would give (the gather and lws[:] =
   * logLikelihood follow as usual).  Re-runs another filter's resampling
   * history in FP64, e.g. that of an FP32 run. */
  const int* imposed_resampled;
  const int* imposed_ancestors;
} orc_pf_cfg;

/*
 * Run the filter.  y is T x nobs (row k = observation at t0+(k+1)dt).
 * Outputs (any may be NULL): logIncrements[T], ess[T], resampled[T],
 * ancestors[T*P] (ancestry stored at each step, identity if not resampled),
 * Xout (final state, column-major ld=P).  Returns ORC_OK or ORC_DEGENERATE.
 */
int orc_pf_run(const orc_model* m, const orc_pf_cfg* cfg, const double* y,
    const double* mean0, double* logZ, double* logIncrements, double* ess,
    int* resampled, int* ancestors, double* Xout);

/* step-wise filter (the state of one BootstrapPF) */
typedef struct orc_pf orc_pf;
orc_pf* orc_pf_new(const orc_model* m, int P, int resampler, double essRel, double t0,
    double dt, uint64_t seed, uint32_t batch);
void orc_pf_free(orc_pf* f);
void orc_pf_init(orc_pf* f, const double* mean0);
/* Schedule (Schedule, share/src/bi/state/Schedule.hpp:214-337, restricted to
 * output and observation times): times[0..T] and observed[0..T-1], both
 * nullable and kept by reference. */
void orc_pf_set_schedule(orc_pf* f, const double* times, const int* observed);
/* Inputs (Forcer, share/src/bi/simulator/Forcer.hpp, as parameter values per
 * step): inputs[16 k .. 16 k + 15] replace the model's parameters during step
 * k (transition and observation density).  Nullable, kept by reference. */
void orc_pf_set_inputs(orc_pf* f, const double* inputs);
/* Imposed resampling history (see orc_pf_cfg), both nullable, kept by reference. */
void orc_pf_impose(orc_pf* f, const int* resampled, const int* ancestors);
/* One step whose resampling at its start is imposed: resampled (0/1) and,
 * when set, the P ancestors.  Otherwise as orc_pf_step. */
int orc_pf_step_imposed(orc_pf* f, const double* y, int resampled, const int* ancestors,
    double* logIncrement);
/* Replace the filter's state after k steps: X (d x P column-major, ld = P),
 * log-weights, running log-likelihood and the ESS of the last correction, so
 * that the next orc_pf_step continues from another filter's state. */
void orc_pf_set_state(orc_pf* f, int k, const double* X, const double* lws,
    double logLikelihood, double ess);
/* The current state (d x P column-major, ld = P) and log-weights (nullable). */
void orc_pf_get_state(const orc_pf* f, double* X, double* lws);
int orc_pf_step(orc_pf* f, const double* y, double* logIncrement, int* resampled, int* ancestors);
double orc_pf_loglikelihood(const orc_pf* f);
double orc_pf_ess(const orc_pf* f);
int orc_pf_size(const orc_pf* f);
double orc_pf_term(const orc_pf* f);
void orc_pf_copy(orc_pf* dst, const orc_pf* src);

/* SMC^2 over theta in log space (MarginalSIR), inner bootstrap filters with
 * particle counts N[j]; see filter.c for the keyed draws. */
typedef struct {
  int B, T, ntheta, nmoves, resampler;
  double essRel_inner, essRel_outer, t0, dt;
  uint64_t seed;
  double prior_mean[8], prior_sd[8], prop_sd[8];
} orc_smc2_cfg;
/* outputs (nullable): logZ, theta_out[B x ntheta], lw_out[B], accepts_out[T+1],
 * ess_out[T], resampled_out[T] */
int orc_smc2_run(const orc_model* base, const orc_smc2_cfg* c, const int* N, const double* y,
    double* logZ, double* theta_out, double* lw_out, int* accepts_out, double* ess_out,
    int* resampled_out);

/* PMMH (MarginalMH) with B independent chains and C iterations (the first is
 * the initial draw), reusing orc_smc2_cfg (B, T, ntheta, resampler,
 * essRel_inner, t0, dt, seed, prior, prop_sd).  Outputs (nullable): theta_out
 * [C x B x ntheta], ll_out [C x B], accepted_out [B] (counts, the initial
 * draw included). */
int orc_pmmh_run(const orc_model* base, const orc_smc2_cfg* c, int C, const int* N,
    const double* y, double* theta_out, double* ll_out, int* accepted_out);

/* ---- multilayer perceptron (mlp.c): GELU z sigma(1.702 z) hidden layers, linear output,
 * flat parameters [W_l (d_{l+1} x d_l, row-major), b_l] for l = 0..L-1 ---- */
double orc_bf16(double x);
size_t orc_mlp_nparams(int L, const int* dims);
void orc_mlp_forward(int L, const int* dims, const double* params, const double* X, int n,
    int bf16, double* Y);
/* loss over n samples and its gradient: ORC_LOSS_MSE (targets dims[L]) or
 * ORC_LOSS_GAUSS (outputs: mean then log variance; targets dims[L] / 2) */
enum { ORC_LOSS_MSE = 0, ORC_LOSS_GAUSS = 1 };
double orc_mlp_loss_grad(int L, const int* dims, const double* params, const double* X,
    const double* T, int n, int loss_kind, int bf16, double* grad);

/* The learned guided proposal's step (guided.c) over P flowed states Phi
 * (d x P column-major, ld = P; d even, the even components observed in y):
 * inputs u = ((Phi - norm[0]) / norm[1], (y - H Phi) / norm[2],
 * (F - norm[3]) / norm[4]), (m, s) = MLP(u) (dims[0] = d + d/2 + 1,
 * dims[L] = 2 d), x' = Phi + q (m + e^{s/2} eps) with the transition-noise
 * normals of (seed, p, step, batch), and the log-weight increment
 * log p(y | x') + log N(x'; Phi, q^2 I) - log N(x'; Phi + q m, q^2 e^s).
 * Outputs Xout (d x P) and inc (P). */
int orc_guided_step(int L, const int* dims, const double* params, int bf16,
    const double* norm, double F, int d, int P, const double* Phi, const double* y,
    uint64_t seed, uint32_t batch, uint32_t step, double q, double obs_sd, int noise_bits,
    double* Xout, double* inc);

/* scalar linear-Gaussian model: exact log-likelihood by the Kalman filter */
double orc_kalman_loglik_lg(const orc_model* m, double dt, int T, const double* y);
/* the same over a schedule: interval k is [times[k], times[k+1]], and step k
 * predicts without an update when observed[k] == 0 */
double orc_kalman_loglik_lg_sched(const orc_model* m, const double* times, const int* observed,
    int T, const double* y);

#ifdef __cplusplus
}
#endif

#endif
