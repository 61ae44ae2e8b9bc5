/*
 * Bootstrap particle filter driver of the oracle, the Kalman filter that
 * pins its log-likelihood conventions, and the SMC^2 sampler built on it.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The filter restates Filter<F>::filter (share/src/bi/filter/Filter.hpp:56-71)
 * with BootstrapPF::step/correct/resample/term
 * (share/src/bi/filter/BootstrapPF.hpp:156-196) and Resampler<R>::resample
 * (share/src/bi/resampler/Resampler.hpp:165-182):
 *   - the schedule is t_k = t0 + k dt, k = 0..T, observed for k >= 1;
 *   - resample at t_k when observed and ess < essRel * P (strict);
 *     then the in-place gather and lws[:] = logLikelihood;
 *   - predict to t_{k+1}, then correct: lws += log p(y | x),
 *     ess, lW = ess_reduce(lws), logIncrements[k] = lW - logLikelihood,
 *     logLikelihood = lW;
 *   - term: logZ = logsumexp(lws) - log P.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

struct orc_pf {
  orc_model m;
  int P, resampler;
  double essRel, t0, dt;
  uint64_t seed;
  uint32_t batch;
  double *X, *lws, *Ws, *U;
  int *Os, *os, *as, *a0;
  double logLikelihood, ess;
  int k;
  const double* times;  /* nullable: t0 + k dt */
  const int* observed;  /* nullable: every step observed */
  const double* inputs; /* nullable: T x 16 parameters per step */
  const int* imp_r;     /* nullable: imposed resampling flags per step */
  const int* imp_a;     /* nullable: imposed ancestors, T x P */
};

orc_pf* orc_pf_new(const orc_model* m, int P, int resampler, double essRel, double t0,
                   double dt, uint64_t seed, uint32_t batch) {
  orc_pf* f = (orc_pf*)calloc(1, sizeof(orc_pf));
  f->m = *m;
  f->P = P;
  f->resampler = resampler;
  f->essRel = essRel;
  f->t0 = t0;
  f->dt = dt;
  f->seed = seed;
  f->batch = batch;
  f->X = (double*)malloc(sizeof(double) * (size_t)m->d * P);
  f->lws = (double*)calloc((size_t)P, sizeof(double));
  f->Ws = (double*)malloc(sizeof(double) * (size_t)P);
  f->U = (double*)malloc(sizeof(double) * (size_t)P);
  f->Os = (int*)malloc(sizeof(int) * (size_t)P);
  f->os = (int*)malloc(sizeof(int) * (size_t)P);
  f->as = (int*)malloc(sizeof(int) * (size_t)P);
  f->a0 = (int*)malloc(sizeof(int) * (size_t)P);
  return f;
}

void orc_pf_free(orc_pf* f) {
  if (!f) return;
  free(f->X); free(f->lws); free(f->Ws); free(f->U);
  free(f->Os); free(f->os); free(f->as); free(f->a0);
  free(f);
}

/* Filter::init: initial samples, lws = 0, logLikelihood = 0 */
void orc_pf_init(orc_pf* f, const double* mean0) {
  orc_init_samples(&f->m, f->seed, f->batch, f->P, f->X, f->P, mean0);
  memset(f->lws, 0, sizeof(double) * (size_t)f->P);
  f->logLikelihood = 0.0;
  f->ess = 0.0;
  f->k = 0;
}

void orc_pf_set_schedule(orc_pf* f, const double* times, const int* observed) {
  f->times = times;
  f->observed = observed;
}

void orc_pf_set_inputs(orc_pf* f, const double* inputs) { f->inputs = inputs; }

void orc_pf_impose(orc_pf* f, const int* resampled, const int* ancestors) {
  f->imp_r = resampled;
  f->imp_a = ancestors;
}

void orc_pf_set_state(orc_pf* f, int k, const double* X, const double* lws,
                      double logLikelihood, double ess) {
  memcpy(f->X, X, sizeof(double) * (size_t)f->m.d * f->P);
  memcpy(f->lws, lws, sizeof(double) * (size_t)f->P);
  f->logLikelihood = logLikelihood;
  f->ess = ess;
  f->k = k;
}

void orc_pf_get_state(const orc_pf* f, double* X, double* lws) {
  if (X) memcpy(X, f->X, sizeof(double) * (size_t)f->m.d * f->P);
  if (lws) memcpy(lws, f->lws, sizeof(double) * (size_t)f->P);
}

static double pf_time(const orc_pf* f, int k) {
  return f->times ? f->times[k] : f->t0 + k * f->dt;
}

static int pf_observed(const orc_pf* f, int k) { return f->observed ? f->observed[k] != 0 : 1; }

/* One BootstrapPF::step from t_k to t_{k+1} with observation y_{k+1}.
 * imp_r < 0: the filter decides its resampling itself; otherwise imp_r (0/1)
 * imposes it, with the P ancestors imp_a when set.  resampled/ancestors
 * (nullable) report the resampling at t_k.  Returns ORC_OK or
 * ORC_DEGENERATE; *logIncrement gets lW - logLikelihood. */
static int pf_step(orc_pf* f, const double* y, double* logIncrement, int* resampled,
                   int* ancestors, int imp_r, const int* imp_a) {
  const int P = f->P, D = f->m.d, k = f->k;
  const double t1 = pf_time(f, k), t2 = pf_time(f, k + 1);
  /* Resampler::resample (share/src/bi/resampler/Resampler.hpp:165-182): only at
   * an observed element, when the ESS fell below the threshold */
  const int imposed = imp_r >= 0;
  const int r = imposed ? imp_r != 0
                        : k >= 1 && pf_observed(f, k - 1) && f->ess < f->essRel * P;
  int rc = ORC_OK;
  if (r && imposed) {
    /* another filter's resampling at t_k: its ancestors, then the same
     * gather and reset of the log-weights */
    memcpy(f->as, imp_a, sizeof(int) * (size_t)P);
    orc_gather_rows(f->as, P, D, f->X, P);
    for (int i = 0; i < P; ++i) f->lws[i] = f->logLikelihood;
  } else if (r) {
    orc_sumexpu_inclusive_scan(f->lws, P, f->Ws);
    if (f->resampler == ORC_SYSTEMATIC) {
      double a;
      orc_uniforms(f->seed, (uint32_t)k, ORC_TAG_SYSTEMATIC, f->batch, 1, &a);
      rc = orc_systematic_cumulative_offspring(f->Ws, P, P, a, f->Os);
    } else if (f->resampler == ORC_STRATIFIED) {
      orc_uniforms(f->seed, (uint32_t)k, ORC_TAG_STRATIFIED, f->batch, P, f->U);
      rc = orc_stratified_cumulative_offspring(f->Ws, P, P, f->U, f->Os);
    } else {
      orc_uniforms(f->seed, (uint32_t)k, ORC_TAG_MULTINOMIAL, f->batch, P, f->U);
      rc = orc_multinomial_ancestors(f->Ws, P, P, f->U, f->a0);
      if (rc == ORC_OK) {
        orc_ancestors_to_offspring(f->a0, P, P, f->os);
        orc_offspring_to_cumulative(f->os, P, f->Os);
      }
    }
    if (rc != ORC_OK) {
      f->logLikelihood = -INFINITY;  /* a degenerate filter has no likelihood */
      return rc;
    }
    orc_cumulative_offspring_to_ancestors_permute(f->Os, P, f->as);
    orc_gather_rows(f->as, P, D, f->X, P);
    for (int i = 0; i < P; ++i) f->lws[i] = f->logLikelihood;
  } else {
    for (int i = 0; i < P; ++i) f->as[i] = i;
  }
  if (ancestors) memcpy(ancestors, f->as, sizeof(int) * (size_t)P);
  if (resampled) *resampled = r;

  orc_model mk = f->m;  /* the inputs in force over [t_k, t_{k+1}) */
  if (f->inputs) memcpy(mk.par, f->inputs + (size_t)16 * k, sizeof(double) * 16);
  orc_transition(&mk, f->seed, f->batch, (uint32_t)(k + 1), t1, t2, P, f->X, P, NULL);
  /* BootstrapPF::correct (share/src/bi/filter/BootstrapPF.hpp:171-181): only
   * at an observed element */
  if (pf_observed(f, k)) {
    orc_obs_logdensity(&mk, y, P, f->X, P, f->lws);
    double lW;
    f->ess = orc_ess_reduce(f->lws, P, &lW);
    if (logIncrement) *logIncrement = lW - f->logLikelihood;
    f->logLikelihood = lW;
  } else if (logIncrement) {
    *logIncrement = 0.0;
  }
  f->k = k + 1;
  return ORC_OK;
}

int orc_pf_step(orc_pf* f, const double* y, double* logIncrement, int* resampled,
                int* ancestors) {
  if (f->imp_r != NULL && f->imp_a != NULL)
    return pf_step(f, y, logIncrement, resampled, ancestors, f->imp_r[f->k] != 0,
                   f->imp_a + (size_t)f->k * f->P);
  return pf_step(f, y, logIncrement, resampled, ancestors, -1, NULL);
}

int orc_pf_step_imposed(orc_pf* f, const double* y, int resampled, const int* ancestors,
                        double* logIncrement) {
  return pf_step(f, y, logIncrement, NULL, NULL, resampled != 0, ancestors);
}

double orc_pf_loglikelihood(const orc_pf* f) { return f->logLikelihood; }
double orc_pf_ess(const orc_pf* f) { return f->ess; }
int orc_pf_size(const orc_pf* f) { return f->P; }

/* BootstrapPF::term */
double orc_pf_term(const orc_pf* f) {
  return orc_logsumexp_reduce(f->lws, f->P) - log((double)f->P);
}

/* Copy the filter state of src into dst (model, particles, weights,
 * scalars), keeping dst's random-stream batch id. */
void orc_pf_copy(orc_pf* dst, const orc_pf* src) {
  const uint32_t batch = dst->batch;
  if (dst->P != src->P || dst->m.d != src->m.d) {
    orc_pf* tmp = orc_pf_new(&src->m, src->P, src->resampler, src->essRel, src->t0, src->dt,
                             src->seed, batch);
    free(dst->X); free(dst->lws); free(dst->Ws); free(dst->U);
    free(dst->Os); free(dst->os); free(dst->as); free(dst->a0);
    *dst = *tmp;
    free(tmp);
  }
  dst->m = src->m;
  dst->P = src->P;
  memcpy(dst->X, src->X, sizeof(double) * (size_t)src->m.d * src->P);
  memcpy(dst->lws, src->lws, sizeof(double) * (size_t)src->P);
  dst->logLikelihood = src->logLikelihood;
  dst->ess = src->ess;
  dst->k = src->k;
  dst->batch = batch;
}

int orc_pf_run(const orc_model* m, const orc_pf_cfg* cfg, const double* y,
    const double* mean0, double* logZ, double* logIncrements, double* essOut,
    int* resampledOut, int* ancestorsOut, double* Xout) {
  const int P = cfg->P, T = cfg->T, nobs = orc_nobs(m);
  orc_pf* f = orc_pf_new(m, P, cfg->resampler, cfg->essRel, cfg->t0, cfg->dt, cfg->seed,
                         cfg->batch);
  int rc = ORC_OK;
  orc_pf_init(f, mean0);
  orc_pf_set_schedule(f, cfg->times, cfg->observed);
  orc_pf_set_inputs(f, cfg->inputs);
  orc_pf_impose(f, cfg->imposed_resampled, cfg->imposed_ancestors);
  for (int k = 0; k < T; ++k) {
    double li;
    int r;
    rc = orc_pf_step(f, y + (size_t)k * nobs, &li, &r,
                     ancestorsOut ? ancestorsOut + (size_t)k * P : NULL);
    if (rc != ORC_OK) break;
    if (resampledOut) resampledOut[k] = r;
    if (essOut) essOut[k] = f->ess;
    if (logIncrements) logIncrements[k] = li;
  }
  if (rc == ORC_OK && logZ) *logZ = orc_pf_term(f);
  if (Xout) memcpy(Xout, f->X, sizeof(double) * (size_t)m->d * P);
  orc_pf_free(f);
  return rc;
}

/* exact log-likelihood of the scalar LG model: x' = A x + q e, y = x + r e,
 * A = exp(-lambda dt), x0 ~ N(m0, s0^2), observations at k = 1..T */
double orc_kalman_loglik_lg(const orc_model* m, double dt, int T, const double* y) {
  const double A = exp(-m->par[0] * dt), q = m->par[1], r = m->par[2];
  double mean = m->par[3], var = m->par[4] * m->par[4], ll = 0.0;
  for (int k = 0; k < T; ++k) {
    mean = A * mean;
    var = A * A * var + q * q;
    const double S = var + r * r, v = y[k] - mean;
    ll += -0.5 * log(2.0 * M_PI * S) - 0.5 * v * v / S;
    const double K = var / S;
    mean = mean + K * v;
    var = (1.0 - K) * var;
  }
  return ll;
}

double orc_kalman_loglik_lg_sched(const orc_model* m, const double* times, const int* observed,
                                  int T, const double* y) {
  const double q = m->par[1], r = m->par[2];
  double mean = m->par[3], var = m->par[4] * m->par[4], ll = 0.0;
  for (int k = 0; k < T; ++k) {
    const double A = exp(-m->par[0] * (times[k + 1] - times[k]));
    mean = A * mean;
    var = A * A * var + q * q;
    if (observed && !observed[k]) continue;
    const double S = var + r * r, v = y[k] - mean;
    ll += -0.5 * log(2.0 * M_PI * S) - 0.5 * v * v / S;
    const double K = var / S;
    mean = mean + K * v;
    var = (1.0 - K) * var;
  }
  return ll;
}

/* ------------------------------------------------------------------ SMC^2 */
/*
 * MarginalSIR::sample (share/src/bi/sampler/MarginalSIR.hpp:281-320):
 * init (:322-344), then per observation move (:395-488), step (:346-368)
 * and interact (:370-393).  theta lives in log space with a Gaussian prior
 * per component and a Gaussian random-walk proposal (the model's
 * proposal_parameter block, share/src/bi/simulator/Simulator.hpp:336-378),
 * so prior and proposal Jacobians cancel in the acceptance ratio.
 * Keyed draws: prior normals {j, 0, THETA, 0}; proposal normals
 * {j, k, PROPOSAL, m}; accept uniform {j, k, ACCEPT, m}; outer systematic
 * offset {0, k, OUTER, 0}.  Inner filter of slot j draws with batch id j;
 * the proposal filter of (k, m, j) with batch id B (1 + m + nmoves k) + j.
 */
static double log_prior(const orc_smc2_cfg* c, const double* lth) {
  double lp = 0.0;
  for (int i = 0; i < c->ntheta; ++i) {
    const double z = (lth[i] - c->prior_mean[i]) / c->prior_sd[i];
    lp += -0.5 * z * z;
  }
  return lp;
}

static void model_with(const orc_model* base, const orc_smc2_cfg* c, const double* lth,
                       orc_model* out) {
  *out = *base;
  for (int i = 0; i < c->ntheta; ++i) out->par[i] = exp(lth[i]);
}

/* outer systematic resampling of the slots (host layout permutation) */
static void outer_ancestors(const double* lw, int B, double a, int* as, double* Ws, int* Os) {
  orc_sumexpu_inclusive_scan(lw, B, Ws);
  orc_systematic_cumulative_offspring(Ws, B, B, a, Os);
  orc_cumulative_offspring_to_ancestors_permute(Os, B, as);
}

/* move: PMMH on every slot (MarginalSIR.hpp:395-488); returns accepts */
static int smc2_move(const orc_model* base, const orc_smc2_cfg* c, const double* y, int k,
                     orc_pf** f, double* lth, double* lth2) {
  const int B = c->B, nt = c->ntheta, nobs = orc_nobs(base);
  int naccept = 0;
  for (int m = 0; m < c->nmoves; ++m) {
    for (int j = 0; j < B; ++j) {
      double z[8], u;
      orc_normals(c->seed, (uint32_t)j, (uint32_t)k, ORC_TAG_PROPOSAL, (uint32_t)m, nt, 32, z);
      for (int i = 0; i < nt; ++i)
        lth2[(size_t)j * nt + i] = lth[(size_t)j * nt + i] + c->prop_sd[i] * z[i];
      orc_model mj;
      model_with(base, c, lth2 + (size_t)j * nt, &mj);
      const uint32_t pb = (uint32_t)(B * (1 + m + c->nmoves * k) + j);
      orc_pf* p = orc_pf_new(&mj, f[j]->P, c->resampler, c->essRel_inner, c->t0, c->dt,
                             c->seed, pb);
      orc_pf_init(p, NULL);
      int rc = ORC_OK;
      for (int s = 0; s < k && rc == ORC_OK; ++s)
        rc = orc_pf_step(p, y + (size_t)s * nobs, NULL, NULL, NULL);
      const double ll2 = rc == ORC_OK ? p->logLikelihood : -INFINITY;
      const double ll1 = f[j]->logLikelihood;
      orc_uniform_at(c->seed, (uint32_t)j, (uint32_t)k, ORC_TAG_ACCEPT, (uint32_t)m, &u);
      int accept;
      if (!isfinite(ll2)) accept = 0;
      else if (!isfinite(ll1)) accept = 1;
      else {
        const double logratio = ll2 - ll1 + log_prior(c, lth2 + (size_t)j * nt) -
                                log_prior(c, lth + (size_t)j * nt);
        accept = log(u) < logratio;
      }
      if (accept) {
        orc_pf_copy(f[j], p);
        memcpy(lth + (size_t)j * nt, lth2 + (size_t)j * nt, sizeof(double) * nt);
        ++naccept;
      }
      orc_pf_free(p);
    }
  }
  return naccept;
}

int orc_smc2_run(const orc_model* base, const orc_smc2_cfg* c, const int* N, const double* y,
                 double* logZ, double* theta_out, double* lw_out, int* accepts_out,
                 double* ess_out, int* resampled_out) {
  const int B = c->B, T = c->T, nt = c->ntheta, nobs = orc_nobs(base);
  orc_pf** f = (orc_pf**)calloc((size_t)B, sizeof(orc_pf*));
  orc_pf** g = (orc_pf**)calloc((size_t)B, sizeof(orc_pf*));
  double* lth = (double*)malloc(sizeof(double) * (size_t)B * nt);
  double* lth2 = (double*)malloc(sizeof(double) * (size_t)B * nt);
  double* lw = (double*)calloc((size_t)B, sizeof(double));
  double* Ws = (double*)malloc(sizeof(double) * (size_t)B);
  int* Os = (int*)malloc(sizeof(int) * (size_t)B);
  int* as = (int*)malloc(sizeof(int) * (size_t)B);
  double logLik = 0.0;
  int lastResample = 0;

  /* init (:322-344): theta from the prior, inner filters from their initial laws */
  for (int j = 0; j < B; ++j) {
    double z[8];
    orc_normals(c->seed, (uint32_t)j, 0, ORC_TAG_THETA, 0, nt, 32, z);
    for (int i = 0; i < nt; ++i) lth[(size_t)j * nt + i] = c->prior_mean[i] + c->prior_sd[i] * z[i];
    orc_model mj;
    model_with(base, c, lth + (size_t)j * nt, &mj);
    f[j] = orc_pf_new(&mj, N[j], c->resampler, c->essRel_inner, c->t0, c->dt, c->seed,
                      (uint32_t)j);
    orc_pf_init(f[j], NULL);
  }

  for (int k = 0; k <= T; ++k) {
    const int naccept = lastResample ? smc2_move(base, c, y, k, f, lth, lth2) : 0;
    if (accepts_out) accepts_out[k] = naccept;
    if (k == T) break;  /* the final move (:301-303) follows the last observation */

    /* step (:346-368): every inner filter, lw += logIncrement */
    for (int j = 0; j < B; ++j) {
      double li;
      if (orc_pf_step(f[j], y + (size_t)k * nobs, &li, NULL, NULL) != ORC_OK) li = -INFINITY;
      lw[j] += li;
    }

    /* interact (:370-393): outer ESS and marginal likelihood, then resampling */
    double lW;
    const double ess = orc_ess_reduce(lw, B, &lW);
    if (ess_out) ess_out[k] = ess;
    logLik = lW;
    const int r = ess < c->essRel_outer * B;
    if (resampled_out) resampled_out[k] = r;
    if (r) {
      double a;
      orc_uniforms(c->seed, (uint32_t)k, ORC_TAG_OUTER, 0, 1, &a);
      outer_ancestors(lw, B, a, as, Ws, Os);
      for (int j = 0; j < B; ++j) {
        g[j] = orc_pf_new(&f[j]->m, 1, c->resampler, c->essRel_inner, c->t0, c->dt, c->seed,
                          (uint32_t)j);
        orc_pf_copy(g[j], f[as[j]]);
        memcpy(lth2 + (size_t)j * nt, lth + (size_t)as[j] * nt, sizeof(double) * nt);
      }
      for (int j = 0; j < B; ++j) {
        orc_pf_free(f[j]);
        f[j] = g[j];
        g[j] = NULL;
        lw[j] = logLik;
      }
      memcpy(lth, lth2, sizeof(double) * (size_t)B * nt);
    }
    lastResample = r;
  }
  if (logZ) *logZ = logLik;
  if (theta_out)
    for (size_t i = 0; i < (size_t)B * nt; ++i) theta_out[i] = exp(lth[i]);
  if (lw_out) memcpy(lw_out, lw, sizeof(double) * B);
  for (int j = 0; j < B; ++j) orc_pf_free(f[j]);
  free(f); free(g); free(lth); free(lth2); free(lw); free(Ws); free(Os); free(as);
  return ORC_OK;
}

/* ------------------------------------------------------------------ PMMH */
/*
 * MarginalMH::sample (share/src/bi/sampler/MarginalMH.hpp:198-215) for B
 * independent chains, each with inner filters of N[j] particles:
 * init (:217-226) filters theta_0 from the prior, then every iteration
 * proposes (:228-244) theta' = theta + prop_sd z in log space, filters it over
 * the whole schedule, and accepts (:246-266) with probability
 * min(1, exp(ll' - ll + log prior' - log prior)); a non-finite ll' is
 * rejected, a non-finite ll accepts.  Keyed draws: prior normals
 * {j, 0, THETA, 0}; proposal normals {j, c, PROPOSAL, 0}; acceptance uniform
 * {j, c, ACCEPT, 0}; the filter of iteration c of chain j draws with batch id
 * B c + j.
 */
static double pmmh_loglik(const orc_model* base, const orc_smc2_cfg* c, const double* lth,
                          int N, const double* y, uint32_t batch) {
  orc_model mj;
  model_with(base, c, lth, &mj);
  const int nobs = orc_nobs(base);
  orc_pf* f = orc_pf_new(&mj, N, c->resampler, c->essRel_inner, c->t0, c->dt, c->seed, batch);
  orc_pf_init(f, NULL);
  int rc = ORC_OK;
  for (int s = 0; s < c->T && rc == ORC_OK; ++s) rc = orc_pf_step(f, y + (size_t)s * nobs, NULL, NULL, NULL);
  const double ll = rc == ORC_OK ? f->logLikelihood : -INFINITY;
  orc_pf_free(f);
  return ll;
}

int orc_pmmh_run(const orc_model* base, const orc_smc2_cfg* c, int C, const int* N,
                 const double* y, double* theta_out, double* ll_out, int* accepted_out) {
  const int B = c->B, nt = c->ntheta;
  double* lth = (double*)malloc(sizeof(double) * (size_t)B * nt);
  double* lth2 = (double*)malloc(sizeof(double) * (size_t)nt);
  double* ll = (double*)malloc(sizeof(double) * (size_t)B);
  for (int j = 0; j < B; ++j) {
    double z[8];
    orc_normals(c->seed, (uint32_t)j, 0, ORC_TAG_THETA, 0, nt, 32, z);
    for (int i = 0; i < nt; ++i) lth[(size_t)j * nt + i] = c->prior_mean[i] + c->prior_sd[i] * z[i];
    ll[j] = pmmh_loglik(base, c, lth + (size_t)j * nt, N[j], y, (uint32_t)j);
    if (accepted_out) accepted_out[j] = 1;
  }
  for (int it = 0; it < C; ++it) {
    if (it > 0) {
      for (int j = 0; j < B; ++j) {
        double z[8], u;
        orc_normals(c->seed, (uint32_t)j, (uint32_t)it, ORC_TAG_PROPOSAL, 0, nt, 32, z);
        for (int i = 0; i < nt; ++i) lth2[i] = lth[(size_t)j * nt + i] + c->prop_sd[i] * z[i];
        const double ll2 = pmmh_loglik(base, c, lth2, N[j], y, (uint32_t)(B * it + j));
        int accept;
        if (!isfinite(ll2)) accept = 0;
        else if (!isfinite(ll[j])) accept = 1;
        else {
          const double logratio = ll2 - ll[j] + log_prior(c, lth2) - log_prior(c, lth + (size_t)j * nt);
          orc_uniform_at(c->seed, (uint32_t)j, (uint32_t)it, ORC_TAG_ACCEPT, 0, &u);
          accept = log(u) < logratio;
        }
        if (accept) {
          memcpy(lth + (size_t)j * nt, lth2, sizeof(double) * nt);
          ll[j] = ll2;
          if (accepted_out) ++accepted_out[j];
        }
      }
    }
    for (int j = 0; j < B; ++j) {
      if (theta_out)
        for (int i = 0; i < nt; ++i)
          theta_out[((size_t)it * B + j) * nt + i] = exp(lth[(size_t)j * nt + i]);
      if (ll_out) ll_out[(size_t)it * B + j] = ll[j];
    }
  }
  free(lth); free(lth2); free(ll);
  return ORC_OK;
}
