/*
 * Bootstrap particle filter driver of the oracle and the Kalman filter
 * that pins its log-likelihood conventions.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * orc_pf_run restates Filter<F>::filter (share/src/bi/filter/Filter.hpp:56-71)
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

int orc_pf_run(const orc_model* m, const orc_pf_cfg* cfg, const double* y,
    const double* mean0, double* logZ, double* logIncrements, double* essOut,
    int* resampledOut, int* ancestorsOut, double* Xout) {
  const int P = cfg->P, T = cfg->T, D = m->d, nobs = orc_nobs(m);
  double* X = (double*)malloc(sizeof(double) * (size_t)D * P);
  double* lws = (double*)calloc((size_t)P, sizeof(double));
  double* Ws = (double*)malloc(sizeof(double) * (size_t)P);
  double* U = (double*)malloc(sizeof(double) * (size_t)P);
  int* Os = (int*)malloc(sizeof(int) * (size_t)P);
  int* os = (int*)malloc(sizeof(int) * (size_t)P);
  int* as = (int*)malloc(sizeof(int) * (size_t)P);
  int* a0 = (int*)malloc(sizeof(int) * (size_t)P);
  int rc = ORC_OK;
  double logLikelihood = 0.0, ess = 0.0;

  orc_init_samples(m, cfg->seed, cfg->batch, P, X, P, mean0);

  for (int k = 0; k < T; ++k) {
    const double t1 = cfg->t0 + k * cfg->dt, t2 = cfg->t0 + (k + 1) * cfg->dt;
    const int r = k >= 1 && ess < cfg->essRel * P;
    if (r) {
      orc_sumexpu_inclusive_scan(lws, P, Ws);
      if (cfg->resampler == ORC_SYSTEMATIC) {
        double a;
        orc_uniforms(cfg->seed, (uint32_t)k, ORC_TAG_SYSTEMATIC, cfg->batch, 1, &a);
        rc = orc_systematic_cumulative_offspring(Ws, P, P, a, Os);
      } else if (cfg->resampler == ORC_STRATIFIED) {
        orc_uniforms(cfg->seed, (uint32_t)k, ORC_TAG_STRATIFIED, cfg->batch, P, U);
        rc = orc_stratified_cumulative_offspring(Ws, P, P, U, Os);
      } else {
        orc_uniforms(cfg->seed, (uint32_t)k, ORC_TAG_MULTINOMIAL, cfg->batch, P, U);
        rc = orc_multinomial_ancestors(Ws, P, P, U, a0);
        if (rc == ORC_OK) {
          orc_ancestors_to_offspring(a0, P, P, os);
          orc_offspring_to_cumulative(os, P, Os);
        }
      }
      if (rc != ORC_OK) break;
      orc_cumulative_offspring_to_ancestors_permute(Os, P, as);
      orc_gather_rows(as, P, D, X, P);
      for (int i = 0; i < P; ++i) lws[i] = logLikelihood;
    } else {
      for (int i = 0; i < P; ++i) as[i] = i;
    }
    if (ancestorsOut) memcpy(ancestorsOut + (size_t)k * P, as, sizeof(int) * (size_t)P);
    if (resampledOut) resampledOut[k] = r;

    orc_transition(m, cfg->seed, cfg->batch, (uint32_t)(k + 1), t1, t2, P, X, P, NULL);
    orc_obs_logdensity(m, y + (size_t)k * nobs, P, X, P, lws);
    double lW;
    ess = orc_ess_reduce(lws, P, &lW);
    if (essOut) essOut[k] = ess;
    if (logIncrements) logIncrements[k] = lW - logLikelihood;
    logLikelihood = lW;
  }
  if (rc == ORC_OK && logZ) *logZ = orc_logsumexp_reduce(lws, P) - log((double)P);
  if (Xout) memcpy(Xout, X, sizeof(double) * (size_t)D * P);

  free(X); free(lws); free(Ws); free(U); free(Os); free(os); free(as); free(a0);
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
