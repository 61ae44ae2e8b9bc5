/*
 * Model functions and ODE integrators of the oracle.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Models (the generated Model<name> of the reference is the semantic spec):
 *  - PZ  (docs/src/users.tex:36-85): alpha ~ N(mu,sigma) drawn once per
 *    transition (share/tt/cpp/action/gaussian.hpp.tt:32-46), then
 *    dP/dt = alpha P - c P Z, dZ/dt = e c P Z - m_l Z - m_q Z^2, and
 *    P_obs ~ log_normal(log P, obs_sd) (share/tt/cpp/action/gaussian.hpp.tt:48-60).
 *  - L96 : dx_i/dt = (x_{i+1} - x_{i-2}) x_{i-1} - x_i + F (cyclic), then
 *    additive noise x += noise_sd * eps per observation interval; obs on
 *    even dimensions, y = x_{2m} + N(0, obs_sd^2).
 *  - LG  : dx/dt = -lambda x, then x += noise_sd * eps; y = x + N(0, obs_sd^2).
 *    Its exact log-likelihood is the Kalman filter's (filter.c).
 *
 * Integrators restate the reference host loops:
 *  - RK4: share/src/bi/host/ode/RK4IntegratorHost.hpp:42-100 with the stage
 *    arithmetic of share/src/bi/ode/RK4Stage.hpp:26-76.
 *  - DOPRI5: share/src/bi/host/ode/DOPRI5IntegratorHost.hpp:42-142 with the
 *    stage arithmetic of share/src/bi/ode/DOPRI5Stage.hpp:26-141 and the
 *    controller constants of share/src/bi/host/ode/IntegratorConstants.cpp:77-87.
 *    host_quirk=1 reproduces the reference host evaluating k7 at x5 (x6 is
 *    never stored before stageErr, DOPRI5IntegratorHost.hpp:100-106);
 *    host_quirk=0 evaluates k7 at x6 (textbook, as the reference GPU kernel
 *    share/src/bi/cuda/ode/DOPRI5KernelGPU.cuh:114-120 does).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define HALF_LOG_TWO_PI 0.91893853320467274178
#define MAXD 256

int orc_nobs(const orc_model* m) {
  switch (m->model) {
  case ORC_MODEL_PZ: return 1;
  case ORC_MODEL_L96: return m->d / 2;
  default: return 1;
  }
}

/* right-hand sides, with the operand order of the model's expressions */
void orc_rhs(const orc_model* m, double alpha, double t, const double* x, double* dx) {
  (void)t;
  const double* par = m->par;
  if (m->model == ORC_MODEL_PZ) {
    const double c = par[2], e = par[3], ml = par[4], mq = par[5];
    const double P = x[0], Z = x[1];
    dx[0] = alpha * P - c * P * Z;
    dx[1] = e * c * P * Z - ml * Z - mq * Z * Z;
  } else if (m->model == ORC_MODEL_L96) {
    const int D = m->d;
    const double F = par[0];
    for (int i = 0; i < D; ++i) {
      const double xp1 = x[(i + 1) % D];
      const double xm2 = x[(i + D - 2) % D];
      const double xm1 = x[(i + D - 1) % D];
      dx[i] = (xp1 - xm2) * xm1 - x[i] + F;
    }
  } else {
    const double lambda = par[0];
    for (int i = 0; i < m->d; ++i) dx[i] = -lambda * x[i];
  }
}

/* RK4IntegratorHost::update for one particle (RK4IntegratorHost.hpp:63-97) */
static int rk4(const orc_model* m, double alpha, double t1, double t2, double* x) {
  const int N = m->d;
  double x1[MAXD], x2[MAXD], x3[MAXD], x4[MAXD], k[MAXD];
  double t = t1, h = m->h0;
  int n = 0;
  while (t < t2) {
    if (t + 1.01 * h - t2 > 0.0) {
      h = t2 - t;
      if (h <= 0.0) { t = t2; break; }
    }
    orc_rhs(m, alpha, t, x, k);                       /* stage1 */
    for (int i = 0; i < N; ++i) {
      x1[i] = 0.5 * k[i];
      x4[i] = (1.0 / 6.0) * k[i];
      x1[i] = h * x1[i] + x[i];
    }
    orc_rhs(m, alpha, t + 0.5 * h, x1, k);            /* stage2 */
    for (int i = 0; i < N; ++i) {
      x2[i] = 0.5 * k[i];
      x4[i] += (1.0 / 3.0) * k[i];
      x2[i] = h * x2[i] + x[i];
    }
    orc_rhs(m, alpha, t + 0.5 * h, x2, k);            /* stage3 */
    for (int i = 0; i < N; ++i) {
      x3[i] = 1.0 * k[i];
      x4[i] += (1.0 / 3.0) * k[i];
      x3[i] = h * x3[i] + x[i];
    }
    orc_rhs(m, alpha, t + h, x3, k);                  /* stage4 */
    for (int i = 0; i < N; ++i) {
      x4[i] += (1.0 / 6.0) * k[i];
      x4[i] = h * x4[i] + x[i];
      x[i] = x4[i];
    }
    t += h;
    ++n;
  }
  return n;
}

/* DOPRI5IntegratorHost::update for one particle (DOPRI5IntegratorHost.hpp:65-139) */
static int dopri5(const orc_model* m, double alpha, double t1, double t2, double* x0) {
  const int N = m->d;
  /* controller constants: IntegratorConstants.cpp:77-87 and :45-61 */
  const double safe = 0.9, facl = 0.2, facr = 10.0, beta = 0.04;
  const double logsafe = log(safe);
  const double expo = 0.5 * (0.2 - beta * 0.75);
  const double atol = m->atol, rtol = m->rtol;
  /* Dormand-Prince coefficients: DOPRI5Stage.hpp:27-136 */
  const double a21 = 0.2, a31 = 3.0 / 40.0, a41 = 44.0 / 45.0, a51 = 19372.0 / 6561.0,
      a61 = 9017.0 / 3168.0, a71 = 35.0 / 384.0, e1 = 71.0 / 57600.0;
  const double c2 = 0.2, a32 = 9.0 / 40.0, a42 = -56.0 / 15.0, a52 = -25360.0 / 2187.0,
      a62 = -355.0 / 33.0;
  const double c3 = 0.3, a43 = 32.0 / 9.0, a53 = 64448.0 / 6561.0, a63 = 46732.0 / 5247.0,
      a73 = 500.0 / 1113.0, e3 = -71.0 / 16695.0;
  const double c4 = 0.8, a54 = -212.0 / 729.0, a64 = 49.0 / 176.0, a74 = 125.0 / 192.0,
      e4 = 71.0 / 1920.0;
  const double c5 = 8.0 / 9.0, a65 = -5103.0 / 18656.0, a75 = -2187.0 / 6784.0,
      e5 = -17253.0 / 339200.0;
  const double a76 = 11.0 / 84.0, e6 = 22.0 / 525.0;
  const double e7 = -1.0 / 40.0;

  double x1[MAXD], x2[MAXD], x3[MAXD], x4[MAXD], x5[MAXD], x6[MAXD], err[MAXD];
  double k1[MAXD], k[MAXD], k7[MAXD];
  double t = t1, h = m->h0, logfacold = log(1.0e-4), e2;
  int k1in = 0, n = 0;

  while (t < t2 && n < m->nsteps) {
    if (t + 1.01 * h - t2 > 0.0) {
      h = t2 - t;
      if (h <= 0.0) { t = t2; break; }
    }
    /* stage1 */
    if (!k1in) orc_rhs(m, alpha, t, x0, k1);
    for (int i = 0; i < N; ++i) {
      x1[i] = a21 * k1[i]; x2[i] = a31 * k1[i]; x3[i] = a41 * k1[i];
      x4[i] = a51 * k1[i]; x5[i] = a61 * k1[i]; x6[i] = a71 * k1[i];
      err[i] = e1 * k1[i];
      x1[i] = h * x1[i] + x0[i];
    }
    k1in = 1;
    /* stage2 */
    orc_rhs(m, alpha, t + c2 * h, x1, k);
    for (int i = 0; i < N; ++i) {
      x2[i] += a32 * k[i]; x3[i] += a42 * k[i]; x4[i] += a52 * k[i]; x5[i] += a62 * k[i];
      x2[i] = h * x2[i] + x0[i];
    }
    /* stage3 */
    orc_rhs(m, alpha, t + c3 * h, x2, k);
    for (int i = 0; i < N; ++i) {
      x3[i] += a43 * k[i]; x4[i] += a53 * k[i]; x5[i] += a63 * k[i]; x6[i] += a73 * k[i];
      err[i] += e3 * k[i];
      x3[i] = h * x3[i] + x0[i];
    }
    /* stage4 */
    orc_rhs(m, alpha, t + c4 * h, x3, k);
    for (int i = 0; i < N; ++i) {
      x4[i] += a54 * k[i]; x5[i] += a64 * k[i]; x6[i] += a74 * k[i]; err[i] += e4 * k[i];
      x4[i] = h * x4[i] + x0[i];
    }
    /* stage5 */
    orc_rhs(m, alpha, t + c5 * h, x4, k);
    for (int i = 0; i < N; ++i) {
      x5[i] += a65 * k[i]; x6[i] += a75 * k[i]; err[i] += e5 * k[i];
      x5[i] = h * x5[i] + x0[i];
    }
    /* stage6 */
    orc_rhs(m, alpha, t + h, x5, k);
    for (int i = 0; i < N; ++i) {
      x6[i] += a76 * k[i]; err[i] += e6 * k[i];
      x6[i] = h * x6[i] + x0[i];
    }
    /* stageErr: k7 at x6 (textbook) or at x5 (reference host) */
    orc_rhs(m, alpha, t + h, m->host_quirk ? x5 : x6, k7);
    for (int i = 0; i < N; ++i) err[i] += e7 * k7[i];

    e2 = 0.0;
    for (int i = 0; i < N; ++i) {
      double e = err[i] * h / (atol + rtol * fmax(fabs(x0[i]), fabs(x6[i])));
      e2 += e * e;
    }
    e2 /= N;

    if (e2 <= 1.0) {
      t += h;
      memcpy(x0, x6, sizeof(double) * N);
      memcpy(k1, k7, sizeof(double) * N);
    }
    if (t < t2) {
      double logfac11 = expo * log(e2);
      if (e2 > 1.0) {
        h *= fmax(facl, exp(logsafe - logfac11));
      } else {
        double fac = exp(beta * logfacold + logsafe - logfac11);
        fac = fmin(facr, fmax(facl, fac));
        h *= fac;
        logfacold = 0.5 * log(fmax(e2, 1.0e-8));
      }
    }
    ++n;
  }
  return n;
}

int orc_flow(const orc_model* m, double alpha, double t1, double t2, double* x) {
  if (!(fabs(t2 - t1) > 0.0)) return 0;  /* DOPRI5Integrator.hpp:51 */
  return m->integrator == ORC_DOPRI5 ? dopri5(m, alpha, t1, t2, x) : rk4(m, alpha, t1, t2, x);
}

void orc_init_samples(const orc_model* m, uint64_t seed, uint32_t batch, int P, double* X,
    int ld, const double* mean0) {
  const int D = m->d;
#pragma omp parallel for schedule(static)
  for (int p = 0; p < P; ++p) {
    double z[MAXD];
    orc_normals(seed, (uint32_t)p, 0, ORC_TAG_INIT, batch, D, m->noise_bits, z);
    if (m->model == ORC_MODEL_PZ) {
      X[p] = exp(m->par[7] + m->par[8] * z[0]);
      X[ld + p] = exp(m->par[9] + m->par[10] * z[1]);
    } else if (m->model == ORC_MODEL_L96) {
      for (int j = 0; j < D; ++j) X[(size_t)j * ld + p] = mean0[j] + m->par[3] * z[j];
    } else {
      for (int j = 0; j < D; ++j) X[(size_t)j * ld + p] = m->par[3] + m->par[4] * z[j];
    }
  }
}

/*
 * Model::transitionSamples for all particles: the noise pdf_ block, then the
 * ode block (share/tt/cpp/block/transition.hpp.tt:40-44, block/ode.hpp.tt:39-54),
 * OpenMP over particles as StaticSamplerHost::samples
 * (share/src/bi/host/updater/StaticSamplerHost.hpp:43-65) and the integrators do.
 */
void orc_transition(const orc_model* m, uint64_t seed, uint32_t batch, uint32_t step,
    double t1, double t2, int P, double* X, int ld, int* nsteps_out) {
  const int D = m->d;
#pragma omp parallel for schedule(static)
  for (int p = 0; p < P; ++p) {
    double x[MAXD], z[MAXD];
    int n;
    for (int j = 0; j < D; ++j) x[j] = X[(size_t)j * ld + p];
    if (m->model == ORC_MODEL_PZ) {
      orc_normals(seed, (uint32_t)p, step, ORC_TAG_NOISE, batch, 1, m->noise_bits, z);
      double alpha = m->par[0] + m->par[1] * z[0];
      n = orc_flow(m, alpha, t1, t2, x);
    } else {
      n = orc_flow(m, 0.0, t1, t2, x);
      orc_normals(seed, (uint32_t)p, step, ORC_TAG_NOISE, batch, D, m->noise_bits, z);
      for (int j = 0; j < D; ++j) x[j] = x[j] + m->par[1] * z[j];
    }
    for (int j = 0; j < D; ++j) X[(size_t)j * ld + p] = x[j];
    if (nsteps_out) nsteps_out[p] = n;
  }
}

/*
 * Model::observationLogDensities: lws[p] += sum over observed variables of
 * the Gaussian / log-normal log-density (share/tt/cpp/action/gaussian.hpp.tt:48-70),
 * OpenMP over particles as SparseStaticLogDensityHost
 * (share/src/bi/host/updater/SparseStaticLogDensityHost.hpp:46-68).
 */
void orc_obs_logdensity(const orc_model* m, const double* y, int P, const double* X, int ld,
    double* lws) {
#pragma omp parallel for schedule(static)
  for (int p = 0; p < P; ++p) {
    double lp = lws[p];
    if (m->model == ORC_MODEL_PZ) {
      const double sd = m->par[6];
      const double mu = log(X[p]);
      const double z = (log(y[0]) - mu) / sd;
      lp += -0.5 * (z * z) - HALF_LOG_TWO_PI - log(sd * y[0]);
    } else if (m->model == ORC_MODEL_L96) {
      const double sd = m->par[2];
      const int nobs = m->d / 2;
      for (int o = 0; o < nobs; ++o) {
        const double z = (y[o] - X[(size_t)(2 * o) * ld + p]) / sd;
        lp += -0.5 * (z * z) - HALF_LOG_TWO_PI - log(sd);
      }
    } else {
      const double sd = m->par[2];
      const double z = (y[0] - X[p]) / sd;
      lp += -0.5 * (z * z) - HALF_LOG_TWO_PI - log(sd);
    }
    lws[p] = lp;
  }
}

/* Model::observationSamples for one particle (gaussian.hpp.tt:32-46) */
void orc_obs_sample(const orc_model* m, uint64_t seed, uint32_t batch, uint32_t step,
    const double* x, double* y) {
  double z[MAXD];
  const int nobs = orc_nobs(m);
  orc_normals(seed, 0, step, ORC_TAG_OBS, batch, nobs, m->noise_bits, z);
  if (m->model == ORC_MODEL_PZ) {
    y[0] = exp(log(x[0]) + m->par[6] * z[0]);
  } else if (m->model == ORC_MODEL_L96) {
    for (int o = 0; o < nobs; ++o) y[o] = x[2 * o] + m->par[2] * z[o];
  } else {
    y[0] = x[0] + m->par[2] * z[0];
  }
}
