/*
 * The learned guided proposal's step, restated in FP64.  TEST INFRASTRUCTURE
 * ONLY (see oracle.h).
 *
 * The proposal has no counterpart in the reference; its weight correction
 * follows the auxiliary-weight pattern of BridgePF / LookaheadPF
 * (share/src/bi/filter/BridgePF.hpp:114-144, LookaheadPF.hpp:91-124): the
 * log-weight gains log p(y | x') + log p(x' | x) - log q(x' | x, y).
 *
 * For each particle with flowed state Phi (the exact flow of its ancestor, d
 * components; observations y of the even components):
 *   u = ((Phi - x_shift) / x_scale, (y - H Phi) / innov_scale,
 *        (F - par_shift) / par_scale)                   (d + d/2 + 1 inputs)
 *   (m, s) = MLP(u)                                      (2 d outputs)
 *   x' = Phi + q (m + e^{s/2} eps),  eps = the transition-noise normals
 *   inc = sum_o log N(y_o; x'_{2o}, obs_sd^2)
 *         + sum_i [ -(m_i + e^{s_i/2} eps_i)^2 / 2 + eps_i^2 / 2 + s_i / 2 ],
 * the last sum being log N(x'; Phi, q^2 I) - log N(x'; Phi + q m, q^2 e^s).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

static const double HALF_LOG_TWO_PI = 0.91893853320467274178;

int orc_guided_step(int L, const int* dims, const double* params, int bf16,
                    const double* norm, double F, int d, int P, const double* Phi,
                    const double* y, uint64_t seed, uint32_t batch, uint32_t step, double q,
                    double obs_sd, int noise_bits, double* Xout, double* inc) {
  const int nobs = d / 2, nin = d + nobs + 1;
  if (d % 2 != 0 || dims[0] != nin || dims[L] != 2 * d) return ORC_EINVAL;
  double* U = (double*)malloc(sizeof(double) * (size_t)P * nin);
  double* O = (double*)malloc(sizeof(double) * (size_t)P * 2 * d);
  for (int p = 0; p < P; ++p) {
    double* u = U + (size_t)p * nin;
    for (int i = 0; i < d; ++i) u[i] = (Phi[(size_t)i * P + p] - norm[0]) / norm[1];
    for (int o = 0; o < nobs; ++o) u[d + o] = (y[o] - Phi[(size_t)(2 * o) * P + p]) / norm[2];
    u[d + nobs] = (F - norm[3]) / norm[4];
  }
  orc_mlp_forward(L, dims, params, U, P, bf16, O);
  const double c_obs = -HALF_LOG_TWO_PI - log(obs_sd);
#pragma omp parallel for schedule(static)
  for (int p = 0; p < P; ++p) {
    double* z = (double*)malloc(sizeof(double) * (size_t)d);
    orc_normals(seed, (uint32_t)p, step, ORC_TAG_NOISE, batch, d, noise_bits, z);
    const double* o = O + (size_t)p * 2 * d;
    double w = 0.0;
    for (int i = 0; i < d; ++i) {
      const double m = o[i], s = o[d + i];
      const double dd = m + exp(0.5 * s) * z[i];
      const double x = Phi[(size_t)i * P + p] + q * dd;
      Xout[(size_t)i * P + p] = x;
      w += -0.5 * dd * dd + 0.5 * z[i] * z[i] + 0.5 * s;
      if (i % 2 == 0) {
        const double e = (y[i / 2] - x) / obs_sd;
        w += -0.5 * (e * e) + c_obs;
      }
    }
    inc[p] = w;
    free(z);
  }
  free(U);
  free(O);
  return ORC_OK;
}
