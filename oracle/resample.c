/*
 * Reductions, scans and resamplers of the oracle.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference's thrust calls run on thrust's sequential host backend
 * (no THRUST_HOST_SYSTEM is set, share/tt/build/Makefile.am.tt:21-27), so
 * reductions and scans here are left-to-right folds.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* nanexp: NaN gives zero (share/src/bi/math/function.hpp:158-160) */
static double nanexp(double x) { return isnan(x) ? 0.0 : exp(x); }

/* max_element with nan_less_functor: NaN below everything
 * (share/src/bi/primitive/vector_primitive.hpp:860-872, functor.hpp:43-53) */
double orc_max_reduce(const double* x, int n) {
  double mx = x[0];
  for (int i = 1; i < n; ++i) {
    int less;
    if (isnan(mx)) less = !isnan(x[i]);
    else if (isnan(x[i])) less = 0;
    else less = mx < x[i];
    if (less) mx = x[i];
  }
  return mx;
}

/* logsumexp_reduce (vector_primitive.hpp:932-943) */
double orc_logsumexp_reduce(const double* x, int n) {
  const double mx = orc_max_reduce(x, n);
  double s = 0.0;
  for (int i = 0; i < n; ++i) s = s + nanexp(x[i] - mx);
  return mx + log(s);
}

/* ess_reduce (vector_primitive.hpp:959-974, functors functor.hpp:634-658) */
double orc_ess_reduce(const double* lws, int n, double* lW) {
  const double mx = orc_max_reduce(lws, n);
  double s1 = 0.0, s2 = 0.0;
  for (int i = 0; i < n; ++i) {
    const double z = nanexp(lws[i] - mx);
    s1 = s1 + z;
    s2 = s2 + z * z;
  }
  if (lW) *lW = mx + log(s1) - log((double)n);
  return s1 * s1 / s2;
}

/* sumexpu_inclusive_scan (vector_primitive.hpp:1075-1084); returns the max */
double orc_sumexpu_inclusive_scan(const double* lws, int n, double* Ws) {
  const double mx = orc_max_reduce(lws, n);
  double s = 0.0;
  for (int i = 0; i < n; ++i) {
    s = s + nanexp(lws[i] - mx);
    Ws[i] = s;
  }
  return mx;
}

/*
 * SystematicResampler::cumulativeOffspring with W = Ws[P-1]
 * (share/src/bi/resampler/SystematicResampler.hpp:110-130 and the functor
 * resample_cumulative_offspring, :47-50): O_i = min(n, (int)(Ws_i / W * n + a)).
 */
int orc_systematic_cumulative_offspring(const double* Ws, int P, int n, double a, int* Os) {
  const double W = Ws[P - 1];
  if (!(W > 0.0)) return ORC_DEGENERATE;
  for (int i = 0; i < P; ++i) {
    const double v = Ws[i] / W * (double)n + a;
    const int o = (int)v;
    Os[i] = o < n ? o : n;
  }
  return ORC_OK;
}

/*
 * StratifiedResamplerHost::op (share/src/bi/host/resampler/StratifiedResamplerHost.hpp:25-74),
 * guarded by StratifiedResampler::cumulativeOffspring (StratifiedResampler.hpp:83-106).
 */
int orc_stratified_cumulative_offspring(const double* Ws, int P, int n, const double* alphas,
    int* Os) {
  const double W = Ws[P - 1];
  if (!(W > 0.0)) return ORC_DEGENERATE;
  for (int i = 0; i < P; ++i) {
    const double reach = Ws[i] / W * (double)n;
    int k = (int)reach;
    if (k > n - 1) k = n - 1;
    const int o = (int)(reach + alphas[k]);
    Os[i] = o < n ? o : n;
  }
  return ORC_OK;
}

/*
 * Multinomial ancestors: a_j = #{i : Ws_i <= W U_j}, the upper-bound search
 * that the reference host performs in log space over sorted uniforms
 * (share/src/bi/host/resampler/MultinomialResamplerHost.hpp:47-57), here on
 * iid uniforms U_j in [0,1).  The clamp to P-1 is inactive for U_j < 1.
 */
int orc_multinomial_ancestors(const double* Ws, int P, int n, const double* U, int* as) {
  const double W = Ws[P - 1];
  if (!(W > 0.0)) return ORC_DEGENERATE;
  for (int j = 0; j < n; ++j) {
    const double u = W * U[j];
    int lo = 0, hi = P; /* first index with Ws[i] > u */
    while (lo < hi) {
      const int mid = lo + (hi - lo) / 2;
      if (Ws[mid] <= u) lo = mid + 1; else hi = mid;
    }
    as[j] = lo < P ? lo : P - 1;
  }
  return ORC_OK;
}

/* ResamplerHost::ancestorsToOffspring (share/src/bi/host/resampler/ResamplerHost.hpp:55-68) */
void orc_ancestors_to_offspring(const int* as, int n, int P, int* os) {
  memset(os, 0, sizeof(int) * (size_t)P);
  for (int j = 0; j < n; ++j) ++os[as[j]];
}

void orc_offspring_to_cumulative(const int* os, int P, int* Os) {
  int s = 0;
  for (int i = 0; i < P; ++i) { s += os[i]; Os[i] = s; }
}

/* ResamplerHost::cumulativeOffspringToAncestors (ResamplerHost.hpp:109-131) */
void orc_cumulative_offspring_to_ancestors(const int* Os, int P, int* as) {
  for (int i = 0; i < P; ++i) {
    const int O1 = i > 0 ? Os[i - 1] : 0;
    for (int j = O1; j < Os[i]; ++j) as[j] = i;
  }
}

/*
 * ResamplerHost::cumulativeOffspringToAncestorsPermute (ResamplerHost.hpp:133-168):
 * a particle with offspring keeps its own slot; its remaining copies, taken
 * in ascending particle order, fill the slots of particles without offspring
 * in ascending slot order.  Written as a two-pointer walk.
 */
void orc_cumulative_offspring_to_ancestors_permute(const int* Os, int P, int* as) {
  int k = 0; /* next candidate free slot */
  for (int i = 0; i < P; ++i) {
    int o = Os[i] - (i > 0 ? Os[i - 1] : 0);
    if (o > 0) { as[i] = i; --o; }
    while (o > 0) {
      while (Os[k] - (k > 0 ? Os[k - 1] : 0) > 0) ++k;
      as[k++] = i;
      --o;
    }
  }
}

/* ResamplerHost::offspringToAncestorsPermute (ResamplerHost.hpp:86-107) */
void orc_offspring_to_ancestors_permute(const int* os, int P, int* as) {
  int k = 0;
  for (int i = 0; i < P; ++i) {
    int o = os[i];
    if (o > 0) { as[i] = i; --o; }
    while (o > 0) {
      while (os[k] > 0) ++k;
      as[k++] = i;
      --o;
    }
  }
}

/* ResamplerHost::permute, the in-place swap form (ResamplerHost.hpp:170-190) */
void orc_permute(int* as, int P) {
  for (int i = 0; i < P; ++i) {
    const int k = as[i];
    if (k < P && k != i && as[k] != k) {
      const int j = as[k];
      as[k] = k;
      as[i] = j;
      --i;
    }
  }
}

/* gather_rows in place (share/src/bi/host/primitive/matrix_primitive.hpp:65-75,
 * share/src/bi/state/State.hpp:905-909); needs as[as[i]] == as[i]. */
void orc_gather_rows(const int* as, int P, int ncols, double* X, int ld) {
  for (int j = 0; j < ncols; ++j) {
    double* col = X + (size_t)j * ld;
    for (int i = 0; i < P; ++i) col[i] = col[as[i]];
  }
}
