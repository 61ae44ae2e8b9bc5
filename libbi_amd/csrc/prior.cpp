// Priors of the parameter blocks of generated models, on the host.
//
// LibBi's samplers draw theta from the model's parameter block
// (Model::parameterSamples, share/tt/cpp/model.hpp.tt) and evaluate its
// log-density in the Metropolis-Hastings ratio (MarginalMH.hpp:217-266,
// MarginalSIR.hpp:395-488).  Here the block's priors are a table written by
// the model generator (one {distribution, a, b} per param element, constant
// arguments); draws are keyed Philox words as every other draw of the
// library: parameter i of theta-particle j takes the words of the stream
// {j, 0, THETA, 0} in order from block 2^20 (i + 1) (the gamma sampler's
// rejected attempts move on along them).
#include <cmath>
#include <cstring>

#include "internal.hpp"
#include "generated_priors.hpp"

namespace bihip {
namespace {

constexpr double HALF_LOG_TWO_PI_H = 0.91893853320467274178;
constexpr double TWO_PI_H = 6.283185307179586476925286766559;

const GenPrior* prior_table(const bihip_model* m, int* count) {
  const int i = m->model - BIHIP_MODEL_GENERATED;
  if (i < 0 || i >= GEN_PRIOR_MODELS) return nullptr;
  *count = GEN_PRIOR_COUNT[i];
  return GEN_PRIORS[i];
}

struct Words {
  uint64_t seed;
  uint32_t index;
  uint32_t block;
  uint32_t r[4];
  int used = 4;
  uint32_t next() {
    if (used == 4) {
      philox_host_block(seed, index, 0u, BIHIP_TAG_THETA, 0u, block++, r);
      used = 0;
    }
    return r[used++];
  }
  // u in (0, 1]
  double open_uniform() { return 1.0 - (double)next() * 0x1.0p-32; }
  double uniform() {  // [0, 1) from 53 bits
    const uint64_t hi = next(), lo = next();
    return (double)(((hi << 32) | lo) >> 11) * 0x1.0p-53;
  }
  double normal() {  // Box-Muller, one of the pair
    const double u1 = open_uniform();
    const double u2 = (double)next() * 0x1.0p-32;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(TWO_PI_H * u2);
  }
};

// Marsaglia and Tsang (2000), "A simple method for generating gamma
// variables": shape k >= 1 directly, k < 1 as G(k + 1) U^(1/k); scale 1.
double gamma_unit(double k, Words& w) {
  if (k < 1.0) {
    const double g = gamma_unit(k + 1.0, w);
    return g * std::pow(w.open_uniform(), 1.0 / k);
  }
  const double d = k - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
  for (int t = 0; t < 1 << 16; ++t) {
    const double z = w.normal();
    const double v0 = 1.0 + c * z;
    if (v0 <= 0.0) continue;
    const double v = v0 * v0 * v0;
    const double u = w.open_uniform();
    if (std::log(u) < 0.5 * z * z + d - d * v + d * std::log(v)) return d * v;
  }
  return d;  // not reached for any shape in practice (acceptance above 95%)
}

double sample_one(const GenPrior& p, Words& w) {
  switch (p.dist) {
    case 0: return p.a + p.b * w.normal();
    case 1: return std::exp(p.a + p.b * w.normal());
    case 2: return p.a + (p.b - p.a) * w.uniform();
    case 3: return -std::log(1.0 - w.uniform()) / p.a;
    case 4: return p.b * gamma_unit(p.a, w);
    case 5: return p.b / gamma_unit(p.a, w);
    case 6: {
      const double x = gamma_unit(p.a, w), y = gamma_unit(p.b, w);
      return x / (x + y);
    }
  }
  return std::nan("");
}

// log-densities as share/src/bi/pdf/functor.hpp and the action templates write them
double logdensity_one(const GenPrior& p, double x) {
  const double inf = INFINITY;
  switch (p.dist) {
    case 0: {
      const double z = (x - p.a) / p.b;
      return -0.5 * (z * z) - HALF_LOG_TWO_PI_H - std::log(p.b);
    }
    case 1: {
      if (!(x > 0.0)) return -inf;
      const double z = (std::log(x) - p.a) / p.b;
      return -0.5 * (z * z) - HALF_LOG_TWO_PI_H - std::log(p.b * x);
    }
    case 2: return x >= p.a && x <= p.b ? -std::log(p.b - p.a) : -inf;
    case 3: return x >= 0.0 ? std::log(p.a) - p.a * x : -inf;
    case 4:
      if (!(x > 0.0)) return -inf;
      return (p.a - 1.0) * std::log(x) - x / p.b - (std::lgamma(p.a) + p.a * std::log(p.b));
    case 5: {
      if (!(x > 0.0)) return -inf;
      const double sc = 1.0 / p.b, ix = 1.0 / x;
      return (p.a - 1.0) * std::log(ix) - ix / sc - (std::lgamma(p.a) + p.a * std::log(sc)) -
             2.0 * std::log(x);
    }
    case 6:
      if (!(x > 0.0 && x < 1.0)) return -inf;
      return (p.a - 1.0) * std::log(x) + (p.b - 1.0) * std::log(1.0 - x) -
             (std::lgamma(p.a) + std::lgamma(p.b) - std::lgamma(p.a + p.b));
  }
  return -inf;
}

}  // namespace

int prior_check(const bihip_model* m, int n, bool positive) {
  int count = 0;
  const GenPrior* t = prior_table(m, &count);
  if (t == nullptr) return set_error(BIHIP_EINVAL, "the model is not generated from a model file");
  if (n > count) return set_error(BIHIP_EINVAL, "the model has %d param elements, not %d", count, n);
  for (int i = 0; i < n; ++i) {
    if (t[i].dist < 0)
      return set_error(BIHIP_EINVAL, "param element %d has no prior in the parameter block", i);
    const bool pos = t[i].dist == 1 || t[i].dist == 3 || t[i].dist == 4 || t[i].dist == 5 ||
                     t[i].dist == 6 || (t[i].dist == 2 && t[i].a >= 0.0);
    if (positive && !pos)
      return set_error(BIHIP_EINVAL,
                       "param element %d: its prior puts mass on values <= 0, and the sampler "
                       "moves in log space", i);
  }
  return BIHIP_OK;
}

void prior_sample(const bihip_model* m, int n, uint64_t seed, uint32_t index, double* theta) {
  int count = 0;
  const GenPrior* t = prior_table(m, &count);
  for (int i = 0; i < n; ++i) {
    Words w{seed, index, (uint32_t)(i + 1) << 20};
    theta[i] = t != nullptr && i < count ? sample_one(t[i], w) : std::nan("");
  }
}

double prior_logdensity(const bihip_model* m, int n, const double* theta) {
  int count = 0;
  const GenPrior* t = prior_table(m, &count);
  double lp = 0.0;
  for (int i = 0; i < n; ++i) lp += t != nullptr && i < count ? logdensity_one(t[i], theta[i])
                                                               : -INFINITY;
  return lp;
}

}  // namespace bihip

using namespace bihip;

extern "C" {

int bihip_model_prior_sample(const bihip_model* m, uint64_t seed, uint32_t index, int n,
                             double* theta_host) {
  if (m == nullptr || theta_host == nullptr) return set_error(BIHIP_EINVAL, "NULL argument");
  if (int rc = prior_check(m, n, false)) return rc;
  prior_sample(m, n, seed, index, theta_host);
  return BIHIP_OK;
}

int bihip_model_prior_logdensity(const bihip_model* m, int n, const double* theta_host,
                                 double* lp) {
  if (m == nullptr || theta_host == nullptr || lp == nullptr)
    return set_error(BIHIP_EINVAL, "NULL argument");
  if (int rc = prior_check(m, n, false)) return rc;
  *lp = prior_logdensity(m, n, theta_host);
  return BIHIP_OK;
}

}  // extern "C"
