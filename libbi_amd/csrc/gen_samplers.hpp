// Samplers of the generated models' draws beyond gaussian, uniform and
// exponential (share/tt/cpp/action/{gamma,inverse_gamma,beta,poisson,negbin,
// binomial}.hpp.tt `sample`): gamma by Marsaglia and Tsang (2000), beta and
// inverse gamma from gammas, Poisson by multiplication of uniforms below a
// mean of 10 and by Hormann's transformed rejection (PTRS, 1993) above,
// binomial by inversion below a mean of 10 and by Hormann's BTRS (1993)
// above, negative binomial as a Poisson of a gamma.  Every draw reads its
// words in order from its own keyed Philox stream (its first block given by
// the caller), so a rejection consumes words of that draw only; each
// rejection loop is bounded (256 attempts, far beyond what any of these
// methods needs at their acceptance rates above 0.85); a draw that reaches
// the bound returns NaN, so the failure reaches the weights (a NaN weight is
// zero) instead of biasing the draw.  Arithmetic is double.
//
// Needs u32x4 and philox_block(seed, index, step, tag, batch, block).
#pragma once

namespace bihip {

struct DrawStream {
  uint64_t seed;
  uint32_t index, step, tag, batch, block;
  uint32_t r0, r1, r2, r3;
  int used;
  __device__ __forceinline__ DrawStream(uint64_t s, uint32_t i, uint32_t st, uint32_t tg,
                                        uint32_t b, uint32_t blk)
      : seed(s), index(i), step(st), tag(tg), batch(b), block(blk), r0(0), r1(0), r2(0), r3(0),
        used(4) {}
  __device__ __forceinline__ uint32_t word() {
    if (used == 4) {
      const u32x4 q = philox_block(seed, index, step, tag, batch, block++);
      r0 = q.x;
      r1 = q.y;
      r2 = q.z;
      r3 = q.w;
      used = 0;
    }
    const uint32_t v = used == 0 ? r0 : (used == 1 ? r1 : (used == 2 ? r2 : r3));
    ++used;
    return v;
  }
  // [0, 1) from 53 bits
  __device__ __forceinline__ double uniform() {
    const uint64_t hi = word(), lo = word();
    return (double)(((hi << 32) | lo) >> 11) * 0x1.0p-53;
  }
  // (0, 1]
  __device__ __forceinline__ double open_uniform() { return 1.0 - (double)word() * 0x1.0p-32; }
  __device__ __forceinline__ double normal() {
    const double u1 = open_uniform();
    const double u2 = (double)word() * 0x1.0p-32;
    return sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
  }
};

constexpr int GEN_MAX_ATTEMPTS = 256;

// gamma(shape k, scale 1)
__device__ __forceinline__ double gen_gamma_unit(double k, DrawStream& w) {
  double boost = 1.0;
  if (k < 1.0) {  // G(k) = G(k + 1) U^(1/k)
    boost = pow(w.open_uniform(), 1.0 / k);
    k += 1.0;
  }
  const double d = k - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
  double x = __builtin_nan("");
  for (int t = 0; t < GEN_MAX_ATTEMPTS; ++t) {
    const double z = w.normal();
    const double v0 = 1.0 + c * z;
    if (v0 <= 0.0) continue;
    const double v = v0 * v0 * v0;
    const double u = w.open_uniform();
    if (log(u) < 0.5 * z * z + d - d * v + d * log(v)) {
      x = d * v;
      break;
    }
  }
  return x * boost;
}

__device__ __forceinline__ double gen_gamma(double shape, double scale, DrawStream w) {
  return scale * gen_gamma_unit(shape, w);
}

__device__ __forceinline__ double gen_inverse_gamma(double shape, double scale, DrawStream w) {
  return scale / gen_gamma_unit(shape, w);
}

__device__ __forceinline__ double gen_beta(double a, double b, DrawStream w) {
  const double x = gen_gamma_unit(a, w), y = gen_gamma_unit(b, w);
  return x / (x + y);
}

__device__ __forceinline__ double gen_poisson_stream(double lambda, DrawStream& w) {
  if (!(lambda > 0.0)) return 0.0;
  if (lambda < 10.0) {  // multiply uniforms until the product falls below e^-lambda
    const double L = exp(-lambda);
    double p = w.open_uniform();
    int k = 0;
    while (p > L && k < 1000) {
      p *= w.open_uniform();
      ++k;
    }
    return (double)k;
  }
  // PTRS (W. Hormann, "The transformed rejection method for generating Poisson
  // random variables", Insurance: Mathematics and Economics 12, 1993)
  const double slam = sqrt(lambda), loglam = log(lambda);
  const double b = 0.931 + 2.53 * slam, a = -0.059 + 0.02483 * b;
  const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2.0);
  double k = __builtin_nan("");
  for (int t = 0; t < GEN_MAX_ATTEMPTS; ++t) {
    const double U = w.uniform() - 0.5, V = w.uniform();
    const double us = 0.5 - fabs(U);
    const double kk = floor((2.0 * a / us + b) * U + lambda + 0.43);
    if (us >= 0.07 && V <= vr) {
      k = kk;
      break;
    }
    if (kk < 0.0 || (us < 0.013 && V > us)) continue;
    if (log(V) + log(inv_alpha) - log(a / (us * us) + b) <=
        -lambda + kk * loglam - lgamma(kk + 1.0)) {
      k = kk;
      break;
    }
  }
  return k;
}

__device__ __forceinline__ double gen_poisson(double lambda, DrawStream w) {
  return gen_poisson_stream(lambda, w);
}

// negative binomial of mean mu and shape k: Poisson(Gamma(k, mu / k))
__device__ __forceinline__ double gen_negbin(double mu, double k, DrawStream w) {
  if (!(k > 0.0)) return gen_poisson_stream(mu, w);
  const double lambda = (mu / k) * gen_gamma_unit(k, w);
  return gen_poisson_stream(lambda, w);
}

__device__ __forceinline__ double gen_binomial(double n, double p, DrawStream w) {
  n = floor(n);
  if (!(n > 0.0) || !(p > 0.0)) return 0.0;
  if (p >= 1.0) return n;
  const bool flip = p > 0.5;
  const double pp = flip ? 1.0 - p : p, q = 1.0 - pp;
  double x;
  if (n * pp < 10.0) {  // inversion: walk the probabilities up from x = 0
    const double s = pp / q, a = (n + 1.0) * s;
    double r = pow(q, n), u = w.uniform();
    x = 0.0;
    // the probabilities fall faster than geometrically past the mean, so r
    // underflows within a few hundred terms: the walk ends there even when
    // rounding leaves u above their sum
    while (u > r && x < n && r > 0.0) {
      u -= r;
      x += 1.0;
      r *= a / x - s;
    }
  } else {
    // BTRS (W. Hormann, "The generation of binomial random variates", Journal of
    // Statistical Computation and Simulation 46, 1993)
    const double spq = sqrt(n * pp * q);
    const double b = 1.15 + 2.53 * spq, a = -0.0873 + 0.0248 * b + 0.01 * pp;
    const double c = n * pp + 0.5, vr = 0.92 - 4.2 / b;
    const double alpha = (2.83 + 5.1 / b) * spq, lpq = log(pp / q);
    const double m = floor((n + 1.0) * pp);
    const double h = lgamma(m + 1.0) + lgamma(n - m + 1.0);
    x = __builtin_nan("");
    for (int t = 0; t < GEN_MAX_ATTEMPTS; ++t) {
      const double u = w.uniform() - 0.5, v = w.uniform();
      const double us = 0.5 - fabs(u);
      const double k = floor((2.0 * a / us + b) * u + c);
      if (k < 0.0 || k > n) continue;
      if (us >= 0.07 && v <= vr) {
        x = k;
        break;
      }
      const double lv = log(v * alpha / (a / (us * us) + b));
      if (lv <= h - lgamma(k + 1.0) - lgamma(n - k + 1.0) + (k - m) * lpq) {
        x = k;
        break;
      }
    }
  }
  return flip ? n - x : x;
}

}  // namespace bihip
