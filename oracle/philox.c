/*
 * Philox4x32-10 and the variate conversions shared with the HIP library.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Philox4x32-10 follows Salmon, Moraes, Dror & Shaw, "Parallel random
 * numbers: as easy as 1, 2, 3", SC'11, section 4.  It replaces the
 * reference's Random facade (share/src/bi/random/Random.hpp:44-380), whose
 * host uniforms are boost::uniform_real on [lo,hi) and whose Gaussians are
 * boost::normal_distribution (share/src/bi/host/random/RngHost.hpp:103-152).
 * Parity unpinned against those exact streams: they depend on Boost and on
 * the OpenMP thread count (share/src/bi/host/random/RandomHost.cpp:12-27).
 */
#include <math.h>
#include "oracle.h"

#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
    uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += PHILOX_W0;
    k1 += PHILOX_W1;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void orc_philox_block(uint64_t seed, uint32_t index, uint32_t step, uint32_t tag,
    uint32_t batch, uint32_t block, uint32_t out[4]) {
  uint32_t ctr[4] = { index, step, tag | (batch << 8), block };
  uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) };
  orc_philox4x32_10(ctr, key, out);
}

/* uniform on [0,1) with 53 random bits: the top 53 bits of (r0:r1) */
double orc_u53(uint32_t r0, uint32_t r1) {
  uint64_t m = (((uint64_t)r0 << 32) | r1) >> 11;
  return (double)m * 0x1.0p-53;
}

void orc_uniforms(uint64_t seed, uint32_t step, uint32_t tag, uint32_t batch, int n, double* u) {
  for (int j = 0; j < n; ++j) {
    uint32_t r[4];
    orc_philox_block(seed, (uint32_t)j, step, tag, batch, 0, r);
    u[j] = orc_u53(r[0], r[1]);
  }
}

void orc_uniform_at(uint64_t seed, uint32_t index, uint32_t step, uint32_t tag, uint32_t batch,
    double* u) {
  uint32_t r[4];
  orc_philox_block(seed, index, step, tag, batch, 0, r);
  *u = orc_u53(r[0], r[1]);
}

/*
 * Box-Muller pair from two 32-bit words.  bits=32: u1 = 1 - r0*2^-32 in
 * (0,1], u2 = r1*2^-32 in [0,1).  bits=24: the top 24 bits of each word, as
 * the FP32 kernels use, so that both uniforms are exact in float.
 */
static void box_muller(uint32_t r0, uint32_t r1, int bits, double* z0, double* z1) {
  double u1, u2;
  if (bits == 24) {
    u1 = 1.0 - (double)(r0 >> 8) * 0x1.0p-24;
    u2 = (double)(r1 >> 8) * 0x1.0p-24;
  } else {
    u1 = 1.0 - (double)r0 * 0x1.0p-32;
    u2 = (double)r1 * 0x1.0p-32;
  }
  double rad = sqrt(-2.0 * log(u1));
  double ang = 2.0 * M_PI * u2;
  *z0 = rad * cos(ang);
  *z1 = rad * sin(ang);
}

void orc_normals(uint64_t seed, uint32_t index, uint32_t step, uint32_t tag, uint32_t batch,
    int n, int bits, double* z) {
  for (int b = 0; 4 * b < n; ++b) {
    uint32_t r[4];
    double g[4];
    orc_philox_block(seed, index, step, tag, batch, (uint32_t)b, r);
    box_muller(r[0], r[1], bits, &g[0], &g[1]);
    box_muller(r[2], r[3], bits, &g[2], &g[3]);
    for (int i = 0; i < 4 && 4 * b + i < n; ++i) {
      z[4 * b + i] = g[i];
    }
  }
}
