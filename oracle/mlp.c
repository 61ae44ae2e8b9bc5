/*
 * FP64 multilayer perceptron of the oracle: forward pass, mean-squared-error
 * loss and its gradient by backpropagation.  TEST INFRASTRUCTURE ONLY (see
 * oracle.h).
 *
 * The network has no counterpart in the reference (LibBi has no learned
 * components); it is the checker of the HIP BF16 MFMA kernels that train and
 * run the surrogate flow and the guided proposal.  Layers l = 0..L-1 map
 * d_l -> d_{l+1} as z = W_l a + b_l, with W_l stored row-major [d_{l+1}][d_l]
 * followed by b_l in the flat parameter vector; hidden layers apply GELU in
 * its sigmoid form a = z sigma(1.702 z) (Hendrycks & Gimpel 2016, sec. 2), the
 * last layer is linear.
 * Losses: ORC_LOSS_MSE: sum over samples and outputs of (y - t)^2 / (n d_L);
 * ORC_LOSS_GAUSS: the outputs are a mean mu (first m = d_L / 2) and a log
 * variance s (last m) of a diagonal Gaussian for m targets, and the loss is
 * the negative log-likelihood without its constant,
 * sum over samples and targets of [(t - mu)^2 e^{-s} + s] / (2 n m).
 *
 * With bf16 = 1 the oracle rounds, as the GPU kernels do, the inputs, the
 * weights, the hidden activations and the back-propagated errors dL/dy,
 * dL/dz to bfloat16 (round to nearest even) and accumulates in FP64; biases
 * and pre-activations stay unrounded.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* round to bfloat16, nearest even, through float as the kernels' inputs are */
double orc_bf16(double x) {
  float f = (float)x;
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return x; /* NaN unchanged */
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  memcpy(&f, &u, 4);
  return (double)f;
}

static double sig(double z) { return 1.0 / (1.0 + exp(-1.702 * z)); }

static double gelu(double z) { return z * sig(z); }

static double gelu_grad(double z) {
  const double s = sig(z);
  return s + 1.702 * z * s * (1.0 - s);
}

static size_t layer_offset(const int* dims, int l) {
  size_t off = 0;
  for (int i = 0; i < l; ++i) off += (size_t)dims[i + 1] * dims[i] + dims[i + 1];
  return off;
}

size_t orc_mlp_nparams(int L, const int* dims) { return layer_offset(dims, L); }

/* per-sample forward; a[l] holds the input of layer l, z[l] its pre-activation */
static void forward_one(int L, const int* dims, const double* p, const double* x, int bf,
                        double** a, double** z) {
  for (int i = 0; i < dims[0]; ++i) a[0][i] = bf ? orc_bf16(x[i]) : x[i];
  for (int l = 0; l < L; ++l) {
    const double* W = p + layer_offset(dims, l);
    const double* b = W + (size_t)dims[l + 1] * dims[l];
    for (int o = 0; o < dims[l + 1]; ++o) {
      double s = 0.0;
      for (int i = 0; i < dims[l]; ++i) {
        const double w = bf ? orc_bf16(W[(size_t)o * dims[l] + i]) : W[(size_t)o * dims[l] + i];
        s += w * a[l][i];
      }
      s += b[o];
      z[l][o] = s;
      if (l + 1 < L) a[l + 1][o] = bf ? orc_bf16(gelu(s)) : gelu(s);
    }
  }
}

static int max_width(int L, const int* dims) {
  int w = 0;
  for (int l = 0; l <= L; ++l) w = dims[l] > w ? dims[l] : w;
  return w;
}

void orc_mlp_forward(int L, const int* dims, const double* params, const double* X, int n,
                     int bf16, double* Y) {
  const int W = max_width(L, dims);
#pragma omp parallel
  {
    double** a = (double**)malloc(sizeof(double*) * (L + 1));
    double** z = (double**)malloc(sizeof(double*) * (L + 1));
    for (int l = 0; l <= L; ++l) {
      a[l] = (double*)malloc(sizeof(double) * W);
      z[l] = (double*)malloc(sizeof(double) * W);
    }
#pragma omp for
    for (int s = 0; s < n; ++s) {
      forward_one(L, dims, params, X + (size_t)s * dims[0], bf16, a, z);
      memcpy(Y + (size_t)s * dims[L], z[L - 1], sizeof(double) * dims[L]);
    }
    for (int l = 0; l <= L; ++l) {
      free(a[l]);
      free(z[l]);
    }
    free(a);
    free(z);
  }
}

double orc_mlp_loss_grad(int L, const int* dims, const double* params, const double* X,
                         const double* T, int n, int loss_kind, int bf16, double* grad) {
  const size_t np = layer_offset(dims, L);
  const int W = max_width(L, dims), dout = dims[L];
  const int gauss = loss_kind == ORC_LOSS_GAUSS, m = gauss ? dout / 2 : dout;
  const double scale = gauss ? 1.0 / ((double)n * m) : 2.0 / ((double)n * dout);
  double loss = 0.0;
  memset(grad, 0, sizeof(double) * np);
  /* per-sample contributions summed in sample order for determinism */
  double** a = (double**)malloc(sizeof(double*) * (L + 1));
  double** z = (double**)malloc(sizeof(double*) * (L + 1));
  for (int l = 0; l <= L; ++l) {
    a[l] = (double*)malloc(sizeof(double) * W);
    z[l] = (double*)malloc(sizeof(double) * W);
  }
  double* dz = (double*)malloc(sizeof(double) * W);
  double* da = (double*)malloc(sizeof(double) * W);
  for (int s = 0; s < n; ++s) {
    forward_one(L, dims, params, X + (size_t)s * dims[0], bf16, a, z);
    const double* t = T + (size_t)s * m;
    if (gauss) {
      for (int o = 0; o < m; ++o) {
        const double mu = z[L - 1][o], lv = z[L - 1][m + o];
        const double e = t[o] - mu, q = e * e * exp(-lv);
        loss += 0.5 * (q + lv);
        const double gm = -scale * e * exp(-lv), gs = 0.5 * scale * (1.0 - q);
        dz[o] = bf16 ? orc_bf16(gm) : gm;
        dz[m + o] = bf16 ? orc_bf16(gs) : gs;
      }
    } else {
      for (int o = 0; o < dout; ++o) {
        const double e = z[L - 1][o] - t[o];
        loss += e * e;
        dz[o] = bf16 ? orc_bf16(scale * e) : scale * e;
      }
    }
    for (int l = L - 1; l >= 0; --l) {
      const double* Wl = params + layer_offset(dims, l);
      double* gW = grad + layer_offset(dims, l);
      double* gb = gW + (size_t)dims[l + 1] * dims[l];
      for (int o = 0; o < dims[l + 1]; ++o) {
        gb[o] += dz[o];
        for (int i = 0; i < dims[l]; ++i) gW[(size_t)o * dims[l] + i] += dz[o] * a[l][i];
      }
      if (l == 0) break;
      for (int i = 0; i < dims[l]; ++i) {
        double s2 = 0.0;
        for (int o = 0; o < dims[l + 1]; ++o) {
          const double w = bf16 ? orc_bf16(Wl[(size_t)o * dims[l] + i]) : Wl[(size_t)o * dims[l] + i];
          s2 += dz[o] * w;
        }
        da[i] = s2;
      }
      for (int i = 0; i < dims[l]; ++i) {
        const double g = da[i] * gelu_grad(z[l - 1][i]);
        dz[i] = bf16 ? orc_bf16(g) : g;
      }
    }
  }
  for (int l = 0; l <= L; ++l) {
    free(a[l]);
    free(z[l]);
  }
  free(a);
  free(z);
  free(dz);
  free(da);
  return loss / ((double)n * m);
}
