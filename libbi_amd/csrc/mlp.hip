// BF16 MFMA kernels of the multilayer perceptron: the tiled GEMM with fused
// epilogues that the forward and backward passes are built from, the loss,
// column sums, split-K reduction, Adam, and input conversion.
//
// The network has no counterpart in the reference (LibBi has no learned
// components): it is the learned surrogate flow and guided proposal built
// beside the bootstrap filter.  Matrices are row-major; bfloat16 operands,
// FP32 accumulation (v_mfma_f32_32x32x16_bf16).
#include "internal.hpp"
#include "mlp.hpp"

namespace bihip {

namespace {

constexpr int GB_M = 64, GB_N = 64, GB_K = 32, LDS_K = GB_K + 8;  // +8: bank spread

// C = op(A) op(B) over k in [z ksplit, (z + 1) ksplit), 64 x 64 tiles, four
// waves of 32 x 32.  Tiles are staged k-contiguous in LDS whatever the global
// orientation, so every fragment read is one 16-byte LDS load.
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) uint16_t As[GB_M * LDS_K];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[GB_N * LDS_K];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int m0 = blockIdx.x * GB_M, n0 = blockIdx.y * GB_N;
  const int kbeg = blockIdx.z * g.ksplit;
  const int kend = kbeg + g.ksplit < g.K ? kbeg + g.ksplit : g.K;
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
  for (int k0 = kbeg; k0 < kend; k0 += GB_K) {
    if (!TA) {
      const int m = tid >> 2, k8 = (tid & 3) * 8;
      *reinterpret_cast<u16x8*>(&As[m * LDS_K + k8]) =
          *reinterpret_cast<const u16x8*>(g.A + (size_t)(m0 + m) * g.lda + k0 + k8);
    } else {
      const int k = tid >> 3, m8 = (tid & 7) * 8;
      const u16x8 v = *reinterpret_cast<const u16x8*>(g.A + (size_t)(k0 + k) * g.lda + m0 + m8);
#pragma unroll
      for (int j = 0; j < 8; ++j) As[(m8 + j) * LDS_K + k] = v[j];
    }
    if (TB) {
      const int n = tid >> 2, k8 = (tid & 3) * 8;
      *reinterpret_cast<u16x8*>(&Bs[n * LDS_K + k8]) =
          *reinterpret_cast<const u16x8*>(g.B + (size_t)(n0 + n) * g.ldb + k0 + k8);
    } else {
      const int k = tid >> 3, n8 = (tid & 7) * 8;
      const u16x8 v = *reinterpret_cast<const u16x8*>(g.B + (size_t)(k0 + k) * g.ldb + n0 + n8);
#pragma unroll
      for (int j = 0; j < 8; ++j) Bs[(n8 + j) * LDS_K + k] = v[j];
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GB_K; kk += 16) {
      const u16x8 a = *reinterpret_cast<const u16x8*>(
          &As[(wm * 32 + (lane & 31)) * LDS_K + kk + 8 * (lane >> 5)]);
      const u16x8 b = *reinterpret_cast<const u16x8*>(
          &Bs[(wn * 32 + (lane & 31)) * LDS_K + kk + 8 * (lane >> 5)]);
      acc = mfma32(a, b, acc);
    }
    __syncthreads();
  }
  const int col = n0 + wn * 32 + (lane & 31);
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int row = m0 + wm * 32 + acc_row(r, lane);
    const float v = acc[r];
    if constexpr (EPI == EPI_STORE) {
      g.Cf[(size_t)blockIdx.z * g.M * g.ldc + (size_t)row * g.ldc + col] = v;
    } else if constexpr (EPI == EPI_BIAS_GELU) {
      const float z = v + g.bias[col];
      g.Cf[(size_t)row * g.ldc + col] = z;
      g.Cb[(size_t)row * g.ldcb + col] = f2bf(gelu_f(z));
    } else if constexpr (EPI == EPI_BIAS) {
      g.Cf[(size_t)row * g.ldc + col] = v + g.bias[col];
    } else {  // EPI_GELU_GRAD
      g.Cb[(size_t)row * g.ldcb + col] = f2bf(v * gelu_grad_f(g.Zprev[(size_t)row * g.ldz + col]));
    }
  }
}

// Y[i][j] = bf16(X[src(i)][j]) for j < d, 0 for d <= j < ldy, and rows i >= n zero.
__global__ void k_rows_to_bf16(const float* X, long ldx, const long* idx, int n, int rows, int d,
                               uint16_t* Y, int ldy) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long)rows * ldy) return;
  const int i = (int)(g / ldy), j = (int)(g % ldy);
  float v = 0.0f;
  if (i < n && j < d) {
    const long src = idx != nullptr ? idx[i] : i;
    v = X[src * ldx + j];
  }
  Y[g] = f2bf(v);
}

// dY = bf16(2 (Y - T) / (n d)) for the first n rows and d columns (0 elsewhere),
// and per-block partial sums of (Y - T)^2.
__global__ __launch_bounds__(256) void k_mse_grad(const float* Y, int ldy, const float* T, long ldt,
                                                  const long* idx, int n, int rows, int d,
                                                  float scale, uint16_t* dY, float* partial) {
  __shared__ float red[256];
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  float sq = 0.0f;
  if (g < (long)rows * ldy) {
    const int i = (int)(g / ldy), j = (int)(g % ldy);
    float gr = 0.0f;
    if (i < n && j < d) {
      const long src = idx != nullptr ? idx[i] : i;
      const float e = Y[(size_t)i * ldy + j] - T[src * ldt + j];
      sq = e * e;
      gr = scale * e;
    }
    dY[g] = f2bf(gr);
  }
  red[threadIdx.x] = sq;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Gaussian negative log-likelihood: Y holds the mean (columns < m) and the log
// variance (columns m .. 2m - 1) for m targets.  dY = bf16 of the gradient of
// sum [(t - mu)^2 e^{-s} + s] / (2 n m); partial sums of the per-target terms.
__global__ __launch_bounds__(256) void k_gauss_grad(const float* Y, int ldy, const float* T,
                                                    long ldt, const long* idx, int n, int rows,
                                                    int m, float scale, uint16_t* dY,
                                                    float* partial) {
  __shared__ float red[256];
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  float term = 0.0f;
  if (g < (long)rows * ldy) {
    const int i = (int)(g / ldy), j = (int)(g % ldy);
    float gr = 0.0f;
    if (i < n && j < 2 * m) {
      const int o = j < m ? j : j - m;
      const long src = idx != nullptr ? idx[i] : i;
      const float mu = Y[(size_t)i * ldy + o], lv = Y[(size_t)i * ldy + m + o];
      const float e = T[src * ldt + o] - mu, iv = __expf(-lv), q = e * e * iv;
      if (j < m) {
        gr = -scale * e * iv;
        term = 0.5f * (q + lv);
      } else {
        gr = 0.5f * scale * (1.0f - q);
      }
    }
    dY[g] = f2bf(gr);
  }
  red[threadIdx.x] = term;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// part[s][c] = sum of bf16 dZ[r][c] over rows r in [256 s, 256 s + 256), in a
// fixed order.  A workgroup takes 64 columns of one split: each thread reads
// eight columns of eight rows (rows rs, rs + 32, ..., rs + 224 of the split,
// rs = thread / 8) as 16-byte loads issued together, sums its rows in order,
// and the 32 row slots are then added in order through LDS.  Needs cols and
// ld to be multiples of 8 (the padded layer widths are multiples of 64).
__global__ __launch_bounds__(256) void k_colsum_bf16(const uint16_t* dZ, int ld, int rows,
                                                     int cols, float* part) {
  __shared__ float red[32][64 + 1];
  const int cg = threadIdx.x & 7, rs = threadIdx.x >> 3;
  const int c0 = blockIdx.x * 64 + cg * 8;
  const int r0 = blockIdx.y * 256;
  float acc[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) acc[j] = 0.0f;
  if (c0 < cols) {
    uint4 v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int r = r0 + rs + 32 * u;
      v[u] = r < rows ? *reinterpret_cast<const uint4*>(dZ + (size_t)r * ld + c0)
                      : make_uint4(0u, 0u, 0u, 0u);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint32_t w[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        acc[2 * q] += bf2f((uint16_t)(w[q] & 0xFFFFu));
        acc[2 * q + 1] += bf2f((uint16_t)(w[q] >> 16));
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) red[rs][cg * 8 + j] = acc[j];
  __syncthreads();
  if (threadIdx.x < 64) {
    const int c = blockIdx.x * 64 + threadIdx.x;
    float t = 0.0f;
    for (int q = 0; q < 32; ++q) t += red[q][threadIdx.x];
    if (c < cols) part[(size_t)blockIdx.y * cols + c] = t;
  }
}

// out[i] = sum_s part[s][i], added in the order s = 0, 1, ..., S-1.  Each
// thread takes four consecutive outputs (one 16-byte load per split when
// count is a multiple of 4) and issues the loads of eight splits before
// adding them, so the split loads overlap instead of each waiting for the
// previous add; the order of the additions is unchanged.
__global__ void k_sum_splits(const float* part, int S, long count, float* out) {
  const long i = ((long)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (i >= count) return;
  if ((count & 3) == 0) {
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    int z = 0;
    for (; z + 8 <= S; z += 8) {
      float4 v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        v[u] = *reinterpret_cast<const float4*>(part + (size_t)(z + u) * count + i);
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        acc.x += v[u].x;
        acc.y += v[u].y;
        acc.z += v[u].z;
        acc.w += v[u].w;
      }
    }
    for (; z < S; ++z) {
      const float4 v = *reinterpret_cast<const float4*>(part + (size_t)z * count + i);
      acc.x += v.x;
      acc.y += v.y;
      acc.z += v.z;
      acc.w += v.w;
    }
    *reinterpret_cast<float4*>(out + i) = acc;
    return;
  }
  const int n = count - i < 4 ? (int)(count - i) : 4;
  for (int q = 0; q < n; ++q) {
    float s = 0.0f;
    for (int z = 0; z < S; ++z) s += part[(size_t)z * count + i + q];
    out[i + q] = s;
  }
}

// out[0] = sum of x[0 .. n) by one workgroup: strided partial sums, then a
// fixed-order tree in LDS
__global__ __launch_bounds__(256) void k_sum_all(const float* x, int n, float* out) {
  __shared__ float red[256];
  float s = 0.0f;
  for (int i = threadIdx.x; i < n; i += 256) s += x[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// Adam (Kingma & Ba 2015) with bias correction; wbf (nullable) gets the bf16 copy.
__global__ void k_adam(float* w, float* m1, float* m2, const float* grad, long n, float lr,
                       float b1, float b2, float eps, float bc1, float bc2, uint16_t* wbf) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gr = grad[i];
  const float a = b1 * m1[i] + (1.0f - b1) * gr;
  const float b = b2 * m2[i] + (1.0f - b2) * gr * gr;
  m1[i] = a;
  m2[i] = b;
  const float x = w[i] - lr * (a / bc1) / (sqrtf(b / bc2) + eps);
  w[i] = x;
  if (wbf != nullptr) wbf[i] = f2bf(x);
}

__global__ void k_to_bf16(const float* x, long n, uint16_t* y) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

template <bool TA, bool TB>
int gemm_epi(const GemmArgs& g, int epi, dim3 grid, hipStream_t s) {
  switch (epi) {
    case EPI_STORE: hipLaunchKernelGGL((k_gemm<TA, TB, EPI_STORE>), grid, dim3(256), 0, s, g); break;
    case EPI_BIAS_GELU:
      hipLaunchKernelGGL((k_gemm<TA, TB, EPI_BIAS_GELU>), grid, dim3(256), 0, s, g);
      break;
    case EPI_BIAS: hipLaunchKernelGGL((k_gemm<TA, TB, EPI_BIAS>), grid, dim3(256), 0, s, g); break;
    case EPI_GELU_GRAD:
      hipLaunchKernelGGL((k_gemm<TA, TB, EPI_GELU_GRAD>), grid, dim3(256), 0, s, g);
      break;
    default: return set_error(BIHIP_EINVAL, "unknown GEMM epilogue %d", epi);
  }
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

}  // namespace

int launch_gemm(const GemmArgs& g, bool ta, bool tb, int epi, hipStream_t s) {
  // the kernel's tiling assumptions, checked here so that no access leaves the operands
  if (g.M <= 0 || g.N <= 0 || g.K <= 0 || g.M % GB_M || g.N % GB_N || g.K % GB_K ||
      g.ksplit <= 0 || g.ksplit % GB_K || g.lda % 8 || g.ldb % 8)
    return set_error(BIHIP_EINVAL, "GEMM shape %dx%dx%d (split %d, lda %ld, ldb %ld) off the "
                     "64x64x32 tiling", g.M, g.N, g.K, g.ksplit, g.lda, g.ldb);
  if (((uintptr_t)g.A | (uintptr_t)g.B) & 15)
    return set_error(BIHIP_EINVAL, "GEMM operands must be 16-byte aligned");
  const dim3 grid(g.M / GB_M, g.N / GB_N, (g.K + g.ksplit - 1) / g.ksplit);
  if (!ta && tb) return gemm_epi<false, true>(g, epi, grid, s);
  if (!ta && !tb) return gemm_epi<false, false>(g, epi, grid, s);
  if (ta && !tb) return gemm_epi<true, false>(g, epi, grid, s);
  return gemm_epi<true, true>(g, epi, grid, s);
}

int launch_rows_to_bf16(const float* X, long ldx, const long* idx, int n, int rows, int d,
                        uint16_t* Y, int ldy, hipStream_t s) {
  const long tot = (long)rows * ldy;
  if (tot <= 0) return BIHIP_OK;
  hipLaunchKernelGGL(k_rows_to_bf16, dim3(ceil_div(tot, 256)), dim3(256), 0, s, X, ldx, idx, n,
                     rows, d, Y, ldy);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

int launch_mse_grad(const float* Y, int ldy, const float* T, long ldt, const long* idx, int n,
                    int rows, int d, float scale, uint16_t* dY, float* partial, int* nblocks,
                    hipStream_t s) {
  const long tot = (long)rows * ldy;
  *nblocks = ceil_div(tot, 256);
  hipLaunchKernelGGL(k_mse_grad, dim3(*nblocks), dim3(256), 0, s, Y, ldy, T, ldt, idx, n, rows, d,
                     scale, dY, partial);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

int launch_gauss_grad(const float* Y, int ldy, const float* T, long ldt, const long* idx, int n,
                      int rows, int m, float scale, uint16_t* dY, float* partial, int* nblocks,
                      hipStream_t s) {
  const long tot = (long)rows * ldy;
  *nblocks = ceil_div(tot, 256);
  hipLaunchKernelGGL(k_gauss_grad, dim3(*nblocks), dim3(256), 0, s, Y, ldy, T, ldt, idx, n, rows,
                     m, scale, dY, partial);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

int launch_colsum_bf16(const uint16_t* dZ, int ld, int rows, int cols, float* part, float* out,
                       hipStream_t s) {
  if (cols % 8 || ld % 8 || ((uintptr_t)dZ & 15))
    return set_error(BIHIP_EINVAL, "column sums need 8-aligned columns (cols %d, ld %d)", cols, ld);
  const int S = ceil_div(rows, 256);
  hipLaunchKernelGGL(k_colsum_bf16, dim3(ceil_div(cols, 64), S), dim3(256), 0, s, dZ, ld, rows,
                     cols, part);
  BIHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum_splits, dim3(ceil_div(cols, 1024)), dim3(256), 0, s, part, S,
                     (long)cols, out);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

int launch_sum_splits(const float* part, int S, long count, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_splits, dim3(ceil_div(count, 1024)), dim3(256), 0, s, part, S, count,
                     out);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

int launch_sum_all(const float* x, int n, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_all, dim3(1), dim3(256), 0, s, x, n, out);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

int launch_adam(float* w, float* m1, float* m2, const float* grad, long n, float lr, float b1,
                float b2, float eps, float bc1, float bc2, uint16_t* wbf, hipStream_t s) {
  if (n <= 0) return BIHIP_OK;
  hipLaunchKernelGGL(k_adam, dim3(ceil_div(n, 256)), dim3(256), 0, s, w, m1, m2, grad, n, lr, b1,
                     b2, eps, bc1, bc2, wbf);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

int launch_to_bf16(const float* x, long n, uint16_t* y, hipStream_t s) {
  if (n <= 0) return BIHIP_OK;
  hipLaunchKernelGGL(k_to_bf16, dim3(ceil_div(n, 256)), dim3(256), 0, s, x, n, y);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

}  // namespace bihip
