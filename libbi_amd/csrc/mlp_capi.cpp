// C ABI of the BF16 MFMA multilayer perceptron (bihip_mlp_*, declared in include/bihip.h).
//
// Layers l = 0..L-1 map d_l -> d_{l+1}: z = W_l a + b_l, GELU z sigma(1.702 z)
// on hidden layers, linear output.  Every width is padded to a multiple of 64 on the
// device (padded weights, inputs and output errors are zero, so the padding
// never reaches a result).  Parameters are kept in FP32 (the Adam master
// copy) and, for the weights, in bfloat16 for the matrix cores.  The flat host
// layout is, per layer, W_l row-major [d_{l+1}][d_l] then b_l.
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "internal.hpp"

using namespace bihip;

namespace {
constexpr uint32_t TAG_MLP_INIT = 11;
constexpr uint32_t TAG_MLP_BATCH = 12;
int pad64(int d) { return (d + 63) / 64 * 64; }
}  // namespace

struct bihip_mlp {
  bihip_ctx* ctx = nullptr;
  int L = 0;
  std::vector<int> dims, pad;
  std::vector<long> offW, offb;  // offsets in the FP32 arenas: [all W | all b]
  long nW = 0, ntot = 0;
  float *P = nullptr, *G = nullptr, *M1 = nullptr, *M2 = nullptr;
  uint16_t* Wbf = nullptr;       // bf16 weights, same offsets as the W part
  long adam_t = 0;
  int loss = BIHIP_LOSS_MSE;
  // workspace for batches of up to cap rows
  int cap = 0;
  uint16_t* X0 = nullptr;
  std::vector<float*> Z;         // pre-activations / output, per layer
  std::vector<uint16_t*> A;      // bf16 activations, A[l] = input of layer l (l >= 1)
  std::vector<uint16_t*> dZ;     // bf16 errors per layer output
  float* part = nullptr;         // split-K partials
  long part_n = 0;
  float* lossp = nullptr;
  uint16_t* Winf = nullptr;      // inference copy of the weights (permuted columns)
  float* binf = nullptr;         // inference copy of the biases (row order of Winf)
};

namespace {

void free_work(bihip_mlp* m) {
  dfree(m->X0);
  for (auto p : m->Z) dfree(p);
  for (auto p : m->A) dfree(p);
  for (auto p : m->dZ) dfree(p);
  dfree(m->part);
  dfree(m->lossp);
  m->X0 = nullptr;
  m->Z.clear();
  m->A.clear();
  m->dZ.clear();
  m->part = m->lossp = nullptr;
  m->cap = 0;
}

constexpr int KSPLIT = 256;  // batch rows per split of the weight-gradient GEMMs

int ensure_work(bihip_mlp* m, int rows) {
  if (rows <= m->cap) return BIHIP_OK;
  free_work(m);
  const int L = m->L;
  int rc;
  if ((rc = dalloc(&m->X0, (size_t)rows * m->pad[0]))) return rc;
  m->Z.assign(L, nullptr);
  m->A.assign(L, nullptr);
  m->dZ.assign(L, nullptr);
  long maxw = 0;
  for (int l = 0; l < L; ++l) {
    if ((rc = dalloc(&m->Z[l], (size_t)rows * m->pad[l + 1]))) return rc;
    if ((rc = dalloc(&m->dZ[l], (size_t)rows * m->pad[l + 1]))) return rc;
    if (l + 1 < L && (rc = dalloc(&m->A[l + 1], (size_t)rows * m->pad[l + 1]))) return rc;
    maxw = std::max(maxw, (long)m->pad[l + 1] * m->pad[l]);
  }
  const int S = (rows + KSPLIT - 1) / KSPLIT;
  m->part_n = maxw * S;
  if ((rc = dalloc(&m->part, (size_t)m->part_n))) return rc;
  if ((rc = dalloc(&m->lossp, (size_t)rows * m->pad[L] / 256 + 1))) return rc;
  m->cap = rows;
  return BIHIP_OK;
}

// forward over rows (multiple of 64) already in X0
int forward(bihip_mlp* m, int rows, hipStream_t s) {
  const int L = m->L;
  for (int l = 0; l < L; ++l) {
    GemmArgs g{};
    g.A = l == 0 ? m->X0 : m->A[l];
    g.lda = m->pad[l];
    g.B = m->Wbf + m->offW[l];
    g.ldb = m->pad[l];
    g.M = rows;
    g.N = m->pad[l + 1];
    g.K = m->pad[l];
    g.ksplit = g.K;
    g.bias = m->P + m->offb[l];
    g.Cf = m->Z[l];
    g.ldc = m->pad[l + 1];
    g.Cb = l + 1 < L ? m->A[l + 1] : nullptr;
    g.ldcb = m->pad[l + 1];
    int rc = launch_gemm(g, false, true, l + 1 < L ? EPI_BIAS_GELU : EPI_BIAS, s);
    if (rc) return rc;
  }
  return BIHIP_OK;
}

// backward from dZ[L-1] (already set) into the gradient arena G
int backward(bihip_mlp* m, int rows, hipStream_t s) {
  int rc;
  for (int l = m->L - 1; l >= 0; --l) {
    GemmArgs g{};  // dW_l = dZ_l^T A_l, split over the batch
    g.A = m->dZ[l];
    g.lda = m->pad[l + 1];
    g.B = l == 0 ? m->X0 : m->A[l];
    g.ldb = m->pad[l];
    g.M = m->pad[l + 1];
    g.N = m->pad[l];
    g.K = rows;
    g.ksplit = std::min(rows, KSPLIT);
    g.Cf = m->part;
    g.ldc = m->pad[l];
    if ((rc = launch_gemm(g, true, false, EPI_STORE, s))) return rc;
    const int S = (rows + g.ksplit - 1) / g.ksplit;
    if ((rc = launch_sum_splits(m->part, S, (long)g.M * g.N, m->G + m->offW[l], s))) return rc;
    if ((rc = launch_colsum_bf16(m->dZ[l], m->pad[l + 1], rows, m->pad[l + 1], m->part,
                                 m->G + m->offb[l], s)))
      return rc;
    if (l == 0) break;
    GemmArgs h{};  // dZ_{l-1} = (dZ_l W_l) * GELU'(Z_{l-1})
    h.A = m->dZ[l];
    h.lda = m->pad[l + 1];
    h.B = m->Wbf + m->offW[l];
    h.ldb = m->pad[l];
    h.M = rows;
    h.N = m->pad[l];
    h.K = m->pad[l + 1];
    h.ksplit = h.K;
    h.Zprev = m->Z[l - 1];
    h.ldz = m->pad[l];
    h.Cb = m->dZ[l - 1];
    h.ldcb = m->pad[l];
    if ((rc = launch_gemm(h, false, false, EPI_GELU_GRAD, s))) return rc;
  }
  return BIHIP_OK;
}

// output error dL/dy of the configured loss for n valid rows (rows padded);
// per-block partial sums of the loss terms into m->lossp
int output_error(bihip_mlp* m, const float* T, long ldt, const long* idx, int n, int rows,
                 int* nblocks, hipStream_t s) {
  const int L = m->L, d = m->dims[L];
  if (m->loss == BIHIP_LOSS_GAUSS)
    return launch_gauss_grad(m->Z[L - 1], m->pad[L], T, ldt, idx, n, rows, d / 2,
                             1.0f / ((float)n * (float)(d / 2)), m->dZ[L - 1], m->lossp, nblocks,
                             s);
  return launch_mse_grad(m->Z[L - 1], m->pad[L], T, ldt, idx, n, rows, d,
                         2.0f / ((float)n * (float)d), m->dZ[L - 1], m->lossp, nblocks, s);
}

int targets_per_sample(const bihip_mlp* m) {
  return m->loss == BIHIP_LOSS_GAUSS ? m->dims[m->L] / 2 : m->dims[m->L];
}

// loss and output error; *sumsq gets the summed loss terms
int loss_and_error(bihip_mlp* m, const float* T, long ldt, const long* idx, int n, int rows,
                   double* sumsq, hipStream_t s) {
  int nb = 0;
  int rc = output_error(m, T, ldt, idx, n, rows, &nb, s);
  if (rc) return rc;
  if (sumsq) {
    std::vector<float> h(nb);
    BIHIP_HIP(hipMemcpyAsync(h.data(), m->lossp, sizeof(float) * nb, hipMemcpyDeviceToHost, s));
    BIHIP_HIP(hipStreamSynchronize(s));
    double t = 0.0;
    for (float v : h) t += v;
    *sumsq = t;
  }
  return BIHIP_OK;
}

int upload_params(bihip_mlp* m, const float* flat, hipStream_t s) {
  std::vector<float> P(m->ntot, 0.0f);
  long o = 0;
  for (int l = 0; l < m->L; ++l) {
    const int din = m->dims[l], dout = m->dims[l + 1];
    for (int r = 0; r < dout; ++r)
      for (int c = 0; c < din; ++c) P[m->offW[l] + (long)r * m->pad[l] + c] = flat[o++];
    for (int r = 0; r < dout; ++r) P[m->offb[l] + r] = flat[o++];
  }
  BIHIP_HIP(hipMemcpyAsync(m->P, P.data(), sizeof(float) * m->ntot, hipMemcpyHostToDevice, s));
  int rc = launch_to_bf16(m->P, m->nW, m->Wbf, s);
  if (rc) return rc;
  BIHIP_HIP(hipStreamSynchronize(s));
  return BIHIP_OK;
}

void unpack(const bihip_mlp* m, const std::vector<float>& P, float* flat) {
  long o = 0;
  for (int l = 0; l < m->L; ++l) {
    const int din = m->dims[l], dout = m->dims[l + 1];
    for (int r = 0; r < dout; ++r)
      for (int c = 0; c < din; ++c) flat[o++] = P[m->offW[l] + (long)r * m->pad[l] + c];
    for (int r = 0; r < dout; ++r) flat[o++] = P[m->offb[l] + r];
  }
}

}  // namespace

namespace {
uint16_t host_f2bf(float f) {  // round to nearest even
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
}  // namespace

namespace bihip {
int mlp_inference_weights(bihip_mlp* m, const uint16_t* W[MLP_MAX_LAYERS],
                          const float* b[MLP_MAX_LAYERS], int out_interleave) {
  if (m->L > MLP_MAX_LAYERS) return set_error(BIHIP_EINVAL, "too many layers");
  const int L = m->L, dout = m->dims[L];
  if (out_interleave > 0 && (dout % 2 || 2 * ((dout / 2 + out_interleave - 1) / out_interleave) *
                                               out_interleave > m->pad[L]))
    return set_error(BIHIP_EINVAL, "output interleave does not fit the padded width");
  int rc;
  if (m->Winf == nullptr && (rc = dalloc(&m->Winf, m->nW))) return rc;
  if (m->binf == nullptr && (rc = dalloc(&m->binf, m->ntot - m->nW))) return rc;
  hipStream_t s = ctx_stream_of(m->ctx);
  std::vector<float> P(m->ntot);
  BIHIP_HIP(hipMemcpyAsync(P.data(), m->P, sizeof(float) * m->ntot, hipMemcpyDeviceToHost, s));
  BIHIP_HIP(hipStreamSynchronize(s));
  std::vector<uint16_t> H(m->nW, 0);
  std::vector<float> B(m->ntot - m->nW, 0.0f);
  for (int l = 0; l < L; ++l) {
    const int K = m->pad[l], N = m->pad[l + 1];
    for (int n = 0; n < N; ++n) {
      // source row: natural, or for the interleaved output layer block t of
      // 2 out_interleave rows = [m of outputs t*oi .., s of the same outputs]
      int srow = n;
      if (l == L - 1 && out_interleave > 0) {
        const int oi = out_interleave, t = n / (2 * oi), r = n % (2 * oi);
        const int dim = t * oi + (r < oi ? r : r - oi);
        srow = dim < dout / 2 ? (r < oi ? dim : dout / 2 + dim) : -1;
      }
      const long boff = m->offb[l] - m->nW;
      B[boff + n] = srow >= 0 ? P[m->offb[l] + srow] : 0.0f;
      for (int k = 0; k < K; ++k) {
        // position 16 c + p holds input 16 c + perm(p), perm = [0-3, 8-11, 4-7, 12-15]
        const int c = k >> 4, q = k & 15;
        const int src = l == 0 ? k : 16 * c + 8 * ((q & 7) >> 2) + 4 * (q >> 3) + (q & 3);
        H[m->offW[l] + (long)n * K + k] =
            srow >= 0 ? host_f2bf(P[m->offW[l] + (long)srow * K + src]) : (uint16_t)0;
      }
    }
    W[l] = m->Winf + m->offW[l];
    if (b != nullptr) b[l] = m->binf + (m->offb[l] - m->nW);
  }
  BIHIP_HIP(hipMemcpyAsync(m->Winf, H.data(), sizeof(uint16_t) * m->nW, hipMemcpyHostToDevice, s));
  BIHIP_HIP(hipMemcpyAsync(m->binf, B.data(), sizeof(float) * B.size(), hipMemcpyHostToDevice, s));
  BIHIP_HIP(hipStreamSynchronize(s));
  return BIHIP_OK;
}

// Device parameters of an MLP for the fused inference kernels (surrogate.hip).
MlpView mlp_view(const bihip_mlp* m) {
  MlpView v{};
  v.L = m->L;
  for (int l = 0; l <= m->L && l < MLP_MAX_LAYERS + 1; ++l) {
    v.dims[l] = m->dims[l];
    v.pad[l] = m->pad[l];
  }
  for (int l = 0; l < m->L && l < MLP_MAX_LAYERS; ++l) {
    v.W[l] = m->Wbf + m->offW[l];
    v.b[l] = m->P + m->offb[l];
  }
  return v;
}
}  // namespace bihip

extern "C" {

int bihip_mlp_create(bihip_ctx* ctx, int nlayers, const int* dims, uint64_t seed,
                     bihip_mlp** out) {
  int rc = ctx_enter_device(ctx);
  if (rc) return rc;
  if (dims == nullptr || out == nullptr) return set_error(BIHIP_EINVAL, "NULL argument");
  if (nlayers < 1 || nlayers > MLP_MAX_LAYERS)
    return set_error(BIHIP_EINVAL, "nlayers must lie in [1, %d]", MLP_MAX_LAYERS);
  for (int l = 0; l <= nlayers; ++l)
    if (dims[l] < 1 || dims[l] > 4096) return set_error(BIHIP_EINVAL, "width %d outside [1,4096]", dims[l]);
  bihip_mlp* m = new bihip_mlp();
  m->ctx = ctx;
  m->L = nlayers;
  m->dims.assign(dims, dims + nlayers + 1);
  for (int l = 0; l <= nlayers; ++l) m->pad.push_back(pad64(dims[l]));
  long o = 0;
  for (int l = 0; l < nlayers; ++l) {
    m->offW.push_back(o);
    o += (long)m->pad[l + 1] * m->pad[l];
  }
  m->nW = o;
  for (int l = 0; l < nlayers; ++l) {
    m->offb.push_back(o);
    o += m->pad[l + 1];
  }
  m->ntot = o;
  if ((rc = dalloc(&m->P, m->ntot)) || (rc = dalloc(&m->G, m->ntot)) ||
      (rc = dalloc(&m->M1, m->ntot)) || (rc = dalloc(&m->M2, m->ntot)) ||
      (rc = dalloc(&m->Wbf, m->nW))) {
    bihip_mlp_destroy(m);
    return rc;
  }
  hipStream_t s = ctx_stream_of(ctx);
  BIHIP_HIP(hipMemsetAsync(m->G, 0, sizeof(float) * m->ntot, s));
  BIHIP_HIP(hipMemsetAsync(m->M1, 0, sizeof(float) * m->ntot, s));
  BIHIP_HIP(hipMemsetAsync(m->M2, 0, sizeof(float) * m->ntot, s));
  // U(-1/sqrt(d_in), 1/sqrt(d_in)) for weights and biases, keyed Philox draws
  std::vector<float> flat((size_t)bihip_mlp_nparams(m));
  long k = 0;
  for (int l = 0; l < nlayers; ++l) {
    const double bound = 1.0 / std::sqrt((double)dims[l]);
    const long cnt = (long)dims[l + 1] * dims[l] + dims[l + 1];
    for (long i = 0; i < cnt; ++i) {
      const double u = bihip_uniform_host(seed, (uint32_t)i, (uint32_t)l, TAG_MLP_INIT, 0);
      flat[k++] = (float)(bound * (2.0 * u - 1.0));
    }
  }
  if ((rc = upload_params(m, flat.data(), s))) {
    bihip_mlp_destroy(m);
    return rc;
  }
  *out = m;
  return BIHIP_OK;
}

int bihip_mlp_destroy(bihip_mlp* m) {
  if (m == nullptr) return BIHIP_OK;
  (void)ctx_enter_device(m->ctx);
  (void)hipStreamSynchronize(ctx_stream_of(m->ctx));
  free_work(m);
  dfree(m->P);
  dfree(m->G);
  dfree(m->M1);
  dfree(m->M2);
  dfree(m->Wbf);
  dfree(m->Winf);
  dfree(m->binf);
  delete m;
  return BIHIP_OK;
}

long bihip_mlp_nparams(const bihip_mlp* m) {
  if (m == nullptr) return -1;
  long n = 0;
  for (int l = 0; l < m->L; ++l) n += (long)m->dims[l + 1] * m->dims[l] + m->dims[l + 1];
  return n;
}

int bihip_mlp_set_params(bihip_mlp* m, const float* flat_host) {
  if (m == nullptr || flat_host == nullptr) return set_error(BIHIP_EINVAL, "NULL argument");
  int rc = ctx_enter_device(m->ctx);
  if (rc) return rc;
  hipStream_t s = ctx_stream_of(m->ctx);
  if ((rc = upload_params(m, flat_host, s))) return rc;
  BIHIP_HIP(hipMemsetAsync(m->M1, 0, sizeof(float) * m->ntot, s));
  BIHIP_HIP(hipMemsetAsync(m->M2, 0, sizeof(float) * m->ntot, s));
  BIHIP_HIP(hipStreamSynchronize(s));
  m->adam_t = 0;
  return BIHIP_OK;
}

int bihip_mlp_get_params(bihip_mlp* m, float* flat_host) {
  if (m == nullptr || flat_host == nullptr) return set_error(BIHIP_EINVAL, "NULL argument");
  int rc = ctx_enter_device(m->ctx);
  if (rc) return rc;
  hipStream_t s = ctx_stream_of(m->ctx);
  std::vector<float> P(m->ntot);
  BIHIP_HIP(hipMemcpyAsync(P.data(), m->P, sizeof(float) * m->ntot, hipMemcpyDeviceToHost, s));
  BIHIP_HIP(hipStreamSynchronize(s));
  unpack(m, P, flat_host);
  return BIHIP_OK;
}

int bihip_mlp_forward(bihip_mlp* m, const float* X_dev, long ldx, long n, float* Y_dev,
                      long ldy) {
  if (m == nullptr || X_dev == nullptr || Y_dev == nullptr)
    return set_error(BIHIP_EINVAL, "NULL argument");
  if (n < 0 || ldx < m->dims[0] || ldy < m->dims[m->L])
    return set_error(BIHIP_EINVAL, "bad n or leading dimensions");
  int rc = ctx_enter_device(m->ctx);
  if (rc) return rc;
  hipStream_t s = ctx_stream_of(m->ctx);
  const long chunk = 16384;
  if ((rc = ensure_work(m, (int)std::min(chunk, (n + 63) / 64 * 64)))) return rc;
  const int L = m->L;
  for (long r0 = 0; r0 < n; r0 += chunk) {
    const int nr = (int)std::min(chunk, n - r0);
    const int rows = (nr + 63) / 64 * 64;
    if ((rc = launch_rows_to_bf16(X_dev + r0 * ldx, ldx, nullptr, nr, rows, m->dims[0], m->X0,
                                  m->pad[0], s)) ||
        (rc = forward(m, rows, s)))
      return rc;
    BIHIP_HIP(hipMemcpy2DAsync(Y_dev + r0 * ldy, sizeof(float) * ldy, m->Z[L - 1],
                               sizeof(float) * m->pad[L], sizeof(float) * m->dims[L], nr,
                               hipMemcpyDeviceToDevice, s));
  }
  return BIHIP_OK;
}

int bihip_mlp_loss_grad(bihip_mlp* m, const float* X_dev, long ldx, const float* T_dev, long ldt,
                        int n, double* loss_host, float* grad_host) {
  if (m == nullptr || X_dev == nullptr || T_dev == nullptr)
    return set_error(BIHIP_EINVAL, "NULL argument");
  if (n <= 0 || ldx < m->dims[0] || ldt < targets_per_sample(m))
    return set_error(BIHIP_EINVAL, "bad n or leading dimensions");
  int rc = ctx_enter_device(m->ctx);
  if (rc) return rc;
  hipStream_t s = ctx_stream_of(m->ctx);
  const int rows = (n + 63) / 64 * 64;
  if ((rc = ensure_work(m, rows))) return rc;
  double sq = 0.0;
  if ((rc = launch_rows_to_bf16(X_dev, ldx, nullptr, n, rows, m->dims[0], m->X0, m->pad[0], s)) ||
      (rc = forward(m, rows, s)) ||
      (rc = loss_and_error(m, T_dev, ldt, nullptr, n, rows, &sq, s)) ||
      (rc = backward(m, rows, s)))
    return rc;
  if (loss_host) *loss_host = sq / ((double)n * targets_per_sample(m));
  if (grad_host) {
    std::vector<float> G(m->ntot);
    BIHIP_HIP(hipMemcpyAsync(G.data(), m->G, sizeof(float) * m->ntot, hipMemcpyDeviceToHost, s));
    BIHIP_HIP(hipStreamSynchronize(s));
    unpack(m, G, grad_host);
  }
  return BIHIP_OK;
}

int bihip_mlp_train(bihip_mlp* m, const float* X_dev, long ldx, const float* T_dev, long ldt,
                    long n, int batch, int steps, float lr, uint64_t seed, float* loss_host) {
  if (m == nullptr || X_dev == nullptr || T_dev == nullptr)
    return set_error(BIHIP_EINVAL, "NULL argument");
  if (ldt < targets_per_sample(m)) return set_error(BIHIP_EINVAL, "ldt below the targets per sample");
  if (n <= 0 || batch <= 0 || batch % 64 || batch > n || steps < 0 || !(lr > 0.0f))
    return set_error(BIHIP_EINVAL, "need 0 < batch <= n, batch a multiple of 64, lr > 0");
  int rc = ctx_enter_device(m->ctx);
  if (rc) return rc;
  hipStream_t s = ctx_stream_of(m->ctx);
  if ((rc = ensure_work(m, batch))) return rc;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  long* perm_d = nullptr;
  float* loss_d = nullptr;
  if ((rc = dalloc(&perm_d, (size_t)n)) || (rc = dalloc(&loss_d, (size_t)std::max(steps, 1)))) {
    dfree(perm_d);
    return rc;
  }
  std::vector<long> perm(n);
  std::iota(perm.begin(), perm.end(), 0L);
  long pos = n;  // position in the current epoch's permutation
  long epoch = -1;
  for (int it = 0; it < steps && rc == BIHIP_OK; ++it) {
    if (pos + batch > n) {  // new epoch: Fisher-Yates with keyed uniforms
      ++epoch;
      for (long i = n - 1; i > 0; --i) {
        const double u = bihip_uniform_host(seed, (uint32_t)i, (uint32_t)epoch, TAG_MLP_BATCH,
                                            (uint32_t)(i >> 32));
        const long j = std::min(i, (long)(u * (double)(i + 1)));
        std::swap(perm[i], perm[j]);
      }
      // the previous epoch's batches must have read perm_d before it changes
      if (hipStreamSynchronize(s) != hipSuccess ||
          hipMemcpyAsync(perm_d, perm.data(), sizeof(long) * n, hipMemcpyHostToDevice, s) !=
              hipSuccess) {
        rc = set_error(BIHIP_EHIP, "uploading the batch permutation failed");
        break;
      }
      pos = 0;
    }
    const long* idx = perm_d + pos;
    pos += batch;
    int nb = 0;
    if ((rc = launch_rows_to_bf16(X_dev, ldx, idx, batch, batch, m->dims[0], m->X0, m->pad[0],
                                  s)) ||
        (rc = forward(m, batch, s)) ||
        (rc = output_error(m, T_dev, ldt, idx, batch, batch, &nb, s)) ||
        (rc = launch_sum_all(m->lossp, nb, loss_d + it, s)) || (rc = backward(m, batch, s)))
      break;
    ++m->adam_t;
    const float bc1 = 1.0f - std::pow(b1, (float)m->adam_t);
    const float bc2 = 1.0f - std::pow(b2, (float)m->adam_t);
    if ((rc = launch_adam(m->P, m->M1, m->M2, m->G, m->nW, lr, b1, b2, eps, bc1, bc2, m->Wbf, s)) ||
        (rc = launch_adam(m->P + m->nW, m->M1 + m->nW, m->M2 + m->nW, m->G + m->nW,
                          m->ntot - m->nW, lr, b1, b2, eps, bc1, bc2, nullptr, s)))
      break;
  }
  if (rc == BIHIP_OK && loss_host != nullptr && steps > 0) {
    std::vector<float> h(steps);
    if (hipMemcpyAsync(h.data(), loss_d, sizeof(float) * steps, hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      rc = set_error(BIHIP_EHIP, "reading the losses failed");
    const float inv = 1.0f / ((float)batch * (float)targets_per_sample(m));
    for (int it = 0; it < steps && rc == BIHIP_OK; ++it) loss_host[it] = h[it] * inv;
  }
  (void)hipStreamSynchronize(s);
  dfree(perm_d);
  dfree(loss_d);
  return rc;
}

int bihip_mlp_set_loss(bihip_mlp* m, int loss) {
  if (m == nullptr) return set_error(BIHIP_EINVAL, "NULL network");
  if (loss != BIHIP_LOSS_MSE && loss != BIHIP_LOSS_GAUSS)
    return set_error(BIHIP_EINVAL, "unknown loss %d", loss);
  if (loss == BIHIP_LOSS_GAUSS && m->dims[m->L] % 2)
    return set_error(BIHIP_EINVAL, "the Gaussian loss needs an even output width");
  m->loss = loss;
  return BIHIP_OK;
}

int bihip_surrogate_flow(bihip_ctx* ctx, bihip_mlp* m, const bihip_surrogate_norm* norm,
                         double par, const float* Xin_dev, float* Xout_dev, int ld, int P) {
  int rc = ctx_enter_device(ctx);
  if (rc) return rc;
  if (m == nullptr || norm == nullptr || Xin_dev == nullptr || Xout_dev == nullptr)
    return set_error(BIHIP_EINVAL, "NULL argument");
  if (P < 0 || ld < P) return set_error(BIHIP_EINVAL, "need 0 <= P <= ld");
  if (Xin_dev == Xout_dev) return set_error(BIHIP_EINVAL, "the flow is out of place");
  SurrogateLaunch L{};
  L.net = mlp_view(m);
  if ((rc = mlp_inference_weights(m, L.Winf))) return rc;
  L.x_shift = norm->x_shift;
  L.x_scale = norm->x_scale;
  L.dx_scale = norm->dx_scale;
  L.par = par;
  L.par_shift = norm->par_shift;
  L.par_scale = norm->par_scale;
  L.Xin = Xin_dev;
  L.Xout = Xout_dev;
  L.ld = ld;
  L.P = P;
  L.step = false;
  return launch_surrogate(L, ctx_stream_of(ctx));
}

int bihip_guided_propose(bihip_ctx* ctx, bihip_mlp* m, const bihip_guided_norm* norm, double par,
                         float* X_dev, int ld, int P, const double* y_dev, uint64_t seed,
                         uint32_t batch, uint32_t step, double noise_sd, double obs_sd,
                         double* inc_dev) {
  int rc = ctx_enter_device(ctx);
  if (rc) return rc;
  if (m == nullptr || norm == nullptr || X_dev == nullptr || y_dev == nullptr || inc_dev == nullptr)
    return set_error(BIHIP_EINVAL, "NULL argument");
  if (P < 0 || ld < P) return set_error(BIHIP_EINVAL, "need 0 <= P <= ld");
  GuidedLaunch G{};
  G.net = mlp_view(m);
  if ((rc = mlp_inference_weights(m, G.Winf, G.binf, 16))) return rc;
  G.x_shift = norm->x_shift;
  G.x_scale = norm->x_scale;
  G.innov_scale = norm->innov_scale;
  G.par = par;
  G.par_shift = norm->par_shift;
  G.par_scale = norm->par_scale;
  G.X = X_dev;
  G.ld = ld;
  G.P = P;
  G.y = y_dev;
  G.seed = seed;
  G.batch = batch;
  G.step = step;
  G.noise_sd = (float)noise_sd;
  G.obs_sd = obs_sd;
  G.inc = inc_dev;
  return launch_guided(G, ctx_stream_of(ctx));
}

}  // extern "C"
