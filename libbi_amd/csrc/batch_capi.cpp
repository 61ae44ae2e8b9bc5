// C ABI of the batched inner filters (bihip_batch_*, declared in include/bihip.h).
#include <algorithm>
#include <numeric>
#include <vector>

#include "internal.hpp"

using namespace bihip;

struct bihip_batch {
  bihip_ctx* ctx = nullptr;
  bihip_model model{};
  bihip_batch_config cfg{};
  int ntheta = 0, nobs = 1, d = 1, elem = 8;
  long ld = 0, total = 0;
  std::vector<int> N;
  std::vector<long> off;
  double* theta = nullptr;
  int* order = nullptr;
  long* offd = nullptr;
  BatchFilter* f = nullptr;
  void* X[2] = {nullptr, nullptr};
  int cur = 0;
  double* lws = nullptr;
  double* Ws = nullptr;
  int* as = nullptr;
  int* anc = nullptr;
  double* y = nullptr;
  double* hist_li = nullptr;
  double* hist_ess = nullptr;
  int* hist_res = nullptr;
  uint32_t* counter = nullptr;
  int k = 0;
  bool ready = false, have_obs = false;
  int block = 0, grid = 0;
  double tune_ms[3] = {0, 0, 0};
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  bool timing = false;
  std::vector<hipEvent_t> ev;
  size_t ev_used = 0;
};

static BatchLaunch batch_launch(bihip_batch* b) {
  BatchLaunch L{};
  L.model = &b->model;
  L.theta = b->theta;
  L.ntheta = b->ntheta;
  L.order = b->order;
  L.B = b->cfg.B;
  L.f = b->f;
  L.Xin = b->X[b->cur];
  L.Xout = b->X[1 - b->cur];
  L.ld = b->ld;
  L.lws = b->lws;
  L.Ws = b->Ws;
  L.as = b->as;
  L.anc_rec = b->anc;
  L.hist_li = b->hist_li;
  L.hist_ess = b->hist_ess;
  L.hist_res = b->hist_res;
  L.y = b->y;
  L.nobs = b->nobs;
  L.T = b->cfg.T;
  L.t0 = b->cfg.t0;
  L.dt = b->cfg.dt;
  L.seed = b->cfg.seed;
  L.essRel = b->cfg.essRel;
  L.resampler = b->cfg.resampler;
  L.counter = b->counter;
  return L;
}

// The autotuner: one step from the current state with each workgroup size,
// on a snapshot that is restored afterwards; the fastest size is kept.
static int batch_tune(bihip_batch* b) {
  hipStream_t s = ctx_stream_of(b->ctx);
  const int B = b->cfg.B;
  const size_t xbytes = (size_t)b->ld * b->d * b->elem;
  void* xs = nullptr;
  double* ls = nullptr;
  BatchFilter* fs = nullptr;
  double *hl = nullptr, *he = nullptr;
  int* hr = nullptr;
  int rc;
  if ((rc = dalloc((uint8_t**)&xs, xbytes)) || (rc = dalloc(&ls, b->total)) ||
      (rc = dalloc(&fs, B)) || (rc = dalloc(&hl, (size_t)B * b->cfg.T)) ||
      (rc = dalloc(&he, (size_t)B * b->cfg.T)) || (rc = dalloc(&hr, (size_t)B * b->cfg.T))) {
    dfree(xs), dfree(ls), dfree(fs), dfree(hl), dfree(he), dfree(hr);
    return rc;
  }
  auto save = [&]() -> int {
    BIHIP_HIP(hipMemcpyAsync(xs, b->X[b->cur], xbytes, hipMemcpyDeviceToDevice, s));
    BIHIP_HIP(hipMemcpyAsync(ls, b->lws, sizeof(double) * b->total, hipMemcpyDeviceToDevice, s));
    BIHIP_HIP(hipMemcpyAsync(fs, b->f, sizeof(BatchFilter) * B, hipMemcpyDeviceToDevice, s));
    BIHIP_HIP(hipMemcpyAsync(hl, b->hist_li, sizeof(double) * B * b->cfg.T,
                             hipMemcpyDeviceToDevice, s));
    BIHIP_HIP(hipMemcpyAsync(he, b->hist_ess, sizeof(double) * B * b->cfg.T,
                             hipMemcpyDeviceToDevice, s));
    BIHIP_HIP(hipMemcpyAsync(hr, b->hist_res, sizeof(int) * B * b->cfg.T,
                             hipMemcpyDeviceToDevice, s));
    return BIHIP_OK;
  };
  auto restore = [&]() -> int {
    BIHIP_HIP(hipMemcpyAsync(b->X[b->cur], xs, xbytes, hipMemcpyDeviceToDevice, s));
    BIHIP_HIP(hipMemcpyAsync(b->lws, ls, sizeof(double) * b->total, hipMemcpyDeviceToDevice, s));
    BIHIP_HIP(hipMemcpyAsync(b->f, fs, sizeof(BatchFilter) * B, hipMemcpyDeviceToDevice, s));
    BIHIP_HIP(hipMemcpyAsync(b->hist_li, hl, sizeof(double) * B * b->cfg.T,
                             hipMemcpyDeviceToDevice, s));
    BIHIP_HIP(hipMemcpyAsync(b->hist_ess, he, sizeof(double) * B * b->cfg.T,
                             hipMemcpyDeviceToDevice, s));
    BIHIP_HIP(hipMemcpyAsync(b->hist_res, hr, sizeof(int) * B * b->cfg.T,
                             hipMemcpyDeviceToDevice, s));
    return BIHIP_OK;
  };
  const int cand[3] = {256, 512, 1024};
  hipEvent_t e0, e1;
  BIHIP_HIP(hipEventCreate(&e0));
  BIHIP_HIP(hipEventCreate(&e1));
  rc = save();
  double best = 1e300;
  for (int c = 0; c < 3 && rc == BIHIP_OK; ++c) {
    // two runs per size: the first also loads the code object
    for (int rep = 0; rep < 2 && rc == BIHIP_OK; ++rep) {
      if ((rc = restore())) break;
      BIHIP_HIP(hipEventRecord(e0, s));
      int grid = 0;
      if ((rc = launch_batch_step(batch_launch(b), cand[c], s, &grid))) break;
      BIHIP_HIP(hipEventRecord(e1, s));
      BIHIP_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      BIHIP_HIP(hipEventElapsedTime(&ms, e0, e1));
      b->tune_ms[c] = ms;
      if (rep == 1 && ms < best) {
        best = ms;
        b->block = cand[c];
        b->grid = grid;
      }
    }
  }
  if (rc == BIHIP_OK) rc = restore();
  (void)hipStreamSynchronize(s);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  dfree(xs), dfree(ls), dfree(fs), dfree(hl), dfree(he), dfree(hr);
  return rc;
}

static int batch_capture(bihip_batch* b, int parity) {
  hipStream_t cs;
  BIHIP_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipGraph_t g = nullptr;
  const int saved = b->cur;
  b->cur = parity;
  int rc = BIHIP_OK;
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    int grid = 0;
    rc = launch_batch_step(batch_launch(b), b->block, cs, &grid);
    hipError_t e2 = hipStreamEndCapture(cs, &g);
    if (rc == BIHIP_OK && e2 != hipSuccess)
      rc = set_error(BIHIP_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e2));
  } else {
    rc = set_error(BIHIP_EHIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
  }
  b->cur = saved;
  if (rc == BIHIP_OK) {
    hipError_t e3 = hipGraphInstantiate(&b->graph[parity], g, nullptr, nullptr, 0);
    if (e3 != hipSuccess)
      rc = set_error(BIHIP_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e3));
  }
  if (g) (void)hipGraphDestroy(g);
  (void)hipStreamDestroy(cs);
  return rc;
}

extern "C" {

int bihip_batch_create(bihip_ctx* ctx, const bihip_model* m, const bihip_batch_config* cfg,
                       const int* N_host, const double* theta_host, int ntheta,
                       bihip_batch** out) {
  int rc = ctx_enter_device(ctx);
  if (rc) return rc;
  if (m == nullptr || cfg == nullptr || N_host == nullptr || out == nullptr)
    return set_error(BIHIP_EINVAL, "NULL argument");
  if (!batch_model_supported(m))
    return set_error(BIHIP_EINVAL, "batched filters support the PZ and LG models");
  if (cfg->B <= 0 || cfg->T <= 0) return set_error(BIHIP_EINVAL, "need B > 0 and T > 0");
  if (cfg->resampler != BIHIP_SYSTEMATIC && cfg->resampler != BIHIP_STRATIFIED)
    return set_error(BIHIP_EINVAL, "batched filters resample systematically or stratified");
  if (!(cfg->essRel >= 0.0 && cfg->essRel <= 1.0))
    return set_error(BIHIP_EINVAL, "essRel must lie in [0,1]");
  if (!(cfg->dt > 0.0)) return set_error(BIHIP_EINVAL, "dt must be positive");
  if (ntheta < 0 || ntheta > 16 || (ntheta > 0 && theta_host == nullptr))
    return set_error(BIHIP_EINVAL, "ntheta must lie in [0,16] with theta given");
  if (cfg->block != 0 && cfg->block != 256 && cfg->block != 512 && cfg->block != 1024)
    return set_error(BIHIP_EINVAL, "block must be 0 (autotune), 256, 512 or 1024");
  const int B = cfg->B;
  bihip_batch* b = new bihip_batch();
  b->ctx = ctx;
  b->model = *m;
  b->cfg = *cfg;
  b->ntheta = ntheta;
  b->nobs = m->model == BIHIP_MODEL_L96 ? m->d / 2 : 1;
  b->d = m->d;
  b->elem = m->precision == BIHIP_F64 ? 8 : 4;
  b->block = cfg->block;
  b->N.assign(N_host, N_host + B);
  b->off.resize(B);
  long total = 0;
  for (int j = 0; j < B; ++j) {
    if (b->N[j] <= 0 || b->N[j] > batch_nmax()) {
      delete b;
      return set_error(BIHIP_EINVAL, "inner filter %d has N = %d outside [1, %d]", j, N_host[j],
                       batch_nmax());
    }
    b->off[j] = total;
    total += b->N[j];
  }
  b->total = total;
  b->ld = (total + 63) / 64 * 64;
  std::vector<int> ord(B);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return b->N[x] > b->N[y]; });
#define B_ALLOC(ptr, n)                         \
  if ((rc = dalloc(&(ptr), (n))) != BIHIP_OK) { \
    bihip_batch_destroy(b);                     \
    return rc;                                  \
  }
  uint8_t *x0 = nullptr, *x1 = nullptr;
  B_ALLOC(x0, (size_t)b->ld * b->d * b->elem);
  B_ALLOC(x1, (size_t)b->ld * b->d * b->elem);
  b->X[0] = x0;
  b->X[1] = x1;
  B_ALLOC(b->theta, (size_t)B * std::max(ntheta, 1));
  B_ALLOC(b->order, B);
  B_ALLOC(b->offd, B);
  B_ALLOC(b->f, B);
  B_ALLOC(b->lws, total);
  B_ALLOC(b->Ws, total);
  B_ALLOC(b->as, total);
  B_ALLOC(b->y, (size_t)cfg->T * b->nobs);
  B_ALLOC(b->hist_li, (size_t)B * cfg->T);
  B_ALLOC(b->hist_ess, (size_t)B * cfg->T);
  B_ALLOC(b->hist_res, (size_t)B * cfg->T);
  B_ALLOC(b->counter, 1);
  if (cfg->record) B_ALLOC(b->anc, (size_t)cfg->T * b->ld);
#undef B_ALLOC
  hipStream_t s = ctx_stream_of(ctx);
  if (ntheta > 0)
    BIHIP_HIP(hipMemcpyAsync(b->theta, theta_host, sizeof(double) * B * ntheta,
                             hipMemcpyHostToDevice, s));
  BIHIP_HIP(hipMemcpyAsync(b->order, ord.data(), sizeof(int) * B, hipMemcpyHostToDevice, s));
  BIHIP_HIP(hipMemcpyAsync(b->offd, b->off.data(), sizeof(long) * B, hipMemcpyHostToDevice, s));
  BIHIP_HIP(hipStreamSynchronize(s));
  *out = b;
  return BIHIP_OK;
}

int bihip_batch_destroy(bihip_batch* b) {
  if (b == nullptr) return BIHIP_OK;
  (void)ctx_enter_device(b->ctx);
  (void)hipStreamSynchronize(ctx_stream_of(b->ctx));
  for (int i = 0; i < 2; ++i) {
    if (b->graph[i]) (void)hipGraphExecDestroy(b->graph[i]);
    dfree(b->X[i]);
  }
  for (auto e : b->ev) (void)hipEventDestroy(e);
  dfree(b->theta);
  dfree(b->order);
  dfree(b->offd);
  dfree(b->f);
  dfree(b->lws);
  dfree(b->Ws);
  dfree(b->as);
  dfree(b->anc);
  dfree(b->y);
  dfree(b->hist_li);
  dfree(b->hist_ess);
  dfree(b->hist_res);
  dfree(b->counter);
  delete b;
  return BIHIP_OK;
}

int bihip_batch_set_observations(bihip_batch* b, const double* y_host) {
  if (b == nullptr || y_host == nullptr) return set_error(BIHIP_EINVAL, "NULL argument");
  int rc = ctx_enter_device(b->ctx);
  if (rc) return rc;
  hipStream_t s = ctx_stream_of(b->ctx);
  BIHIP_HIP(hipMemcpyAsync(b->y, y_host, sizeof(double) * b->cfg.T * b->nobs,
                           hipMemcpyHostToDevice, s));
  BIHIP_HIP(hipStreamSynchronize(s));
  b->have_obs = true;
  return BIHIP_OK;
}

int bihip_batch_init(bihip_batch* b) {
  if (b == nullptr) return set_error(BIHIP_EINVAL, "NULL batch");
  int rc = ctx_enter_device(b->ctx);
  if (rc) return rc;
  hipStream_t s = ctx_stream_of(b->ctx);
  b->cur = 0;
  b->k = 0;
  if ((rc = launch_batch_init(&b->model, b->theta, b->ntheta, b->offd, b->cfg.B, b->total,
                              b->cfg.seed, b->X[0], b->ld, s)))
    return rc;
  std::vector<BatchFilter> F(b->cfg.B);
  for (int j = 0; j < b->cfg.B; ++j) {
    F[j] = BatchFilter{};
    F[j].status = BIHIP_OK;
    F[j].N = b->N[j];
    F[j].off = b->off[j];
  }
  BIHIP_HIP(hipMemcpyAsync(b->f, F.data(), sizeof(BatchFilter) * F.size(), hipMemcpyHostToDevice,
                           s));
  BIHIP_HIP(hipMemsetAsync(b->lws, 0, sizeof(double) * b->total, s));
  BIHIP_HIP(hipMemsetAsync(b->hist_li, 0, sizeof(double) * b->cfg.B * b->cfg.T, s));
  BIHIP_HIP(hipMemsetAsync(b->hist_ess, 0, sizeof(double) * b->cfg.B * b->cfg.T, s));
  BIHIP_HIP(hipMemsetAsync(b->hist_res, 0, sizeof(int) * b->cfg.B * b->cfg.T, s));
  BIHIP_HIP(hipStreamSynchronize(s));
  b->ev_used = 0;
  b->ready = true;
  return BIHIP_OK;
}

int bihip_batch_step(bihip_batch* b, int n) {
  if (b == nullptr) return set_error(BIHIP_EINVAL, "NULL batch");
  if (!b->ready) return set_error(BIHIP_ESTATE, "bihip_batch_init has not been called");
  if (!b->have_obs) return set_error(BIHIP_ESTATE, "no observations set");
  if (n < 0 || b->k + n > b->cfg.T)
    return set_error(BIHIP_EINVAL, "steps beyond the schedule (k=%d, n=%d, T=%d)", b->k, n,
                     b->cfg.T);
  int rc = ctx_enter_device(b->ctx);
  if (rc) return rc;
  hipStream_t s = ctx_stream_of(b->ctx);
  if (b->block == 0 && n > 0) {
    if ((rc = batch_tune(b))) return rc;
  }
  for (int i = 0; i < n; ++i) {
    if (b->timing) {
      while (b->ev.size() < b->ev_used + 2) {
        hipEvent_t e;
        BIHIP_HIP(hipEventCreate(&e));
        b->ev.push_back(e);
      }
      BIHIP_HIP(hipEventRecord(b->ev[b->ev_used], s));
      int grid = 0;
      if ((rc = launch_batch_step(batch_launch(b), b->block, s, &grid))) return rc;
      b->grid = grid;
      BIHIP_HIP(hipEventRecord(b->ev[b->ev_used + 1], s));
      b->ev_used += 2;
    } else if (b->cfg.use_graph) {
      if (b->graph[b->cur] == nullptr && (rc = batch_capture(b, b->cur))) return rc;
      BIHIP_HIP(hipGraphLaunch(b->graph[b->cur], s));
    } else {
      int grid = 0;
      if ((rc = launch_batch_step(batch_launch(b), b->block, s, &grid))) return rc;
      b->grid = grid;
    }
    b->cur = 1 - b->cur;
    ++b->k;
  }
  return BIHIP_OK;
}

int bihip_batch_outputs(bihip_batch* b, double* ll, double* li, double* ess, int* res,
                        int* status) {
  if (b == nullptr) return set_error(BIHIP_EINVAL, "NULL batch");
  int rc = ctx_enter_device(b->ctx);
  if (rc) return rc;
  hipStream_t s = ctx_stream_of(b->ctx);
  const int B = b->cfg.B, T = b->cfg.T;
  std::vector<BatchFilter> F(B);
  BIHIP_HIP(hipMemcpyAsync(F.data(), b->f, sizeof(BatchFilter) * B, hipMemcpyDeviceToHost, s));
  if (li) BIHIP_HIP(hipMemcpyAsync(li, b->hist_li, sizeof(double) * B * T, hipMemcpyDeviceToHost, s));
  if (ess)
    BIHIP_HIP(hipMemcpyAsync(ess, b->hist_ess, sizeof(double) * B * T, hipMemcpyDeviceToHost, s));
  if (res) BIHIP_HIP(hipMemcpyAsync(res, b->hist_res, sizeof(int) * B * T, hipMemcpyDeviceToHost, s));
  BIHIP_HIP(hipStreamSynchronize(s));
  for (int j = 0; j < B; ++j) {
    if (ll) ll[j] = F[j].logLikelihood;
    if (status) status[j] = F[j].status;
  }
  return BIHIP_OK;
}

int bihip_batch_ancestors(bihip_batch* b, int j, int k, int* as_host) {
  if (b == nullptr || as_host == nullptr) return set_error(BIHIP_EINVAL, "NULL argument");
  if (b->anc == nullptr) return set_error(BIHIP_ESTATE, "ancestors were not recorded");
  if (j < 0 || j >= b->cfg.B || k < 0 || k >= b->k) return set_error(BIHIP_EINVAL, "bad j or k");
  int rc = ctx_enter_device(b->ctx);
  if (rc) return rc;
  hipStream_t s = ctx_stream_of(b->ctx);
  BIHIP_HIP(hipMemcpyAsync(as_host, b->anc + (size_t)k * b->ld + b->off[j],
                           sizeof(int) * b->N[j], hipMemcpyDeviceToHost, s));
  BIHIP_HIP(hipStreamSynchronize(s));
  return BIHIP_OK;
}

int bihip_batch_tuning(bihip_batch* b, int* block, int* grid, double* ms3) {
  if (b == nullptr) return set_error(BIHIP_EINVAL, "NULL batch");
  if (block) *block = b->block;
  if (grid) *grid = b->grid;
  if (ms3)
    for (int i = 0; i < 3; ++i) ms3[i] = b->tune_ms[i];
  return BIHIP_OK;
}

int bihip_batch_set_timing(bihip_batch* b, int enable) {
  if (b == nullptr) return set_error(BIHIP_EINVAL, "NULL batch");
  b->timing = enable != 0;
  b->ev_used = 0;
  return BIHIP_OK;
}

int bihip_batch_kernel_time(bihip_batch* b, double* ms, long* count) {
  if (b == nullptr) return set_error(BIHIP_EINVAL, "NULL batch");
  int rc = ctx_enter_device(b->ctx);
  if (rc) return rc;
  BIHIP_HIP(hipStreamSynchronize(ctx_stream_of(b->ctx)));
  double total = 0.0;
  for (size_t i = 0; i + 1 < b->ev_used; i += 2) {
    float t = 0.f;
    BIHIP_HIP(hipEventElapsedTime(&t, b->ev[i], b->ev[i + 1]));
    total += t;
  }
  if (ms) *ms = total;
  if (count) *count = (long)(b->ev_used / 2);
  return BIHIP_OK;
}

long bihip_batch_particles(bihip_batch* b) { return b ? b->total : -1; }

}  // extern "C"
