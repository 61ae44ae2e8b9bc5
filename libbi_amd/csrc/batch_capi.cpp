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
  long ld = 0, total = 0, cap = 0;  // ld: column stride, >= cap >= total
  std::vector<int> N;
  std::vector<long> off;
  std::vector<double> theta_h;      // host copy of theta (B x ntheta)
  uint32_t stream_base = 0;
  int anc_from = 0;                 // first step whose recorded ancestors hold
  double* theta = nullptr;
  long* aux_src = nullptr;          // whole-filter copies: source offsets,
  int* aux_len = nullptr;           // lengths
  int* aux_mask = nullptr;          // and selection
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
  int items = 4, oversub = 1;  // resampling tile width, grid oversubscription
  double tune_ms[3] = {0, 0, 0};  // best step time per workgroup size
  // every configuration the autotuner timed: (block, items, oversub) and ms
  std::vector<int> tune_cfg;
  std::vector<double> tune_all;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  bool timing = false;
  std::vector<hipEvent_t> ev;
  size_t ev_used = 0;
};

// Offsets (exclusive scan of N) and the largest-first processing order of a
// layout, uploaded to the device; returns the total or -1 beyond capacity.
static long batch_layout(bihip_batch* b, const int* N_host) {
  const int B = b->cfg.B;
  std::vector<long> off(B);
  long total = 0;
  for (int j = 0; j < B; ++j) {
    off[j] = total;
    total += N_host[j];
  }
  if (b->cap > 0 && total > b->cap) return -1;
  b->N.assign(N_host, N_host + B);
  b->off = off;
  b->total = total;
  return total;
}

static int batch_upload_layout(bihip_batch* b, hipStream_t s) {
  const int B = b->cfg.B;
  std::vector<int> ord(B);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return b->N[x] > b->N[y]; });
  BIHIP_HIP(hipMemcpyAsync(b->order, ord.data(), sizeof(int) * B, hipMemcpyHostToDevice, s));
  BIHIP_HIP(hipMemcpyAsync(b->offd, b->off.data(), sizeof(long) * B, hipMemcpyHostToDevice, s));
  BIHIP_HIP(hipStreamSynchronize(s));
  return BIHIP_OK;
}

static int check_sizes(const int* N_host, int B) {
  for (int j = 0; j < B; ++j)
    if (N_host[j] <= 0 || N_host[j] > batch_nmax())
      return set_error(BIHIP_EINVAL, "inner filter %d has N = %d outside [1, %d]", j, N_host[j],
                       batch_nmax());
  return BIHIP_OK;
}

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

// The autotuner: one step from the current state with each configuration
// (workgroup size 256 / 512 / 1024, resampling tile width of the widths
// compiled for the model, persistent grid of one or two times the resident
// workgroups per CU), on a snapshot that is restored afterwards; the fastest
// configuration is kept.
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
  int widths[3];
  const int nw = batch_items_for(&b->model, widths);
  hipEvent_t e0, e1;
  BIHIP_HIP(hipEventCreateWithFlags(&e0, hipEventDisableSystemFence));
  BIHIP_HIP(hipEventCreateWithFlags(&e1, hipEventDisableSystemFence));
  rc = save();
  double best = 1e300;
  b->tune_cfg.clear();
  b->tune_all.clear();
  for (int c = 0; c < 3; ++c) b->tune_ms[c] = 1e300;
  for (int c = 0; c < 3 && rc == BIHIP_OK; ++c)
    for (int w = 0; w < nw && rc == BIHIP_OK; ++w)
      for (int ov = 1; ov <= 2 && rc == BIHIP_OK; ++ov) {
        // two runs per configuration: the first also loads the code object
        for (int rep = 0; rep < 2 && rc == BIHIP_OK; ++rep) {
          if ((rc = restore())) break;
          BIHIP_HIP(hipEventRecord(e0, s));
          int grid = 0;
          if ((rc = launch_batch_step(batch_launch(b), cand[c], s, &grid, widths[w], ov))) break;
          BIHIP_HIP(hipEventRecord(e1, s));
          BIHIP_HIP(hipEventSynchronize(e1));
          float ms = 0.f;
          BIHIP_HIP(hipEventElapsedTime(&ms, e0, e1));
          if (rep == 0) continue;
          b->tune_cfg.insert(b->tune_cfg.end(), {cand[c], widths[w], ov});
          b->tune_all.push_back(ms);
          b->tune_ms[c] = std::min(b->tune_ms[c], (double)ms);
          if (ms < best) {
            best = ms;
            b->block = cand[c];
            b->items = widths[w];
            b->oversub = ov;
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
    rc = launch_batch_step(batch_launch(b), b->block, cs, &grid, b->items, b->oversub);
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
    return set_error(BIHIP_EINVAL,
                     "batched filters support the PZ and LG models and one-lane generated models");
  if (cfg->B <= 0 || cfg->T <= 0) return set_error(BIHIP_EINVAL, "need B > 0 and T > 0");
  if (cfg->resampler != BIHIP_SYSTEMATIC && cfg->resampler != BIHIP_STRATIFIED)
    return set_error(BIHIP_EINVAL, "batched filters resample systematically or stratified");
  if (!(cfg->essRel >= 0.0 && cfg->essRel <= 1.0))
    return set_error(BIHIP_EINVAL, "essRel must lie in [0,1]");
  if (!(cfg->dt > 0.0)) return set_error(BIHIP_EINVAL, "dt must be positive");
  if (int rc = check_interval(m, cfg->t0, cfg->t0 + cfg->dt)) return rc;
  if (ntheta < 0 || ntheta > 16 || (ntheta > 0 && theta_host == nullptr))
    return set_error(BIHIP_EINVAL, "ntheta must lie in [0,16] with theta given");
  if (cfg->block != 0 && cfg->block != 256 && cfg->block != 512 && cfg->block != 1024)
    return set_error(BIHIP_EINVAL, "block must be 0 (autotune), 256, 512 or 1024");
  if (cfg->items != 0) {
    int widths[3];
    const int nw = batch_items_for(m, widths);
    if (std::find(widths, widths + nw, cfg->items) == widths + nw)
      return set_error(BIHIP_EINVAL, "tile width %d not compiled for this model", cfg->items);
  }
  if (cfg->oversub < 0 || cfg->oversub > 8) return set_error(BIHIP_EINVAL, "oversub must lie in [0, 8]");
  const int B = cfg->B;
  bihip_batch* b = new bihip_batch();
  b->ctx = ctx;
  b->model = *m;
  b->cfg = *cfg;
  b->ntheta = ntheta;
  b->nobs = bihip_model_nobs(m);
  b->d = m->d;
  b->elem = m->precision == BIHIP_F64 ? 8 : 4;
  b->block = cfg->block;
  b->items = cfg->items > 0 ? cfg->items : 4;
  b->oversub = cfg->oversub > 0 ? cfg->oversub : 1;
  b->stream_base = cfg->stream_base;
  if ((rc = check_sizes(N_host, B)) != BIHIP_OK) {
    delete b;
    return rc;
  }
  batch_layout(b, N_host);
  b->cap = std::max(cfg->capacity, b->total);
  b->ld = (b->cap + 63) / 64 * 64;
  if (ntheta > 0) b->theta_h.assign(theta_host, theta_host + (size_t)B * ntheta);
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
  B_ALLOC(b->lws, b->cap);
  B_ALLOC(b->Ws, b->cap);
  B_ALLOC(b->as, b->cap);
  B_ALLOC(b->aux_src, B);
  B_ALLOC(b->aux_len, B);
  B_ALLOC(b->aux_mask, B);
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
  if ((rc = batch_upload_layout(b, s)) != BIHIP_OK) {
    bihip_batch_destroy(b);
    return rc;
  }
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
  dfree(b->aux_src);
  dfree(b->aux_len);
  dfree(b->aux_mask);
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
                              b->cfg.seed, b->stream_base, b->X[0], b->ld, s)))
    return rc;
  std::vector<BatchFilter> F(b->cfg.B);
  for (int j = 0; j < b->cfg.B; ++j) {
    F[j] = BatchFilter{};
    F[j].status = BIHIP_OK;
    F[j].N = b->N[j];
    F[j].off = b->off[j];
    F[j].stream = b->stream_base + (uint32_t)j;
  }
  BIHIP_HIP(hipMemcpyAsync(b->f, F.data(), sizeof(BatchFilter) * F.size(), hipMemcpyHostToDevice,
                           s));
  BIHIP_HIP(hipMemsetAsync(b->lws, 0, sizeof(double) * b->total, s));
  BIHIP_HIP(hipMemsetAsync(b->hist_li, 0, sizeof(double) * b->cfg.B * b->cfg.T, s));
  BIHIP_HIP(hipMemsetAsync(b->hist_ess, 0, sizeof(double) * b->cfg.B * b->cfg.T, s));
  BIHIP_HIP(hipMemsetAsync(b->hist_res, 0, sizeof(int) * b->cfg.B * b->cfg.T, s));
  BIHIP_HIP(hipStreamSynchronize(s));
  b->ev_used = 0;
  b->anc_from = 0;
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
        BIHIP_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        b->ev.push_back(e);
      }
      BIHIP_HIP(hipEventRecord(b->ev[b->ev_used], s));
      int grid = 0;
      if ((rc = launch_batch_step(batch_launch(b), b->block, s, &grid, b->items, b->oversub)))
        return rc;
      b->grid = grid;
      BIHIP_HIP(hipEventRecord(b->ev[b->ev_used + 1], s));
      b->ev_used += 2;
    } else if (b->cfg.use_graph) {
      if (b->graph[b->cur] == nullptr && (rc = batch_capture(b, b->cur))) return rc;
      BIHIP_HIP(hipGraphLaunch(b->graph[b->cur], s));
    } else {
      int grid = 0;
      if ((rc = launch_batch_step(batch_launch(b), b->block, s, &grid, b->items, b->oversub)))
        return rc;
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
  if (j < 0 || j >= b->cfg.B || k < b->anc_from || k >= b->k)
    return set_error(BIHIP_EINVAL, "bad j or k (recorded steps %d..%d of the current layout)",
                     b->anc_from, b->k - 1);
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
    for (int i = 0; i < 3; ++i) ms3[i] = b->tune_ms[i] < 1e300 ? b->tune_ms[i] : 0.0;
  return BIHIP_OK;
}

int bihip_batch_tuning_table(bihip_batch* b, int* items, int* oversub, int* n, int* cfg_host,
                             double* ms_host) {
  if (b == nullptr) return set_error(BIHIP_EINVAL, "NULL batch");
  if (items) *items = b->items;
  if (oversub) *oversub = b->oversub;
  const int m = (int)b->tune_all.size();
  if (n) {
    const int cap = *n;
    *n = m;
    for (int i = 0; i < m && i < cap; ++i) {
      if (cfg_host)
        for (int q = 0; q < 3; ++q) cfg_host[3 * i + q] = b->tune_cfg[3 * i + q];
      if (ms_host) ms_host[i] = b->tune_all[i];
    }
  }
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

int bihip_batch_set_theta(bihip_batch* b, const double* theta_host) {
  if (b == nullptr || theta_host == nullptr) return set_error(BIHIP_EINVAL, "NULL argument");
  if (b->ntheta == 0) return set_error(BIHIP_ESTATE, "the batch was created with ntheta = 0");
  int rc = ctx_enter_device(b->ctx);
  if (rc) return rc;
  hipStream_t s = ctx_stream_of(b->ctx);
  const size_t n = (size_t)b->cfg.B * b->ntheta;
  b->theta_h.assign(theta_host, theta_host + n);
  BIHIP_HIP(hipMemcpyAsync(b->theta, b->theta_h.data(), sizeof(double) * n,
                           hipMemcpyHostToDevice, s));
  BIHIP_HIP(hipStreamSynchronize(s));
  return BIHIP_OK;
}

int bihip_batch_set_sizes(bihip_batch* b, const int* N_host) {
  if (b == nullptr || N_host == nullptr) return set_error(BIHIP_EINVAL, "NULL argument");
  int rc = check_sizes(N_host, b->cfg.B);
  if (rc) return rc;
  if ((rc = ctx_enter_device(b->ctx))) return rc;
  hipStream_t s = ctx_stream_of(b->ctx);
  BIHIP_HIP(hipStreamSynchronize(s));
  if (batch_layout(b, N_host) < 0)
    return set_error(BIHIP_EINVAL, "the sizes exceed the batch capacity of %ld particles", b->cap);
  b->ready = false;
  return batch_upload_layout(b, s);
}

int bihip_batch_set_stream_base(bihip_batch* b, uint32_t base) {
  if (b == nullptr) return set_error(BIHIP_EINVAL, "NULL batch");
  if ((uint64_t)base + (uint64_t)b->cfg.B > (1u << 24))
    return set_error(BIHIP_EINVAL, "stream ids must stay below 2^24");
  b->stream_base = base;
  return BIHIP_OK;
}

int bihip_batch_increments(bihip_batch* b, int k, double* li_host) {
  if (b == nullptr || li_host == nullptr) return set_error(BIHIP_EINVAL, "NULL argument");
  if (k < 0 || k >= b->k) return set_error(BIHIP_EINVAL, "step %d not taken yet", k);
  int rc = ctx_enter_device(b->ctx);
  if (rc) return rc;
  hipStream_t s = ctx_stream_of(b->ctx);
  BIHIP_HIP(hipMemcpy2DAsync(li_host, sizeof(double), b->hist_li + k,
                             sizeof(double) * b->cfg.T, sizeof(double), b->cfg.B,
                             hipMemcpyDeviceToHost, s));
  BIHIP_HIP(hipStreamSynchronize(s));
  return BIHIP_OK;
}

int bihip_batch_resample_filters(bihip_batch* b, const int* anc_host) {
  if (b == nullptr || anc_host == nullptr) return set_error(BIHIP_EINVAL, "NULL argument");
  if (!b->ready) return set_error(BIHIP_ESTATE, "bihip_batch_init has not been called");
  const int B = b->cfg.B;
  for (int j = 0; j < B; ++j)
    if (anc_host[j] < 0 || anc_host[j] >= B)
      return set_error(BIHIP_EINVAL, "ancestor %d of filter %d outside [0, %d)", anc_host[j], j,
                       B);
  int rc = ctx_enter_device(b->ctx);
  if (rc) return rc;
  hipStream_t s = ctx_stream_of(b->ctx);
  std::vector<BatchFilter> F(B), G(B);
  BIHIP_HIP(hipMemcpyAsync(F.data(), b->f, sizeof(BatchFilter) * B, hipMemcpyDeviceToHost, s));
  BIHIP_HIP(hipStreamSynchronize(s));
  const std::vector<long> off_old = b->off;
  std::vector<int> Nn(B), len(B);
  std::vector<long> src(B);
  for (int j = 0; j < B; ++j) {
    Nn[j] = b->N[anc_host[j]];
    src[j] = off_old[anc_host[j]];
  }
  const std::vector<int> N_old = b->N;
  if (batch_layout(b, Nn.data()) < 0) {
    batch_layout(b, N_old.data());
    return set_error(BIHIP_EINVAL, "the resampled sizes exceed the batch capacity of %ld "
                     "particles", b->cap);
  }
  for (int j = 0; j < B; ++j) {
    G[j] = F[anc_host[j]];
    G[j].N = b->N[j];
    G[j].off = b->off[j];
    G[j].stream = F[j].stream;  // each filter keeps its own random streams
  }
  if ((rc = batch_upload_layout(b, s))) return rc;
  BIHIP_HIP(hipMemcpyAsync(b->aux_src, src.data(), sizeof(long) * B, hipMemcpyHostToDevice, s));
  BIHIP_HIP(hipMemcpyAsync(b->aux_len, b->N.data(), sizeof(int) * B, hipMemcpyHostToDevice, s));
  // state into the other buffer, weights through the scratch Ws
  if ((rc = launch_batch_copy(b->offd, b->aux_src, b->aux_len, nullptr, B, b->total, b->d,
                              b->elem, b->X[b->cur], b->ld, b->lws, b->X[1 - b->cur], b->ld,
                              b->Ws, s)))
    return rc;
  BIHIP_HIP(hipMemcpyAsync(b->lws, b->Ws, sizeof(double) * b->total, hipMemcpyDeviceToDevice, s));
  BIHIP_HIP(hipMemcpyAsync(b->f, G.data(), sizeof(BatchFilter) * B, hipMemcpyHostToDevice, s));
  if (b->ntheta > 0) {
    std::vector<double> th((size_t)B * b->ntheta);
    for (int j = 0; j < B; ++j)
      std::copy(b->theta_h.begin() + (size_t)anc_host[j] * b->ntheta,
                b->theta_h.begin() + (size_t)(anc_host[j] + 1) * b->ntheta,
                th.begin() + (size_t)j * b->ntheta);
    b->theta_h = th;
    BIHIP_HIP(hipMemcpyAsync(b->theta, b->theta_h.data(), sizeof(double) * th.size(),
                             hipMemcpyHostToDevice, s));
  }
  BIHIP_HIP(hipStreamSynchronize(s));
  b->cur = 1 - b->cur;
  b->anc_from = b->k;
  return BIHIP_OK;
}

int bihip_batch_accept(bihip_batch* dst, bihip_batch* src, const int* accept_host) {
  if (dst == nullptr || src == nullptr || accept_host == nullptr)
    return set_error(BIHIP_EINVAL, "NULL argument");
  if (dst == src) return set_error(BIHIP_EINVAL, "source and destination are the same batch");
  if (dst->ctx != src->ctx) return set_error(BIHIP_EINVAL, "batches of different contexts");
  if (!dst->ready || !src->ready) return set_error(BIHIP_ESTATE, "batch not initialised");
  const int B = dst->cfg.B;
  if (src->cfg.B != B || src->N != dst->N || src->k != dst->k || src->d != dst->d ||
      src->elem != dst->elem || src->ntheta != dst->ntheta)
    return set_error(BIHIP_EINVAL, "the batches differ in layout, position, model or theta");
  int rc = ctx_enter_device(dst->ctx);
  if (rc) return rc;
  hipStream_t s = ctx_stream_of(dst->ctx);
  std::vector<BatchFilter> Fd(B), Fs(B);
  BIHIP_HIP(hipMemcpyAsync(Fd.data(), dst->f, sizeof(BatchFilter) * B, hipMemcpyDeviceToHost, s));
  BIHIP_HIP(hipMemcpyAsync(Fs.data(), src->f, sizeof(BatchFilter) * B, hipMemcpyDeviceToHost, s));
  BIHIP_HIP(hipStreamSynchronize(s));
  std::vector<int> mask(B);
  for (int j = 0; j < B; ++j) {
    mask[j] = accept_host[j] != 0;
    if (!mask[j]) continue;
    const uint32_t stream = Fd[j].stream;
    Fd[j] = Fs[j];
    Fd[j].off = dst->off[j];
    Fd[j].stream = stream;  // the filter keeps drawing from its own streams
    if (dst->ntheta > 0)
      std::copy(src->theta_h.begin() + (size_t)j * dst->ntheta,
                src->theta_h.begin() + (size_t)(j + 1) * dst->ntheta,
                dst->theta_h.begin() + (size_t)j * dst->ntheta);
  }
  BIHIP_HIP(hipMemcpyAsync(dst->aux_mask, mask.data(), sizeof(int) * B, hipMemcpyHostToDevice, s));
  BIHIP_HIP(hipMemcpyAsync(dst->aux_src, src->off.data(), sizeof(long) * B, hipMemcpyHostToDevice,
                           s));
  BIHIP_HIP(hipMemcpyAsync(dst->aux_len, dst->N.data(), sizeof(int) * B, hipMemcpyHostToDevice, s));
  if ((rc = launch_batch_copy(dst->offd, dst->aux_src, dst->aux_len, dst->aux_mask, B, dst->total,
                              dst->d, dst->elem, src->X[src->cur], src->ld, src->lws,
                              dst->X[dst->cur], dst->ld, dst->lws, s)))
    return rc;
  BIHIP_HIP(hipMemcpyAsync(dst->f, Fd.data(), sizeof(BatchFilter) * B, hipMemcpyHostToDevice, s));
  if (dst->ntheta > 0)
    BIHIP_HIP(hipMemcpyAsync(dst->theta, dst->theta_h.data(),
                             sizeof(double) * dst->theta_h.size(), hipMemcpyHostToDevice, s));
  BIHIP_HIP(hipStreamSynchronize(s));
  dst->anc_from = dst->k;
  return BIHIP_OK;
}

}  // extern "C"
