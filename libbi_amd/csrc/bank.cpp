// Banks of bootstrap particle filters advanced by shared launches.
//
// B filters of P particles, one parameter vector per filter: the inner loop
// of MarginalSIR::step (share/src/bi/sampler/MarginalSIR.hpp:346-368), which
// the reference runs serially over theta, for inner filters larger than one
// workgroup and for models of any lane width.  Each launch of a step covers
// every filter -- filter f on row blockIdx.y = f of the grid -- so a step is
// the five launches of one filter's step, not B times them: the resampling
// chain (k_weight_sums, k_offspring_tiles, k_place_tiles, k_place_free),
// the fused transition (or the learned flow / proposal) and one workgroup
// per filter for the correct step.
//
// Layout in HBM (filter f): the state at X + f * d * ld (structure of
// arrays, leading dimension ld = P rounded up to 64), the per-particle
// vectors (lws, Os, as, EF words) at f * ld, the resampling tile words and
// scan state at f times their size, the PFDevice at dev + f, the parameter
// row at par + BANK_ROW * f; histories B x T (and B x T x P ancestors).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.hpp"

using namespace bihip;

struct bihip_pfbank {
  bihip_ctx* ctx = nullptr;
  bihip_model model{};
  bihip_pf_config cfg{};
  int B = 0, ld = 0, elem = 4, nobs = 1, nparts = 0, nparts_cap = 0, tiles = 0;
  void* X[2] = {nullptr, nullptr};
  int cur = 0;
  double* lws = nullptr;
  int* Os = nullptr;
  int* as = nullptr;
  uint64_t* EF = nullptr;
  double* partials = nullptr;
  PFDevice* dev = nullptr;
  double* y = nullptr;
  double* mean0 = nullptr;
  double* hist_li = nullptr;
  double* hist_ess = nullptr;
  int* hist_res = nullptr;
  int* hist_anc = nullptr;
  double* par = nullptr;     // B x BANK_ROW
  float* par_u = nullptr;    // B: normalised parameter input of a learned flow
  std::vector<double> rows;  // host copy of par
  int ntheta = 0;
  int* idx = nullptr;        // B ints: filter indices of whole-filter copies
  int* mask = nullptr;       // B ints: which filters a copy takes
  ScanState* ws_state = nullptr;
  uint64_t* ws_words = nullptr;
  int k = 0;
  bool ready = false, have_obs = false;
  int timing = 0;
  std::vector<hipEvent_t> ev[4];
  size_t ev_used[4] = {0, 0, 0, 0};
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  bool surr = false, guided = false;
  SurrogateLaunch sl{};
  GuidedLaunch gl{};
};

namespace {

BankStrides strides(const bihip_pfbank* b) {
  BankStrides bk;
  bk.x = (long)b->model.d * b->ld;
  bk.p = b->ld;
  bk.words = (long)SCAN_WORDS * b->tiles * (long)sizeof(uint64_t);
  bk.state = (long)sizeof(ScanState);
  return bk;
}

int event_pair(bihip_pfbank* b, int kind, hipEvent_t* e0, hipEvent_t* e1) {
  auto& v = b->ev[kind];
  size_t& u = b->ev_used[kind];
  while (v.size() < u + 2) {
    hipEvent_t e;
    BIHIP_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    v.push_back(e);
  }
  *e0 = v[u];
  *e1 = v[u + 1];
  u += 2;
  return BIHIP_OK;
}

// One observation interval of every filter: resample (gated per filter) ->
// record -> predict + correct -> the per-filter correct step.
int bank_enqueue_step(bihip_pfbank* b, hipStream_t s, int timing) {
  const bool timed = timing == BIHIP_TIMING_ALL;
  const bool timed_trans = timing != 0;
  const int P = b->cfg.P, B = b->B;
  const BankStrides bk = strides(b);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc;
  if (timed) {
    if ((rc = event_pair(b, 1, &e0, &e1))) return rc;
    BIHIP_HIP(hipEventRecord(e0, s));
  }
  ResampleTilesLaunch R{};
  R.mode = b->cfg.resampler;
  R.lws = b->lws;
  R.mx = &b->dev->mx;
  R.P = P;
  R.n = P;
  R.seed = b->cfg.seed;
  R.batch = b->cfg.batch;
  R.gate = b->dev;
  R.status = b->dev;
  R.Os = b->Os;
  R.as = b->as;
  R.EF = b->EF;
  R.ws = ScanWorkspace{b->ws_state, b->ws_words, b->tiles};
  R.banks = B;
  R.bk = bk;
  if ((rc = launch_resample_tiles(R, s))) return rc;
  if (b->hist_anc != nullptr &&
      (rc = launch_record_ancestors(b->as, P, b->dev, b->hist_anc, s, B, b->ld,
                                    (long)b->cfg.T * P)))
    return rc;
  if (timed) BIHIP_HIP(hipEventRecord(e1, s));

  if (timed_trans) {
    if ((rc = event_pair(b, 0, &e0, &e1))) return rc;
    BIHIP_HIP(hipEventRecord(e0, s));
  }
  TransLaunch T{};
  T.model = &b->model;
  T.Xin = b->X[b->cur];
  T.Xout = b->X[1 - b->cur];
  T.ld = b->ld;
  T.P = P;
  T.as = b->as;
  T.lws = b->lws;
  T.seed = b->cfg.seed;
  T.batch = b->cfg.batch;
  T.partials = b->partials;
  T.dev = b->dev;
  T.t0 = b->cfg.t0;
  T.dt = b->cfg.dt;
  T.y_base = b->y;
  T.nobs = b->nobs;
  T.banks = B;
  T.bk = bk;
  T.bank_par = b->par;
  if (b->guided) {
    TransLaunch F = T;  // the exact flow of the gathered particles into Xout
    F.flow_only = 1;
    F.y_base = nullptr;
    F.lws = nullptr;
    F.partials = nullptr;
    if ((rc = launch_transition(F, s))) return rc;
    GuidedLaunch G = b->gl;
    G.X = T.Xout;
    G.ld = b->ld;
    G.P = P;
    G.seed = b->cfg.seed;
    G.batch = b->cfg.batch;
    G.dev = b->dev;
    G.y_base = b->y;
    G.nobs = b->nobs;
    G.lws = b->lws;
    G.partials = b->partials;
    G.banks = B;
    G.bk = bk;
    G.par_u_bank = b->par_u;
    hipEvent_t g0 = nullptr, g1 = nullptr;
    if (timed_trans) {
      if ((rc = event_pair(b, 3, &g0, &g1))) return rc;
      BIHIP_HIP(hipEventRecord(g0, s));
    }
    if ((rc = launch_guided(G, s))) return rc;
    if (timed_trans) BIHIP_HIP(hipEventRecord(g1, s));
  } else if (b->surr) {
    SurrogateLaunch S = b->sl;
    S.Xin = T.Xin;
    S.Xout = T.Xout;
    S.ld = b->ld;
    S.P = P;
    S.step = true;
    S.as = b->as;
    S.lws = b->lws;
    S.partials = b->partials;
    S.dev = b->dev;
    S.seed = b->cfg.seed;
    S.batch = b->cfg.batch;
    S.t0 = b->cfg.t0;
    S.dt = b->cfg.dt;
    S.y_base = b->y;
    S.nobs = b->nobs;
    S.banks = B;
    S.bk = bk;
    S.par_u_bank = b->par_u;
    if ((rc = launch_surrogate(S, s))) return rc;
  } else if ((rc = launch_transition(T, s))) {
    return rc;
  }
  if (timed_trans) BIHIP_HIP(hipEventRecord(e1, s));

  if (timed) {
    if ((rc = event_pair(b, 2, &e0, &e1))) return rc;
    BIHIP_HIP(hipEventRecord(e0, s));
  }
  PFHistory H{b->hist_li, b->hist_ess, b->hist_res};
  if ((rc = launch_correct_bank(b->partials, b->nparts, P, b->cfg.essRel, b->dev, H, b->cfg.T, B,
                                s)))
    return rc;
  if (timed) BIHIP_HIP(hipEventRecord(e1, s));
  return BIHIP_OK;
}

int bank_capture(bihip_pfbank* b, int parity) {
  hipStream_t cs;
  BIHIP_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipGraph_t g = nullptr;
  const int saved = b->cur;
  b->cur = parity;
  int rc = BIHIP_OK;
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    rc = bank_enqueue_step(b, cs, 0);
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

void drop_graphs(bihip_pfbank* b) {
  for (int i = 0; i < 2; ++i) {
    if (b->graph[i]) (void)hipGraphExecDestroy(b->graph[i]);
    b->graph[i] = nullptr;
  }
}

// The learned flows serve FP32 Lorenz-96 filters whose filters differ only in
// the forcing (par[0], the network's parameter input).
int learned_rows_ok(const bihip_pfbank* b) {
  for (int j = 0; j < b->B; ++j)
    for (int i = 1; i < 16; ++i)
      if (b->rows[(size_t)j * BANK_ROW + i] != b->model.par[i])
        return set_error(BIHIP_EINVAL, "a learned flow takes banks whose filters differ only in "
                                       "par[0] (filter %d differs in par[%d])", j, i);
  return BIHIP_OK;
}

int set_par_u(bihip_pfbank* b, double shift, double scale) {
  std::vector<float> u(b->B);
  for (int j = 0; j < b->B; ++j)
    u[j] = (float)((b->rows[(size_t)j * BANK_ROW] - shift) / scale);
  BIHIP_HIP(hipMemcpy(b->par_u, u.data(), sizeof(float) * b->B, hipMemcpyHostToDevice));
  return BIHIP_OK;
}

// Parameter rows of theta (B x ntheta over the model's parameters), on the
// device, with the learned flow's parameter inputs.
int upload_rows(bihip_pfbank* b, const double* theta) {
  for (int j = 0; j < b->B; ++j) {
    bihip_model mj = b->model;
    for (int i = 0; i < b->ntheta; ++i) mj.par[i] = theta[(size_t)j * b->ntheta + i];
    if (b->model.model == BIHIP_MODEL_L96 && (b->model.par[2] > 0.0) != (mj.par[2] > 0.0))
      return set_error(BIHIP_EINVAL, "filter %d: the observation sd must stay positive", j);
    bank_row(&mj, &b->rows[(size_t)j * BANK_ROW]);
  }
  return BIHIP_OK;
}

int sync_rows(bihip_pfbank* b) {
  BIHIP_HIP(hipMemcpy(b->par, b->rows.data(), sizeof(double) * b->rows.size(),
                      hipMemcpyHostToDevice));
  if (b->guided) return set_par_u(b, b->gl.par_shift, b->gl.par_scale);
  if (b->surr) return set_par_u(b, b->sl.par_shift, b->sl.par_scale);
  return BIHIP_OK;
}

bool same_shape(const bihip_pfbank* a, const bihip_pfbank* b) {
  return a->B == b->B && a->cfg.P == b->cfg.P && a->ld == b->ld && a->elem == b->elem &&
         a->model.model == b->model.model && a->model.d == b->model.d &&
         a->model.precision == b->model.precision;
}

}  // namespace

extern "C" {

int bihip_pfbank_create(bihip_ctx* ctx, const bihip_model* m, const bihip_pf_config* cfg, int B,
                        const double* theta_host, int ntheta, bihip_pfbank** out) {
  if (cfg == nullptr || out == nullptr) return set_error(BIHIP_EINVAL, "NULL argument");
  int rc = ctx_enter_device(ctx);
  if (rc) return rc;
  if ((rc = check_model(m))) return rc;
  if (B <= 0) return set_error(BIHIP_EINVAL, "need B > 0");
  if (ntheta < 0 || ntheta > 16 || (ntheta > 0 && theta_host == nullptr))
    return set_error(BIHIP_EINVAL, "theta is B x ntheta with 0 <= ntheta <= 16");
  if (cfg->P <= 0 || cfg->T <= 0) return set_error(BIHIP_EINVAL, "need P > 0 and T > 0");
  if (cfg->resampler != BIHIP_SYSTEMATIC && cfg->resampler != BIHIP_STRATIFIED)
    return set_error(BIHIP_EINVAL, "filter banks resample systematically or stratified");
  if (!(cfg->essRel >= 0.0 && cfg->essRel <= 1.0))
    return set_error(BIHIP_EINVAL, "essRel must lie in [0,1]");
  if (!(cfg->dt > 0.0)) return set_error(BIHIP_EINVAL, "dt must be positive");
  if (cfg->record != 0 && cfg->record != 1)
    return set_error(BIHIP_EINVAL, "a filter bank records nothing or its ancestors (record 0/1)");
  if ((uint64_t)cfg->batch + (uint64_t)B > (1u << 24))
    return set_error(BIHIP_EINVAL, "Philox batch ids must stay below 2^24");
  if ((rc = check_interval(m, cfg->t0, cfg->t0 + cfg->dt))) return rc;
  const long ld = ((long)cfg->P + 63) / 64 * 64;
  if ((long)B * ld * m->d >= (1L << 40)) return set_error(BIHIP_EINVAL, "bank too large");
  if ((long)cfg->P * 8 > (1L << 31)) return set_error(BIHIP_EINVAL, "P too large");

  bihip_pfbank* b = new bihip_pfbank();
  b->ctx = ctx;
  b->model = *m;
  b->cfg = *cfg;
  b->B = B;
  b->ld = (int)ld;
  b->elem = bihip_model_elem_size(m);
  b->nobs = nobs_of(m);
  b->tiles = scan_tiles(cfg->P);
  b->nparts = transition_blocks(m, cfg->P);
  b->nparts_cap = b->nparts;
  b->ntheta = ntheta;
  const bool l96f32 = m->model == BIHIP_MODEL_L96 && m->precision == BIHIP_F32;
  if (l96f32 && m->d == 200) b->nparts_cap = std::max(b->nparts_cap, guided_blocks(cfg->P));
  if (l96f32 && m->d == 40) b->nparts_cap = std::max(b->nparts_cap, surrogate_blocks(cfg->P));
  // parameter rows: the model's, with theta's columns in place
  b->rows.resize((size_t)B * BANK_ROW);
  // the precomputed observation constants of Lorenz-96 (1 / obs_sd, log obs_sd)
  if ((rc = upload_rows(b, theta_host))) {
    delete b;
    return rc;
  }
  const size_t xb = (size_t)B * ld * m->d * b->elem;
  const size_t n = (size_t)B * ld;
#define BK_ALLOC(ptr, cnt)                           \
  if ((rc = dalloc(&(ptr), (cnt))) != BIHIP_OK) {    \
    bihip_pfbank_destroy(b);                         \
    return rc;                                       \
  }
  uint8_t* x0 = nullptr;
  uint8_t* x1 = nullptr;
  BK_ALLOC(x0, xb);
  BK_ALLOC(x1, xb);
  b->X[0] = x0;
  b->X[1] = x1;
  BK_ALLOC(b->lws, n);
  BK_ALLOC(b->Os, n);
  BK_ALLOC(b->as, n);
  BK_ALLOC(b->EF, n);
  BK_ALLOC(b->partials, (size_t)3 * B * b->nparts_cap);
  BK_ALLOC(b->dev, (size_t)B);
  BK_ALLOC(b->y, (size_t)cfg->T * b->nobs);
  BK_ALLOC(b->mean0, (size_t)m->d);
  BK_ALLOC(b->hist_li, (size_t)B * cfg->T);
  BK_ALLOC(b->hist_ess, (size_t)B * cfg->T);
  BK_ALLOC(b->hist_res, (size_t)B * cfg->T);
  if (cfg->record == 1) BK_ALLOC(b->hist_anc, (size_t)B * cfg->T * cfg->P);
  BK_ALLOC(b->par, (size_t)B * BANK_ROW);
  BK_ALLOC(b->par_u, (size_t)B);
  BK_ALLOC(b->ws_state, (size_t)B);
  BK_ALLOC(b->ws_words, (size_t)B * SCAN_WORDS * b->tiles);
  BK_ALLOC(b->idx, (size_t)B);
  BK_ALLOC(b->mask, (size_t)B);
#undef BK_ALLOC
  BIHIP_HIP(hipMemcpy(b->par, b->rows.data(), sizeof(double) * b->rows.size(),
                      hipMemcpyHostToDevice));
  BIHIP_HIP(hipMemset(b->ws_state, 0, sizeof(ScanState) * B));
  BIHIP_HIP(hipMemset(b->ws_words, 0, sizeof(uint64_t) * B * SCAN_WORDS * b->tiles));
  BIHIP_HIP(hipMemset(b->mean0, 0, sizeof(double) * m->d));
  *out = b;
  return BIHIP_OK;
}

int bihip_pfbank_destroy(bihip_pfbank* b) {
  if (b == nullptr) return BIHIP_OK;
  (void)ctx_enter_device(b->ctx);
  (void)hipStreamSynchronize(ctx_stream_of(b->ctx));
  drop_graphs(b);
  for (auto& v : b->ev)
    for (auto e : v) (void)hipEventDestroy(e);
  for (int i = 0; i < 2; ++i) dfree(b->X[i]);
  dfree(b->lws);
  dfree(b->Os);
  dfree(b->as);
  dfree(b->EF);
  dfree(b->partials);
  dfree(b->dev);
  dfree(b->y);
  dfree(b->mean0);
  dfree(b->hist_li);
  dfree(b->hist_ess);
  dfree(b->hist_res);
  dfree(b->hist_anc);
  dfree(b->par);
  dfree(b->par_u);
  dfree(b->ws_state);
  dfree(b->ws_words);
  dfree(b->idx);
  dfree(b->mask);
  delete b;
  return BIHIP_OK;
}

int bihip_pfbank_set_observations(bihip_pfbank* b, const double* y_host) {
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

int bihip_pfbank_init(bihip_pfbank* b, const double* mean0_host) {
  if (b == nullptr) return set_error(BIHIP_EINVAL, "NULL bank");
  int rc = ctx_enter_device(b->ctx);
  if (rc) return rc;
  hipStream_t s = ctx_stream_of(b->ctx);
  if (b->model.model == BIHIP_MODEL_L96) {
    if (mean0_host == nullptr) return set_error(BIHIP_EINVAL, "Lorenz-96 needs mean0");
    BIHIP_HIP(hipMemcpyAsync(b->mean0, mean0_host, sizeof(double) * b->model.d,
                             hipMemcpyHostToDevice, s));
  }
  b->cur = 0;
  b->k = 0;
  if ((rc = launch_init_bank(&b->model, b->cfg.seed, b->cfg.batch, b->cfg.P, b->X[0], b->ld,
                             b->mean0, b->B, (long)b->model.d * b->ld, b->par, s)))
    return rc;
  const size_t n = (size_t)b->B * b->ld, nt = (size_t)b->B * b->cfg.T;
  BIHIP_HIP(hipMemsetAsync(b->lws, 0, sizeof(double) * n, s));
  std::vector<PFDevice> d(b->B);
  for (auto& e : d) {
    e = PFDevice{};
    e.status = BIHIP_OK;
  }
  BIHIP_HIP(hipMemcpyAsync(b->dev, d.data(), sizeof(PFDevice) * b->B, hipMemcpyHostToDevice, s));
  BIHIP_HIP(hipMemsetAsync(b->hist_li, 0, sizeof(double) * nt, s));
  BIHIP_HIP(hipMemsetAsync(b->hist_ess, 0, sizeof(double) * nt, s));
  BIHIP_HIP(hipMemsetAsync(b->hist_res, 0, sizeof(int) * nt, s));
  BIHIP_HIP(hipStreamSynchronize(s));
  for (auto& u : b->ev_used) u = 0;
  b->ready = true;
  return BIHIP_OK;
}

int bihip_pfbank_step(bihip_pfbank* b, int n) {
  if (b == nullptr) return set_error(BIHIP_EINVAL, "NULL bank");
  if (!b->ready) return set_error(BIHIP_ESTATE, "bihip_pfbank_init has not been called");
  if (!b->have_obs) return set_error(BIHIP_ESTATE, "no observations set");
  if (n < 0 || b->k + n > b->cfg.T)
    return set_error(BIHIP_EINVAL, "steps beyond the schedule (k=%d, n=%d, T=%d)", b->k, n,
                     b->cfg.T);
  int rc = ctx_enter_device(b->ctx);
  if (rc) return rc;
  hipStream_t s = ctx_stream_of(b->ctx);
  for (int i = 0; i < n; ++i) {
    if (b->cfg.use_graph && !b->timing) {
      if (b->graph[b->cur] == nullptr && (rc = bank_capture(b, b->cur))) return rc;
      BIHIP_HIP(hipGraphLaunch(b->graph[b->cur], s));
    } else if ((rc = bank_enqueue_step(b, s, b->timing))) {
      return rc;
    }
    b->cur = 1 - b->cur;
    ++b->k;
  }
  return BIHIP_OK;
}

int bihip_pfbank_outputs(bihip_pfbank* b, double* ll_host, double* li_host, double* ess_host,
                         int* res_host, int* status_host) {
  if (b == nullptr) return set_error(BIHIP_EINVAL, "NULL bank");
  int rc = ctx_enter_device(b->ctx);
  if (rc) return rc;
  hipStream_t s = ctx_stream_of(b->ctx);
  BIHIP_HIP(hipStreamSynchronize(s));
  const size_t nt = (size_t)b->B * b->cfg.T;
  if (li_host) BIHIP_HIP(hipMemcpy(li_host, b->hist_li, sizeof(double) * nt, hipMemcpyDeviceToHost));
  if (ess_host) BIHIP_HIP(hipMemcpy(ess_host, b->hist_ess, sizeof(double) * nt, hipMemcpyDeviceToHost));
  if (res_host) BIHIP_HIP(hipMemcpy(res_host, b->hist_res, sizeof(int) * nt, hipMemcpyDeviceToHost));
  if (ll_host || status_host) {
    std::vector<PFDevice> d(b->B);
    BIHIP_HIP(hipMemcpy(d.data(), b->dev, sizeof(PFDevice) * b->B, hipMemcpyDeviceToHost));
    for (int j = 0; j < b->B; ++j) {
      if (ll_host) ll_host[j] = d[j].logLikelihood;
      if (status_host) status_host[j] = d[j].status;
    }
  }
  return BIHIP_OK;
}

int bihip_pfbank_ancestors(bihip_pfbank* b, int j, int k, int* as_host) {
  if (b == nullptr || as_host == nullptr) return set_error(BIHIP_EINVAL, "NULL argument");
  if (b->hist_anc == nullptr) return set_error(BIHIP_ESTATE, "ancestors need record = 1");
  if (j < 0 || j >= b->B || k < 0 || k >= b->k)
    return set_error(BIHIP_EINVAL, "filter %d / step %d out of range", j, k);
  int rc = ctx_enter_device(b->ctx);
  if (rc) return rc;
  BIHIP_HIP(hipStreamSynchronize(ctx_stream_of(b->ctx)));
  const size_t P = (size_t)b->cfg.P;
  BIHIP_HIP(hipMemcpy(as_host, b->hist_anc + ((size_t)j * b->cfg.T + k) * P, sizeof(int) * P,
                      hipMemcpyDeviceToHost));
  return BIHIP_OK;
}

int bihip_pfbank_state(bihip_pfbank* b, int j, void** X_dev, int* ld, double** lws_dev) {
  if (b == nullptr) return set_error(BIHIP_EINVAL, "NULL bank");
  if (j < 0 || j >= b->B) return set_error(BIHIP_EINVAL, "filter %d out of range", j);
  if (X_dev)
    *X_dev = static_cast<uint8_t*>(b->X[b->cur]) + (size_t)j * b->model.d * b->ld * b->elem;
  if (ld) *ld = b->ld;
  if (lws_dev) *lws_dev = b->lws + (size_t)j * b->ld;
  return BIHIP_OK;
}

int bihip_pfbank_position(bihip_pfbank* b) { return b ? b->k : -1; }

int bihip_pfbank_set_surrogate(bihip_pfbank* b, bihip_mlp* mlp, const bihip_surrogate_norm* norm) {
  if (b == nullptr) return set_error(BIHIP_EINVAL, "NULL bank");
  int rc = ctx_enter_device(b->ctx);
  if (rc) return rc;
  if (mlp != nullptr) {
    if (norm == nullptr) return set_error(BIHIP_EINVAL, "NULL normalisation");
    if (b->model.model != BIHIP_MODEL_L96 || b->model.d != 40 || b->model.precision != BIHIP_F32)
      return set_error(BIHIP_EINVAL, "the surrogate flow serves FP32 Lorenz-96 d=40 filters");
    if (b->guided) return set_error(BIHIP_ESTATE, "detach the guided proposal first");
    if ((rc = learned_rows_ok(b))) return rc;
    SurrogateLaunch S{};
    S.net = mlp_view(mlp);
    if (S.net.L != 4 || S.net.dims[0] != 41 || S.net.dims[4] != 40)
      return set_error(BIHIP_EINVAL, "the surrogate must be a 41-256-256-256-40 network");
    if ((rc = mlp_inference_weights(mlp, S.Winf))) return rc;
    S.x_shift = norm->x_shift;
    S.x_scale = norm->x_scale;
    S.dx_scale = norm->dx_scale;
    S.par = b->model.par[0];
    S.par_shift = norm->par_shift;
    S.par_scale = norm->par_scale;
    S.noise_sd = (float)b->model.par[1];
    S.obs_sd = b->model.par[2];
    if ((rc = set_par_u(b, norm->par_shift, norm->par_scale))) return rc;
    b->sl = S;
  }
  BIHIP_HIP(hipStreamSynchronize(ctx_stream_of(b->ctx)));
  drop_graphs(b);
  b->surr = mlp != nullptr;
  b->nparts = b->surr ? surrogate_blocks(b->cfg.P) : transition_blocks(&b->model, b->cfg.P);
  return BIHIP_OK;
}

int bihip_pfbank_set_guided(bihip_pfbank* b, bihip_mlp* mlp, const bihip_guided_norm* norm) {
  if (b == nullptr) return set_error(BIHIP_EINVAL, "NULL bank");
  int rc = ctx_enter_device(b->ctx);
  if (rc) return rc;
  if (mlp != nullptr) {
    if (norm == nullptr) return set_error(BIHIP_EINVAL, "NULL normalisation");
    if (b->model.model != BIHIP_MODEL_L96 || b->model.d != 200 ||
        b->model.precision != BIHIP_F32)
      return set_error(BIHIP_EINVAL, "the guided proposal serves FP32 Lorenz-96 d=200 filters");
    if (b->surr) return set_error(BIHIP_ESTATE, "detach the surrogate flow first");
    if ((rc = learned_rows_ok(b))) return rc;
    GuidedLaunch G{};
    G.net = mlp_view(mlp);
    if (G.net.L != 4 || G.net.dims[0] != 301 || G.net.dims[4] != 400)
      return set_error(BIHIP_EINVAL, "the proposal must be a 301-256-256-256-400 network");
    if ((rc = mlp_inference_weights(mlp, G.Winf, G.binf, 16))) return rc;
    G.x_shift = norm->x_shift;
    G.x_scale = norm->x_scale;
    G.innov_scale = norm->innov_scale;
    G.par = b->model.par[0];
    G.par_shift = norm->par_shift;
    G.par_scale = norm->par_scale;
    G.noise_sd = (float)b->model.par[1];
    G.obs_sd = b->model.par[2];
    if ((rc = set_par_u(b, norm->par_shift, norm->par_scale))) return rc;
    b->gl = G;
  }
  BIHIP_HIP(hipStreamSynchronize(ctx_stream_of(b->ctx)));
  drop_graphs(b);
  b->guided = mlp != nullptr;
  b->nparts = b->guided ? guided_blocks(b->cfg.P) : transition_blocks(&b->model, b->cfg.P);
  return BIHIP_OK;
}

int bihip_pfbank_set_theta(bihip_pfbank* b, const double* theta_host) {
  if (b == nullptr || (b->ntheta > 0 && theta_host == nullptr))
    return set_error(BIHIP_EINVAL, "NULL argument");
  int rc = ctx_enter_device(b->ctx);
  if (rc) return rc;
  BIHIP_HIP(hipStreamSynchronize(ctx_stream_of(b->ctx)));
  std::vector<double> keep = b->rows;
  if ((rc = upload_rows(b, theta_host))) {
    b->rows = keep;
    return rc;
  }
  if ((b->guided || b->surr) && (rc = learned_rows_ok(b))) {
    b->rows = keep;
    return rc;
  }
  return sync_rows(b);
}

int bihip_pfbank_set_stream_base(bihip_pfbank* b, uint32_t base) {
  if (b == nullptr) return set_error(BIHIP_EINVAL, "NULL bank");
  if ((uint64_t)base + (uint64_t)b->B > (1u << 24))
    return set_error(BIHIP_EINVAL, "Philox batch ids must stay below 2^24");
  int rc = ctx_enter_device(b->ctx);
  if (rc) return rc;
  BIHIP_HIP(hipStreamSynchronize(ctx_stream_of(b->ctx)));
  if (b->cfg.batch != base) drop_graphs(b);  // the captured launches hold the batch id
  b->cfg.batch = base;
  return BIHIP_OK;
}

int bihip_pfbank_increments(bihip_pfbank* b, int k, double* li_host) {
  if (b == nullptr || li_host == nullptr) return set_error(BIHIP_EINVAL, "NULL argument");
  if (k < 0 || k >= b->cfg.T) return set_error(BIHIP_EINVAL, "step %d out of range", k);
  int rc = ctx_enter_device(b->ctx);
  if (rc) return rc;
  BIHIP_HIP(hipStreamSynchronize(ctx_stream_of(b->ctx)));
  BIHIP_HIP(hipMemcpy2D(li_host, sizeof(double), b->hist_li + k, sizeof(double) * b->cfg.T,
                        sizeof(double), b->B, hipMemcpyDeviceToHost));
  return BIHIP_OK;
}

int bihip_pfbank_resample_filters(bihip_pfbank* b, const int* anc_host) {
  if (b == nullptr || anc_host == nullptr) return set_error(BIHIP_EINVAL, "NULL argument");
  if (!b->ready) return set_error(BIHIP_ESTATE, "bihip_pfbank_init has not been called");
  for (int j = 0; j < b->B; ++j)
    if (anc_host[j] < 0 || anc_host[j] >= b->B)
      return set_error(BIHIP_EINVAL, "ancestor %d of filter %d out of range", anc_host[j], j);
  int rc = ctx_enter_device(b->ctx);
  if (rc) return rc;
  hipStream_t s = ctx_stream_of(b->ctx);
  BIHIP_HIP(hipMemcpyAsync(b->idx, anc_host, sizeof(int) * b->B, hipMemcpyHostToDevice, s));
  const long xb = (long)b->model.d * b->ld * b->elem, lb = (long)b->ld * sizeof(double);
  // particles into the other state buffer, log-weights through the EF scratch
  if ((rc = launch_filter_copy(b->X[b->cur], b->X[1 - b->cur], xb, b->B, b->idx, nullptr, s)) ||
      (rc = launch_filter_copy(b->lws, b->EF, lb, b->B, b->idx, nullptr, s)) ||
      (rc = launch_filter_copy(b->EF, b->lws, lb, b->B, nullptr, nullptr, s)))
    return rc;
  std::vector<PFDevice> d(b->B), d2(b->B);
  BIHIP_HIP(hipMemcpyAsync(d.data(), b->dev, sizeof(PFDevice) * b->B, hipMemcpyDeviceToHost, s));
  BIHIP_HIP(hipStreamSynchronize(s));
  std::vector<double> r2(b->rows.size());
  for (int j = 0; j < b->B; ++j) {
    d2[j] = d[anc_host[j]];
    std::copy(&b->rows[(size_t)anc_host[j] * BANK_ROW],
              &b->rows[(size_t)anc_host[j] * BANK_ROW] + BANK_ROW, &r2[(size_t)j * BANK_ROW]);
  }
  BIHIP_HIP(hipMemcpy(b->dev, d2.data(), sizeof(PFDevice) * b->B, hipMemcpyHostToDevice));
  b->rows = r2;
  b->cur = 1 - b->cur;
  return sync_rows(b);
}

int bihip_pfbank_accept(bihip_pfbank* dst, const bihip_pfbank* src, const int* accept_host) {
  if (dst == nullptr || src == nullptr || accept_host == nullptr)
    return set_error(BIHIP_EINVAL, "NULL argument");
  if (!same_shape(dst, src)) return set_error(BIHIP_EINVAL, "banks of different shapes");
  if (!dst->ready || !src->ready || dst->k != src->k)
    return set_error(BIHIP_ESTATE, "both banks must be initialised and at the same step");
  int rc = ctx_enter_device(dst->ctx);
  if (rc) return rc;
  hipStream_t s = ctx_stream_of(dst->ctx);
  BIHIP_HIP(hipMemcpyAsync(dst->mask, accept_host, sizeof(int) * dst->B, hipMemcpyHostToDevice,
                           s));
  const long xb = (long)dst->model.d * dst->ld * dst->elem, lb = (long)dst->ld * sizeof(double);
  if ((rc = launch_filter_copy(src->X[src->cur], dst->X[dst->cur], xb, dst->B, nullptr,
                               dst->mask, s)) ||
      (rc = launch_filter_copy(src->lws, dst->lws, lb, dst->B, nullptr, dst->mask, s)))
    return rc;
  std::vector<PFDevice> d(dst->B), e(dst->B);
  BIHIP_HIP(hipMemcpyAsync(d.data(), dst->dev, sizeof(PFDevice) * dst->B, hipMemcpyDeviceToHost,
                           s));
  BIHIP_HIP(hipMemcpyAsync(e.data(), src->dev, sizeof(PFDevice) * dst->B, hipMemcpyDeviceToHost,
                           s));
  BIHIP_HIP(hipStreamSynchronize(s));
  for (int j = 0; j < dst->B; ++j) {
    if (!accept_host[j]) continue;
    d[j] = e[j];
    std::copy(&src->rows[(size_t)j * BANK_ROW], &src->rows[(size_t)j * BANK_ROW] + BANK_ROW,
              &dst->rows[(size_t)j * BANK_ROW]);
  }
  BIHIP_HIP(hipMemcpy(dst->dev, d.data(), sizeof(PFDevice) * dst->B, hipMemcpyHostToDevice));
  return sync_rows(dst);
}

int bihip_pfbank_set_timing(bihip_pfbank* b, int mode) {
  if (b == nullptr) return set_error(BIHIP_EINVAL, "NULL bank");
  if (mode != 0 && mode != BIHIP_TIMING_ALL && mode != BIHIP_TIMING_TRANSITION)
    return set_error(BIHIP_EINVAL, "unknown timing mode %d", mode);
  b->timing = mode;
  for (auto& u : b->ev_used) u = 0;
  return BIHIP_OK;
}

int bihip_pfbank_kernel_time(bihip_pfbank* b, int kind, double* ms, long* count) {
  if (b == nullptr || kind < 0 || kind > 3) return set_error(BIHIP_EINVAL, "bad argument");
  int rc = ctx_enter_device(b->ctx);
  if (rc) return rc;
  BIHIP_HIP(hipStreamSynchronize(ctx_stream_of(b->ctx)));
  double total = 0.0;
  const size_t used = b->ev_used[kind];
  for (size_t i = 0; i + 1 < used; i += 2) {
    float t = 0.f;
    BIHIP_HIP(hipEventElapsedTime(&t, b->ev[kind][i], b->ev[kind][i + 1]));
    total += t;
  }
  if (ms) *ms = total;
  if (count) *count = (long)(used / 2);
  return BIHIP_OK;
}

}  // extern "C"
