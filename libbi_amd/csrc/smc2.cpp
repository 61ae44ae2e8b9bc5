// SMC^2 and PMMH over the batched inner filters (bihip_smc2_*, bihip_pmmh_*,
// declared in include/bihip.h).
//
// MarginalSIR::sample (share/src/bi/sampler/MarginalSIR.hpp:281-320): init
// (:322-344), then for every observation a move (:395-488), a step of every
// inner filter (:346-368) and the outer interaction (:370-393), and a final
// move after the last observation.  The inner filters are one bihip_batch
// (one persistent launch per step for all theta-particles); the proposal
// filters of a move are a second batch run from the start of the schedule.
// The outer level (B log-weights, systematic resampling of whole filters,
// the Metropolis-Hastings decisions) is host arithmetic in FP64, as the
// reference's host vectors of theta-particles are.
//
// theta lives in log space: parameter i of slot j is exp(lth[j][i]), with a
// Gaussian prior N(prior_mean[i], prior_sd[i]^2) on lth[j][i] and a Gaussian
// random-walk proposal of scale prop_sd[i], so prior and proposal Jacobians
// cancel in the acceptance ratio.  Keyed draws: prior normals
// {j, 0, THETA, 0}; proposal normals {j, k, PROPOSAL, m}; acceptance uniform
// {j, k, ACCEPT, m}; outer systematic offset {0, k, OUTER, 0}.  Slot j's
// filter draws with batch id j; the proposal filters of move m at time k with
// batch ids B (1 + m + nmoves k) + j.
#include <cmath>
#include <vector>

#include "internal.hpp"

#pragma clang fp contract(off)

using namespace bihip;

namespace {

// The inner filters of the B theta-particles: one bihip_batch (ragged N_j up to
// 16384, one workgroup per filter) or one filter bank (equal N of any size,
// shared launches).  Both equal bihip_pf with batch id stream_base + j.
struct Inner {
  virtual ~Inner() = default;
  virtual int set_observations(const double* y) = 0;
  virtual int set_sizes(const int* N) = 0;
  virtual int set_theta(const double* theta) = 0;
  virtual int set_stream_base(uint32_t base) = 0;
  virtual int init() = 0;
  virtual int step(int n) = 0;
  virtual int outputs(double* ll, int* status) = 0;
  virtual int increments(int k, double* li) = 0;
  virtual int resample_filters(const int* anc) = 0;
  virtual int accept(Inner* src, const int* acc) = 0;
};

struct BatchInner : Inner {
  bihip_batch* b = nullptr;
  ~BatchInner() override { bihip_batch_destroy(b); }
  int set_observations(const double* y) override { return bihip_batch_set_observations(b, y); }
  int set_sizes(const int* N) override { return bihip_batch_set_sizes(b, N); }
  int set_theta(const double* th) override { return bihip_batch_set_theta(b, th); }
  int set_stream_base(uint32_t base) override { return bihip_batch_set_stream_base(b, base); }
  int init() override { return bihip_batch_init(b); }
  int step(int n) override { return bihip_batch_step(b, n); }
  int outputs(double* ll, int* st) override {
    return bihip_batch_outputs(b, ll, nullptr, nullptr, nullptr, st);
  }
  int increments(int k, double* li) override { return bihip_batch_increments(b, k, li); }
  int resample_filters(const int* anc) override { return bihip_batch_resample_filters(b, anc); }
  int accept(Inner* src, const int* acc) override {
    return bihip_batch_accept(b, static_cast<BatchInner*>(src)->b, acc);
  }
};

struct BankInner : Inner {
  bihip_pfbank* b = nullptr;
  int P = 0;
  std::vector<double> mean0;
  ~BankInner() override { bihip_pfbank_destroy(b); }
  int set_observations(const double* y) override { return bihip_pfbank_set_observations(b, y); }
  int set_sizes(const int* N) override { return BIHIP_OK; }  // every filter keeps P
  int set_theta(const double* th) override { return bihip_pfbank_set_theta(b, th); }
  int set_stream_base(uint32_t base) override { return bihip_pfbank_set_stream_base(b, base); }
  int init() override { return bihip_pfbank_init(b, mean0.empty() ? nullptr : mean0.data()); }
  int step(int n) override { return bihip_pfbank_step(b, n); }
  int outputs(double* ll, int* st) override {
    return bihip_pfbank_outputs(b, ll, nullptr, nullptr, nullptr, st);
  }
  int increments(int k, double* li) override { return bihip_pfbank_increments(b, k, li); }
  int resample_filters(const int* anc) override { return bihip_pfbank_resample_filters(b, anc); }
  int accept(Inner* src, const int* acc) override {
    return bihip_pfbank_accept(b, static_cast<BankInner*>(src)->b, acc);
  }
};

}  // namespace

struct bihip_smc2 {
  bihip_ctx* ctx = nullptr;
  bihip_model model{};
  bihip_smc2_config cfg{};
  Inner* slots = nullptr;
  Inner* prop = nullptr;
  bool bank = false;
  std::vector<int> N0, N;  // sizes at creation, current sizes per slot
  std::vector<double> lth, lth2, lw, theta;
  std::vector<int> accepts, resampled;
  std::vector<double> ess;
  double logZ = 0.0;
  bool have_obs = false, done = false;
};

namespace {

constexpr double TWO_PI = 6.283185307179586476925286766559;

double host_u53(uint32_t r0, uint32_t r1) {
  const uint64_t m = ((((uint64_t)r0) << 32) | r1) >> 11;
  return (double)m * 0x1.0p-53;
}

// n standard normals at {index, step, tag|batch<<8, block}: Box-Muller in
// FP64 on the 32-bit word pairs (0,1) and (2,3) of each Philox block, the
// first uniform 1 - r0 2^-32 in (0,1], the second r1 2^-32.
void host_normals(uint64_t seed, uint32_t index, uint32_t step, uint32_t tag, uint32_t batch,
                  int n, double* z) {
  for (int b = 0; 4 * b < n; ++b) {
    uint32_t r[4];
    philox_host_block(seed, index, step, tag, batch, (uint32_t)b, r);
    for (int h = 0; h < 2; ++h) {
      const double u1 = 1.0 - (double)r[2 * h] * 0x1.0p-32;
      const double u2 = (double)r[2 * h + 1] * 0x1.0p-32;
      const double rad = std::sqrt(-2.0 * std::log(u1));
      const double ang = TWO_PI * u2;
      if (4 * b + 2 * h < n) z[4 * b + 2 * h] = rad * std::cos(ang);
      if (4 * b + 2 * h + 1 < n) z[4 * b + 2 * h + 1] = rad * std::sin(ang);
    }
  }
}

double host_uniform(uint64_t seed, uint32_t index, uint32_t step, uint32_t tag, uint32_t batch) {
  uint32_t r[4];
  philox_host_block(seed, index, step, tag, batch, 0u, r);
  return host_u53(r[0], r[1]);
}

double nanexp_h(double x) { return std::isnan(x) ? 0.0 : std::exp(x); }

// max with NaN below every number (vector_primitive.hpp:860-872)
double nan_max_h(const std::vector<double>& x) {
  double mx = x[0];
  for (size_t i = 1; i < x.size(); ++i) {
    const bool less = std::isnan(mx) ? !std::isnan(x[i]) : (!std::isnan(x[i]) && mx < x[i]);
    if (less) mx = x[i];
  }
  return mx;
}

// ess_reduce over the outer log-weights, left-to-right sums
// (vector_primitive.hpp:959-974); *lW = log mean weight.
double outer_ess(const std::vector<double>& lw, double* lW) {
  const double mx = nan_max_h(lw);
  double s1 = 0.0, s2 = 0.0;
  for (double v : lw) {
    const double w = nanexp_h(v - mx);
    s1 = s1 + w;
    s2 = s2 + w * w;
  }
  *lW = mx + std::log(s1) - std::log((double)lw.size());
  return s1 * s1 / s2;
}

// Systematic resampling of the B slots (SystematicResampler.hpp:110-130)
// into ancestors in the permuted layout: a slot with offspring keeps itself,
// the extra copies fill the slots without offspring in ascending order.
bool outer_ancestors(const std::vector<double>& lw, double a, std::vector<int>& as) {
  const int B = (int)lw.size();
  const double mx = nan_max_h(lw);
  std::vector<double> Ws(B);
  double s = 0.0;
  for (int i = 0; i < B; ++i) {
    s = s + nanexp_h(lw[i] - mx);
    Ws[i] = s;
  }
  const double W = Ws[B - 1];
  if (!(W > 0.0)) return false;
  std::vector<int> o(B);
  int prev = 0;
  for (int i = 0; i < B; ++i) {
    const double v = Ws[i] / W * (double)B + a;
    int O = (int)v;
    if (O > B) O = B;
    o[i] = O - prev;
    prev = O;
  }
  as.assign(B, -1);
  std::vector<int> free_slots;
  for (int i = 0; i < B; ++i) {
    if (o[i] > 0) as[i] = i;
    else free_slots.push_back(i);
  }
  size_t next = 0;
  for (int i = 0; i < B; ++i)
    for (int c = 1; c < o[i]; ++c) as[free_slots[next++]] = i;
  return true;
}

// log prior density of lth (up to a constant): the Gaussian on lth, or the
// model file's prior of theta = exp(lth) times the Jacobian exp(sum lth)
double log_prior(const bihip_model& m, const bihip_smc2_config& c, const double* lth) {
  double lp = 0.0;
  if (c.prior_from_model) {
    double th[8];
    for (int i = 0; i < c.ntheta; ++i) {
      th[i] = std::exp(lth[i]);
      lp += lth[i];
    }
    return prior_logdensity(&m, c.ntheta, th) + lp;
  }
  for (int i = 0; i < c.ntheta; ++i) {
    const double z = (lth[i] - c.prior_mean[i]) / c.prior_sd[i];
    lp += -0.5 * z * z;
  }
  return lp;
}

// theta-particle j's draw from the prior, as lth
void prior_lth(const bihip_model& m, const bihip_smc2_config& c, int j, double* lth) {
  const int nt = c.ntheta;
  if (c.prior_from_model) {
    double th[8];
    prior_sample(&m, nt, c.seed, (uint32_t)j, th);
    for (int i = 0; i < nt; ++i) lth[i] = std::log(th[i]);
    return;
  }
  double z[8];
  host_normals(c.seed, (uint32_t)j, 0u, BIHIP_TAG_THETA, 0u, nt, z);
  for (int i = 0; i < nt; ++i) lth[i] = c.prior_mean[i] + c.prior_sd[i] * z[i];
}

int check_prior(const bihip_model* m, const bihip_smc2_config& c) {
  if (c.prior_from_model) {
    if (int rc = prior_check(m, c.ntheta, true)) return rc;
    for (int i = 0; i < c.ntheta; ++i)
      if (!(c.prop_sd[i] >= 0.0)) return set_error(BIHIP_EINVAL, "prop_sd must be non-negative");
    return BIHIP_OK;
  }
  for (int i = 0; i < c.ntheta; ++i)
    if (!(c.prior_sd[i] > 0.0) || !(c.prop_sd[i] >= 0.0))
      return set_error(BIHIP_EINVAL, "prior_sd must be positive and prop_sd non-negative");
  return BIHIP_OK;
}

int filter_logliks(Inner* b, int B, std::vector<double>& ll) {
  std::vector<int> st(B);
  ll.resize(B);
  int rc = b->outputs(ll.data(), st.data());
  if (rc) return rc;
  for (int j = 0; j < B; ++j)
    if (st[j] != BIHIP_OK) ll[j] = -INFINITY;
  return BIHIP_OK;
}

// PMMH move of every slot (MarginalSIR.hpp:395-488); returns accepts in *nacc.
int smc2_move(bihip_smc2* s, int k, int* nacc) {
  const bihip_smc2_config& c = s->cfg;
  const int B = c.B, nt = c.ntheta;
  int rc;
  *nacc = 0;
  std::vector<double> ll1, ll2, th((size_t)B * nt);
  std::vector<int> acc(B);
  for (int m = 0; m < c.nmoves; ++m) {
    for (int j = 0; j < B; ++j) {
      double z[8];
      host_normals(c.seed, (uint32_t)j, (uint32_t)k, BIHIP_TAG_PROPOSAL, (uint32_t)m, nt, z);
      for (int i = 0; i < nt; ++i) {
        const size_t q = (size_t)j * nt + i;
        s->lth2[q] = s->lth[q] + c.prop_sd[i] * z[i];
        th[q] = std::exp(s->lth2[q]);
      }
    }
    const uint32_t base = (uint32_t)(B * (1 + m + c.nmoves * k));
    if ((rc = s->prop->set_sizes(s->N.data())) || (rc = s->prop->set_theta(th.data())) ||
        (rc = s->prop->set_stream_base(base)) || (rc = s->prop->init()) ||
        (rc = s->prop->step(k)) || (rc = filter_logliks(s->prop, B, ll2)) ||
        (rc = filter_logliks(s->slots, B, ll1)))
      return rc;
    for (int j = 0; j < B; ++j) {
      bool accept;
      if (!std::isfinite(ll2[j])) {
        accept = false;
      } else if (!std::isfinite(ll1[j])) {
        accept = true;
      } else {
        const double u =
            host_uniform(c.seed, (uint32_t)j, (uint32_t)k, BIHIP_TAG_ACCEPT, (uint32_t)m);
        const double logratio = ll2[j] - ll1[j] + log_prior(s->model, c, &s->lth2[(size_t)j * nt]) -
                                log_prior(s->model, c, &s->lth[(size_t)j * nt]);
        accept = std::log(u) < logratio;
      }
      acc[j] = accept ? 1 : 0;
      if (accept) {
        for (int i = 0; i < nt; ++i) s->lth[(size_t)j * nt + i] = s->lth2[(size_t)j * nt + i];
        ++*nacc;
      }
    }
    if ((rc = s->slots->accept(s->prop, acc.data()))) return rc;
  }
  return BIHIP_OK;
}

}  // namespace

extern "C" {

static int smc2_check(const bihip_model* m, const bihip_smc2_config& c) {
  if (c.B <= 0 || c.T <= 0) return set_error(BIHIP_EINVAL, "need B > 0 and T > 0");
  if (c.ntheta < 1 || c.ntheta > 8) return set_error(BIHIP_EINVAL, "ntheta must lie in [1,8]");
  if (c.nmoves < 0) return set_error(BIHIP_EINVAL, "nmoves must be >= 0");
  if (!(c.essRel_outer >= 0.0 && c.essRel_outer <= 1.0))
    return set_error(BIHIP_EINVAL, "essRel_outer must lie in [0,1]");
  if (int rc = check_prior(m, c)) return rc;
  if ((uint64_t)c.B * (1 + (uint64_t)c.nmoves * (c.T + 1)) > (1u << 24))
    return set_error(BIHIP_EINVAL, "B (1 + nmoves (T + 1)) random-stream ids exceed 2^24");
  return BIHIP_OK;
}

static bihip_smc2* smc2_new(bihip_ctx* ctx, const bihip_model* m, const bihip_smc2_config& c) {
  bihip_smc2* s = new bihip_smc2();
  s->ctx = ctx;
  s->model = *m;
  s->cfg = c;
  s->lth.assign((size_t)c.B * c.ntheta, 0.0);
  s->lth2 = s->lth;
  s->theta = s->lth;
  s->lw.assign(c.B, 0.0);
  s->accepts.assign(c.T + 1, 0);
  s->resampled.assign(c.T, 0);
  s->ess.assign(c.T, 0.0);
  return s;
}

int bihip_smc2_create(bihip_ctx* ctx, const bihip_model* m, const bihip_smc2_config* cfg,
                      const int* N_host, bihip_smc2** out) {
  if (ctx == nullptr || m == nullptr || cfg == nullptr || N_host == nullptr || out == nullptr)
    return set_error(BIHIP_EINVAL, "NULL argument");
  const bihip_smc2_config& c = *cfg;
  if (int rc = smc2_check(m, c)) return rc;
  int maxN = 0;
  for (int j = 0; j < c.B; ++j) maxN = N_host[j] > maxN ? N_host[j] : maxN;
  bihip_smc2* s = smc2_new(ctx, m, c);
  s->N0.assign(N_host, N_host + c.B);
  s->N = s->N0;
  bihip_batch_config bc{};
  bc.B = c.B;
  bc.T = c.T;
  bc.resampler = c.resampler;
  bc.essRel = c.essRel_inner;
  bc.t0 = c.t0;
  bc.dt = c.dt;
  bc.seed = c.seed;
  bc.record = 0;
  bc.use_graph = c.use_graph;
  bc.block = c.block;
  bc.capacity = (long)c.B * maxN;  // outer resampling can repeat the largest filter
  bc.stream_base = 0;
  auto* slots = new BatchInner();
  auto* prop = new BatchInner();
  s->slots = slots;
  s->prop = prop;
  int rc = bihip_batch_create(ctx, m, &bc, N_host, s->theta.data(), c.ntheta, &slots->b);
  if (rc == BIHIP_OK)
    rc = bihip_batch_create(ctx, m, &bc, N_host, s->theta.data(), c.ntheta, &prop->b);
  if (rc != BIHIP_OK) {
    bihip_smc2_destroy(s);
    return rc;
  }
  *out = s;
  return BIHIP_OK;
}

int bihip_smc2_create_bank(bihip_ctx* ctx, const bihip_model* m, const bihip_smc2_config* cfg,
                           int P, const double* mean0_host, bihip_smc2** out) {
  if (ctx == nullptr || m == nullptr || cfg == nullptr || out == nullptr)
    return set_error(BIHIP_EINVAL, "NULL argument");
  const bihip_smc2_config& c = *cfg;
  if (int rc = smc2_check(m, c)) return rc;
  if (P <= 0) return set_error(BIHIP_EINVAL, "need P > 0");
  if (m->model == BIHIP_MODEL_L96 && mean0_host == nullptr)
    return set_error(BIHIP_EINVAL, "Lorenz-96 needs mean0");
  bihip_smc2* s = smc2_new(ctx, m, c);
  s->bank = true;
  s->N0.assign(c.B, P);
  s->N = s->N0;
  bihip_pf_config pc{};
  pc.P = P;
  pc.T = c.T;
  pc.resampler = c.resampler;
  pc.essRel = c.essRel_inner;
  pc.t0 = c.t0;
  pc.dt = c.dt;
  pc.seed = c.seed;
  pc.batch = 0;
  pc.record = 0;
  pc.use_graph = c.use_graph;
  auto* slots = new BankInner();
  auto* prop = new BankInner();
  s->slots = slots;
  s->prop = prop;
  for (BankInner* b : {slots, prop}) {
    b->P = P;
    if (mean0_host != nullptr) b->mean0.assign(mean0_host, mean0_host + m->d);
  }
  int rc = bihip_pfbank_create(ctx, m, &pc, c.B, s->theta.data(), c.ntheta, &slots->b);
  if (rc == BIHIP_OK)
    rc = bihip_pfbank_create(ctx, m, &pc, c.B, s->theta.data(), c.ntheta, &prop->b);
  if (rc != BIHIP_OK) {
    bihip_smc2_destroy(s);
    return rc;
  }
  *out = s;
  return BIHIP_OK;
}

int bihip_smc2_set_surrogate(bihip_smc2* s, bihip_mlp* mlp, const bihip_surrogate_norm* norm) {
  if (s == nullptr) return set_error(BIHIP_EINVAL, "NULL sampler");
  if (!s->bank) return set_error(BIHIP_ESTATE, "the surrogate flow runs in filter banks");
  int rc = bihip_pfbank_set_surrogate(static_cast<BankInner*>(s->slots)->b, mlp, norm);
  if (rc == BIHIP_OK) rc = bihip_pfbank_set_surrogate(static_cast<BankInner*>(s->prop)->b, mlp, norm);
  return rc;
}

int bihip_smc2_set_guided(bihip_smc2* s, bihip_mlp* mlp, const bihip_guided_norm* norm) {
  if (s == nullptr) return set_error(BIHIP_EINVAL, "NULL sampler");
  if (!s->bank) return set_error(BIHIP_ESTATE, "the guided proposal runs in filter banks");
  int rc = bihip_pfbank_set_guided(static_cast<BankInner*>(s->slots)->b, mlp, norm);
  if (rc == BIHIP_OK) rc = bihip_pfbank_set_guided(static_cast<BankInner*>(s->prop)->b, mlp, norm);
  return rc;
}

int bihip_smc2_destroy(bihip_smc2* s) {
  if (s == nullptr) return BIHIP_OK;
  delete s->prop;
  delete s->slots;
  delete s;
  return BIHIP_OK;
}

int bihip_smc2_set_observations(bihip_smc2* s, const double* y_host) {
  if (s == nullptr || y_host == nullptr) return set_error(BIHIP_EINVAL, "NULL argument");
  int rc = s->slots->set_observations(y_host);
  if (rc == BIHIP_OK) rc = s->prop->set_observations(y_host);
  s->have_obs = rc == BIHIP_OK;
  return rc;
}

int bihip_smc2_run(bihip_smc2* s) {
  if (s == nullptr) return set_error(BIHIP_EINVAL, "NULL sampler");
  if (!s->have_obs) return set_error(BIHIP_ESTATE, "no observations set");
  const bihip_smc2_config& c = s->cfg;
  const int B = c.B, T = c.T, nt = c.ntheta;
  int rc;
  // init (:322-344): theta from the prior, every inner filter from its initial law
  for (int j = 0; j < B; ++j) {
    prior_lth(s->model, c, j, &s->lth[(size_t)j * nt]);
    for (int i = 0; i < nt; ++i) {
      const size_t q = (size_t)j * nt + i;
      s->theta[q] = std::exp(s->lth[q]);
    }
  }
  s->N = s->N0;  // a second run starts from the sizes given at creation
  if ((rc = s->slots->set_sizes(s->N.data())) || (rc = s->slots->set_theta(s->theta.data())) ||
      (rc = s->slots->set_stream_base(0u)) || (rc = s->slots->init()))
    return rc;
  std::fill(s->lw.begin(), s->lw.end(), 0.0);
  std::fill(s->accepts.begin(), s->accepts.end(), 0);
  std::fill(s->resampled.begin(), s->resampled.end(), 0);
  std::fill(s->ess.begin(), s->ess.end(), 0.0);
  s->logZ = 0.0;
  bool lastResample = false;
  std::vector<double> li(B);
  std::vector<int> st(B), as;
  for (int k = 0; k <= T; ++k) {
    int nacc = 0;
    if (lastResample && (rc = smc2_move(s, k, &nacc))) return rc;
    s->accepts[k] = nacc;
    if (k == T) break;  // the final move (:301-303) follows the last observation
    // step (:346-368): every inner filter, lw += logIncrement
    if ((rc = s->slots->step(1)) || (rc = s->slots->increments(k, li.data())) ||
        (rc = s->slots->outputs(nullptr, st.data())))
      return rc;
    for (int j = 0; j < B; ++j) s->lw[j] += st[j] == BIHIP_OK ? li[j] : -INFINITY;
    // interact (:370-393): outer ESS and marginal likelihood, then resampling
    double lW;
    const double e = outer_ess(s->lw, &lW);
    s->ess[k] = e;
    s->logZ = lW;
    const bool r = e < c.essRel_outer * B;
    s->resampled[k] = r ? 1 : 0;
    if (r) {
      const double a = host_uniform(c.seed, 0u, (uint32_t)k, BIHIP_TAG_OUTER, 0u);
      if (!outer_ancestors(s->lw, a, as))
        return set_error(BIHIP_DEGENERATE, "outer weights are all zero at step %d", k);
      if ((rc = s->slots->resample_filters(as.data()))) return rc;
      std::vector<double> l2(s->lth.size());
      std::vector<int> N2(B);
      for (int j = 0; j < B; ++j) {
        for (int i = 0; i < nt; ++i) l2[(size_t)j * nt + i] = s->lth[(size_t)as[j] * nt + i];
        N2[j] = s->N[as[j]];
      }
      s->lth = l2;
      s->N = N2;
      std::fill(s->lw.begin(), s->lw.end(), lW);
    }
    lastResample = r;
  }
  for (size_t q = 0; q < s->lth.size(); ++q) s->theta[q] = std::exp(s->lth[q]);
  s->done = true;
  return BIHIP_OK;
}

int bihip_smc2_outputs(bihip_smc2* s, double* logZ, double* theta_host, double* lw_host,
                       int* accepts_host, double* ess_host, int* resampled_host) {
  if (s == nullptr) return set_error(BIHIP_EINVAL, "NULL sampler");
  if (!s->done) return set_error(BIHIP_ESTATE, "bihip_smc2_run has not completed");
  if (logZ) *logZ = s->logZ;
  if (theta_host) std::copy(s->theta.begin(), s->theta.end(), theta_host);
  if (lw_host) std::copy(s->lw.begin(), s->lw.end(), lw_host);
  if (accepts_host) std::copy(s->accepts.begin(), s->accepts.end(), accepts_host);
  if (ess_host) std::copy(s->ess.begin(), s->ess.end(), ess_host);
  if (resampled_host) std::copy(s->resampled.begin(), s->resampled.end(), resampled_host);
  return BIHIP_OK;
}

// ---------------------------------------------------------------- PMMH
// MarginalMH::sample (share/src/bi/sampler/MarginalMH.hpp:198-266) for B
// independent chains: every iteration filters all B proposals together (one
// batch: a persistent launch per observation; or one filter bank: shared
// launches per step) and accepts on the host.  Keyed draws as SMC^2's: prior
// normals {j, 0, THETA, 0}, proposal normals {j, c, PROPOSAL, 0}, acceptance
// uniform {j, c, ACCEPT, 0}; the filter of iteration c of chain j draws with
// batch id B c + j.

}  // extern "C"

struct bihip_pmmh {
  bihip_model model{};
  bihip_smc2_config cfg{};
  int C = 0;
  Inner* inner = nullptr;  // the B proposals' filters
  bool bank = false;
  std::vector<double> theta, ll;  // C x B x ntheta, C x B
  std::vector<int> accepted;
  bool have_obs = false, done = false;
};

namespace {
int pmmh_filter(bihip_pmmh* s, const std::vector<double>& lth, uint32_t base,
                std::vector<double>& ll) {
  const int B = s->cfg.B, nt = s->cfg.ntheta;
  std::vector<double> th((size_t)B * nt);
  for (size_t q = 0; q < th.size(); ++q) th[q] = std::exp(lth[q]);
  int rc;
  if ((rc = s->inner->set_theta(th.data())) || (rc = s->inner->set_stream_base(base)) ||
      (rc = s->inner->init()) || (rc = s->inner->step(s->cfg.T)))
    return rc;
  std::vector<int> st(B);
  ll.resize(B);
  if ((rc = s->inner->outputs(ll.data(), st.data()))) return rc;
  for (int j = 0; j < B; ++j)
    if (st[j] != BIHIP_OK) ll[j] = -INFINITY;
  return BIHIP_OK;
}
}  // namespace

extern "C" {

static int pmmh_check(const bihip_model* m, const bihip_smc2_config& c, int C) {
  if (c.B <= 0 || c.T <= 0 || C <= 0) return set_error(BIHIP_EINVAL, "need B, T, C > 0");
  if (c.ntheta < 1 || c.ntheta > 8) return set_error(BIHIP_EINVAL, "ntheta must lie in [1,8]");
  if (int rc = check_prior(m, c)) return rc;
  if ((uint64_t)c.B * ((uint64_t)C + 1) > (1u << 24))
    return set_error(BIHIP_EINVAL, "B (C + 1) random-stream ids exceed 2^24");
  return BIHIP_OK;
}

static bihip_pmmh* pmmh_new(const bihip_model* m, const bihip_smc2_config& c, int C) {
  bihip_pmmh* s = new bihip_pmmh();
  s->model = *m;
  s->cfg = c;
  s->C = C;
  s->theta.assign((size_t)C * c.B * c.ntheta, 0.0);
  s->ll.assign((size_t)C * c.B, 0.0);
  s->accepted.assign(c.B, 0);
  return s;
}

int bihip_pmmh_create(bihip_ctx* ctx, const bihip_model* m, const bihip_smc2_config* cfg, int C,
                      const int* N_host, bihip_pmmh** out) {
  if (ctx == nullptr || m == nullptr || cfg == nullptr || N_host == nullptr || out == nullptr)
    return set_error(BIHIP_EINVAL, "NULL argument");
  const bihip_smc2_config& c = *cfg;
  if (int rc = pmmh_check(m, c, C)) return rc;
  bihip_pmmh* s = pmmh_new(m, c, C);
  auto* inner = new BatchInner();
  s->inner = inner;
  bihip_batch_config bc{};
  bc.B = c.B;
  bc.T = c.T;
  bc.resampler = c.resampler;
  bc.essRel = c.essRel_inner;
  bc.t0 = c.t0;
  bc.dt = c.dt;
  bc.seed = c.seed;
  bc.use_graph = c.use_graph;
  bc.block = c.block;
  std::vector<double> th0((size_t)c.B * c.ntheta, 1.0);
  int rc = bihip_batch_create(ctx, m, &bc, N_host, th0.data(), c.ntheta, &inner->b);
  if (rc != BIHIP_OK) {
    bihip_pmmh_destroy(s);
    return rc;
  }
  *out = s;
  return BIHIP_OK;
}

int bihip_pmmh_create_bank(bihip_ctx* ctx, const bihip_model* m, const bihip_smc2_config* cfg,
                           int C, int P, const double* mean0_host, bihip_pmmh** out) {
  if (ctx == nullptr || m == nullptr || cfg == nullptr || out == nullptr)
    return set_error(BIHIP_EINVAL, "NULL argument");
  const bihip_smc2_config& c = *cfg;
  if (int rc = pmmh_check(m, c, C)) return rc;
  if (P <= 0) return set_error(BIHIP_EINVAL, "need P > 0");
  if (m->model == BIHIP_MODEL_L96 && mean0_host == nullptr)
    return set_error(BIHIP_EINVAL, "Lorenz-96 needs mean0");
  bihip_pmmh* s = pmmh_new(m, c, C);
  auto* inner = new BankInner();
  s->inner = inner;
  s->bank = true;
  inner->P = P;
  if (mean0_host != nullptr) inner->mean0.assign(mean0_host, mean0_host + m->d);
  bihip_pf_config pc{};
  pc.P = P;
  pc.T = c.T;
  pc.resampler = c.resampler;
  pc.essRel = c.essRel_inner;
  pc.t0 = c.t0;
  pc.dt = c.dt;
  pc.seed = c.seed;
  pc.batch = 0;
  pc.record = 0;
  pc.use_graph = c.use_graph;
  std::vector<double> th0((size_t)c.B * c.ntheta, 1.0);
  int rc = bihip_pfbank_create(ctx, m, &pc, c.B, th0.data(), c.ntheta, &inner->b);
  if (rc != BIHIP_OK) {
    bihip_pmmh_destroy(s);
    return rc;
  }
  *out = s;
  return BIHIP_OK;
}

int bihip_pmmh_set_surrogate(bihip_pmmh* s, bihip_mlp* mlp, const bihip_surrogate_norm* norm) {
  if (s == nullptr) return set_error(BIHIP_EINVAL, "NULL sampler");
  if (!s->bank) return set_error(BIHIP_ESTATE, "the surrogate flow runs in filter banks");
  return bihip_pfbank_set_surrogate(static_cast<BankInner*>(s->inner)->b, mlp, norm);
}

int bihip_pmmh_set_guided(bihip_pmmh* s, bihip_mlp* mlp, const bihip_guided_norm* norm) {
  if (s == nullptr) return set_error(BIHIP_EINVAL, "NULL sampler");
  if (!s->bank) return set_error(BIHIP_ESTATE, "the guided proposal runs in filter banks");
  return bihip_pfbank_set_guided(static_cast<BankInner*>(s->inner)->b, mlp, norm);
}

int bihip_pmmh_destroy(bihip_pmmh* s) {
  if (s == nullptr) return BIHIP_OK;
  delete s->inner;
  delete s;
  return BIHIP_OK;
}

int bihip_pmmh_set_observations(bihip_pmmh* s, const double* y_host) {
  if (s == nullptr || y_host == nullptr) return set_error(BIHIP_EINVAL, "NULL argument");
  const int rc = s->inner->set_observations(y_host);
  s->have_obs = rc == BIHIP_OK;
  return rc;
}

int bihip_pmmh_run(bihip_pmmh* s) {
  if (s == nullptr) return set_error(BIHIP_EINVAL, "NULL sampler");
  if (!s->have_obs) return set_error(BIHIP_ESTATE, "no observations set");
  const bihip_smc2_config& c = s->cfg;
  const int B = c.B, nt = c.ntheta;
  std::vector<double> lth((size_t)B * nt), lth2((size_t)B * nt), ll, ll2;
  for (int j = 0; j < B; ++j) prior_lth(s->model, c, j, &lth[(size_t)j * nt]);
  int rc = pmmh_filter(s, lth, 0u, ll);
  if (rc) return rc;
  std::fill(s->accepted.begin(), s->accepted.end(), 1);
  for (int it = 0; it < s->C; ++it) {
    if (it > 0) {
      for (int j = 0; j < B; ++j) {
        double z[8];
        host_normals(c.seed, (uint32_t)j, (uint32_t)it, BIHIP_TAG_PROPOSAL, 0u, nt, z);
        for (int i = 0; i < nt; ++i)
          lth2[(size_t)j * nt + i] = lth[(size_t)j * nt + i] + c.prop_sd[i] * z[i];
      }
      if ((rc = pmmh_filter(s, lth2, (uint32_t)(B * it), ll2))) return rc;
      for (int j = 0; j < B; ++j) {
        bool accept;
        if (!std::isfinite(ll2[j])) {
          accept = false;
        } else if (!std::isfinite(ll[j])) {
          accept = true;
        } else {
          const double logratio = ll2[j] - ll[j] + log_prior(s->model, c, &lth2[(size_t)j * nt]) -
                                  log_prior(s->model, c, &lth[(size_t)j * nt]);
          const double u =
              host_uniform(c.seed, (uint32_t)j, (uint32_t)it, BIHIP_TAG_ACCEPT, 0u);
          accept = std::log(u) < logratio;
        }
        if (accept) {
          for (int i = 0; i < nt; ++i) lth[(size_t)j * nt + i] = lth2[(size_t)j * nt + i];
          ll[j] = ll2[j];
          ++s->accepted[j];
        }
      }
    }
    for (int j = 0; j < B; ++j) {
      for (int i = 0; i < nt; ++i)
        s->theta[((size_t)it * B + j) * nt + i] = std::exp(lth[(size_t)j * nt + i]);
      s->ll[(size_t)it * B + j] = ll[j];
    }
  }
  s->done = true;
  return BIHIP_OK;
}

int bihip_pmmh_outputs(bihip_pmmh* s, double* theta_host, double* ll_host, int* accepted_host) {
  if (s == nullptr) return set_error(BIHIP_EINVAL, "NULL sampler");
  if (!s->done) return set_error(BIHIP_ESTATE, "bihip_pmmh_run has not completed");
  if (theta_host) std::copy(s->theta.begin(), s->theta.end(), theta_host);
  if (ll_host) std::copy(s->ll.begin(), s->ll.end(), ll_host);
  if (accepted_host) std::copy(s->accepted.begin(), s->accepted.end(), accepted_host);
  return BIHIP_OK;
}

}  // extern "C"
