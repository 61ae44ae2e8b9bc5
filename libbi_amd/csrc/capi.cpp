// C ABI of the particle-filter library (declared in include/bihip.h).
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "internal.hpp"

namespace bihip {

static thread_local std::string g_error;

int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_error = buf;
  return code;
}


static int ws_alloc(ScanWorkspace* ws, int tiles) {
  ws->max_tiles = tiles;
  int rc;
  if ((rc = dalloc(&ws->state, 1))) return rc;
  if ((rc = dalloc(&ws->words, (size_t)SCAN_WORDS * tiles))) return rc;
  BIHIP_HIP(hipMemset(ws->state, 0, sizeof(ScanState)));
  BIHIP_HIP(hipMemset(ws->words, 0, sizeof(uint64_t) * SCAN_WORDS * tiles));
  return BIHIP_OK;
}

static int rw_alloc(ReduceWork* w, int nb) {
  w->max_blocks = reduce_blocks(nb);
  int rc;
  // 6 epoch-tagged words per block (k_reduce), then 3 plain doubles per block
  // (the two-launch combination)
  if ((rc = dalloc(&w->words, (size_t)9 * w->max_blocks))) return rc;
  if ((rc = dalloc(&w->state, 1))) return rc;
  BIHIP_HIP(hipMemset(w->words, 0, sizeof(uint64_t) * 9 * w->max_blocks));
  BIHIP_HIP(hipMemset(w->state, 0, sizeof(ScanState)));
  return BIHIP_OK;
}

static void rw_free(ReduceWork* w) {
  dfree(w->words);
  dfree(w->state);
  *w = ReduceWork{};
}

static void ws_free(ScanWorkspace* ws) {
  dfree(ws->state);
  dfree(ws->words);
  *ws = ScanWorkspace{};
}

int check_model(const bihip_model* m) {
  if (m == nullptr) return set_error(BIHIP_EINVAL, "model is NULL");
  if (!model_supported(m))
    return set_error(BIHIP_EINVAL,
                     "unsupported model (model %d, d %d, integrator %d, precision %d)",
                     m->model, m->d, m->integrator, m->precision);
  if (m->integrator == BIHIP_DOPRI5 && m->nsteps <= 0)
    return set_error(BIHIP_EINVAL, "nsteps must be positive");
  if (!(m->h0 > 0.0)) return set_error(BIHIP_EINVAL, "h0 must be positive");
  if (m->integrator == BIHIP_DOPRI5 && m->nsteps > RK4_MAX_STEPS)
    return set_error(BIHIP_EINVAL, "nsteps above %d", RK4_MAX_STEPS);
  return BIHIP_OK;
}

// An interval must take a bounded number of fixed RK4 steps and, for a
// generated model, a bounded number of its discrete-time steps.
int check_interval(const bihip_model* m, double t1, double t2) {
  if (!(t1 <= t2)) return set_error(BIHIP_EINVAL, "need t1 <= t2");
  if (m->integrator == BIHIP_RK4 && (t2 - t1) / m->h0 > 0.5 * RK4_MAX_STEPS)
    return set_error(BIHIP_EINVAL, "interval needs more than %d RK4 steps", RK4_MAX_STEPS / 2);
  const GenInfo* g = generated_info(m->model);
  if (g != nullptr && (t2 - t1) / g->delta > GEN_MAX_DELTAS)
    return set_error(BIHIP_EINVAL, "interval spans more than %d multiples of the model's delta",
                     GEN_MAX_DELTAS);
  return BIHIP_OK;
}

int nobs_of(const bihip_model* m) {
  if (const GenInfo* g = generated_info(m->model)) return g->nobs;
  switch (m->model) {
    case BIHIP_MODEL_L96: return m->d / 2;
    default: return 1;
  }
}

}  // namespace bihip

using namespace bihip;

// ================================================================ context
struct bihip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // scratch for the standalone primitives (grown on demand)
  double* partials = nullptr;
  int partials_cap = 0;
  double* out = nullptr;  // 5 doubles
  ReduceWork rw{};
  ScanWorkspace ws{};
  uint64_t* EF = nullptr;
  int EF_cap = 0;
  uint8_t* mc = nullptr;  // the chunked multinomial's scratch
  size_t mc_cap = 0;
};

static int ctx_enter(bihip_ctx* ctx) {
  if (ctx == nullptr) return set_error(BIHIP_EINVAL, "context is NULL");
  BIHIP_HIP(hipSetDevice(ctx->device));
  return BIHIP_OK;
}

namespace bihip {
int ctx_enter_device(bihip_ctx* ctx) { return ctx_enter(ctx); }
hipStream_t ctx_stream_of(bihip_ctx* ctx) { return ctx->stream; }
}  // namespace bihip

static int ctx_partials(bihip_ctx* ctx, int nb) {
  if (nb > ctx->partials_cap) {
    BIHIP_HIP(hipStreamSynchronize(ctx->stream));
    dfree(ctx->partials);
    rw_free(&ctx->rw);
    ctx->partials_cap = 0;
    int rc = dalloc(&ctx->partials, (size_t)3 * nb);
    if (rc) return rc;
    if ((rc = rw_alloc(&ctx->rw, nb))) return rc;
    ctx->partials_cap = nb;
  }
  return BIHIP_OK;
}

static int ctx_scan(bihip_ctx* ctx, int P) {
  const int nt = scan_tiles(P);
  if (nt > ctx->ws.max_tiles) {
    BIHIP_HIP(hipStreamSynchronize(ctx->stream));
    ws_free(&ctx->ws);
    int rc = ws_alloc(&ctx->ws, nt);
    if (rc) return rc;
  }
  return BIHIP_OK;
}

static int ctx_EF(bihip_ctx* ctx, int P) {
  if (P > ctx->EF_cap) {
    BIHIP_HIP(hipStreamSynchronize(ctx->stream));
    dfree(ctx->EF);
    ctx->EF_cap = 0;
    int rc = dalloc(&ctx->EF, P);
    if (rc) return rc;
    ctx->EF_cap = P;
  }
  return BIHIP_OK;
}

static int ctx_mc(bihip_ctx* ctx, int P, int n, McWork* w) {
  const size_t bytes = McWork::bytes(P, n, true);
  if (bytes > ctx->mc_cap) {
    BIHIP_HIP(hipStreamSynchronize(ctx->stream));
    dfree(ctx->mc);
    ctx->mc_cap = 0;
    int rc = dalloc(&ctx->mc, bytes);
    if (rc) return rc;
    ctx->mc_cap = bytes;
  }
  w->carve(ctx->mc, P, n, true);
  return BIHIP_OK;
}

extern "C" {

const char* bihip_last_error(void) { return g_error.c_str(); }

int bihip_ctx_create(int device, bihip_ctx** out) {
  if (out == nullptr) return set_error(BIHIP_EINVAL, "out is NULL");
  int n = 0;
  BIHIP_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return set_error(BIHIP_EINVAL, "no device %d", device);
  BIHIP_HIP(hipSetDevice(device));
  // the device's address of the static scratch of the allocation-free entry
  // point, looked up once per device with the device current
  int rc = noalloc_scratch_init(device);
  if (rc) return rc;
  bihip_ctx* c = new bihip_ctx();
  c->device = device;
  rc = dalloc(&c->out, 5);
  if (rc) {
    delete c;
    return rc;
  }
  *out = c;
  return BIHIP_OK;
}

int bihip_ctx_destroy(bihip_ctx* ctx) {
  if (ctx == nullptr) return BIHIP_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  dfree(ctx->partials);
  dfree(ctx->out);
  dfree(ctx->EF);
  dfree(ctx->mc);
  rw_free(&ctx->rw);
  ws_free(&ctx->ws);
  delete ctx;
  return BIHIP_OK;
}

int bihip_ctx_set_stream(bihip_ctx* ctx, void* stream) {
  if (ctx == nullptr) return set_error(BIHIP_EINVAL, "context is NULL");
  ctx->stream = static_cast<hipStream_t>(stream);
  return BIHIP_OK;
}

int bihip_ctx_synchronize(bihip_ctx* ctx) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  BIHIP_HIP(hipStreamSynchronize(ctx->stream));
  return BIHIP_OK;
}

int bihip_model_nobs(const bihip_model* m) { return m ? nobs_of(m) : 0; }

int bihip_model_elem_size(const bihip_model* m) {
  return m && m->precision == BIHIP_F64 ? 8 : 4;
}

// ================================================================ memory
int bihip_malloc(bihip_ctx* ctx, size_t bytes, void** ptr) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if (ptr == nullptr) return set_error(BIHIP_EINVAL, "ptr is NULL");
  BIHIP_HIP(hipMalloc(ptr, bytes ? bytes : 1));
  return BIHIP_OK;
}

int bihip_free(bihip_ctx* ctx, void* ptr) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  BIHIP_HIP(hipStreamSynchronize(ctx->stream));
  if (ptr) BIHIP_HIP(hipFree(ptr));
  return BIHIP_OK;
}

int bihip_memcpy_to_device(bihip_ctx* ctx, void* dst, const void* src, size_t bytes) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  BIHIP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  BIHIP_HIP(hipStreamSynchronize(ctx->stream));
  return BIHIP_OK;
}

int bihip_memcpy_to_host(bihip_ctx* ctx, void* dst, const void* src, size_t bytes) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  BIHIP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  BIHIP_HIP(hipStreamSynchronize(ctx->stream));
  return BIHIP_OK;
}

int bihip_memcpy_device(bihip_ctx* ctx, void* dst, const void* src, size_t bytes) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  BIHIP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return BIHIP_OK;
}

int bihip_fill(bihip_ctx* ctx, double* x, int n, double v) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if (n < 0) return set_error(BIHIP_EINVAL, "n < 0");
  return launch_fill_f64(x, n, v, ctx->stream);
}

int bihip_iota(bihip_ctx* ctx, int* x, int n) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if (n < 0) return set_error(BIHIP_EINVAL, "n < 0");
  return launch_iota(x, n, ctx->stream);
}

// ================================================================ random numbers
// Philox4x32-10 on the host (Salmon et al., SC'11), as the device's.

static void host_philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = (uint32_t)p1;
    c[2] = n2;
    c[3] = (uint32_t)p0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

double bihip_uniform_host(uint64_t seed, uint32_t index, uint32_t step, uint32_t tag,
                          uint32_t batch) {
  uint32_t c[4];
  bihip::philox_host_block(seed, index, step, tag, batch, 0u, c);
  const uint64_t m = ((((uint64_t)c[0]) << 32) | c[1]) >> 11;
  return (double)m * 0x1.0p-53;
}

int bihip_philox(bihip_ctx* ctx, const uint32_t* ctr, uint32_t k0, uint32_t k1, int n,
                 uint32_t* out) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if (n < 0) return set_error(BIHIP_EINVAL, "n < 0");
  return launch_philox(ctr, k0, k1, n, out, ctx->stream);
}

int bihip_uniforms(bihip_ctx* ctx, uint64_t seed, uint32_t step, uint32_t tag, uint32_t batch,
                   int n, double* u) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if (n < 0) return set_error(BIHIP_EINVAL, "n < 0");
  return launch_uniforms(seed, step, tag, batch, n, u, ctx->stream);
}

int bihip_normals(bihip_ctx* ctx, uint64_t seed, uint32_t index, uint32_t step, uint32_t tag,
                  uint32_t batch, int n, int precision, void* z) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if (n < 0) return set_error(BIHIP_EINVAL, "n < 0");
  return launch_normals(seed, index, step, tag, batch, n, precision, z, ctx->stream);
}

// ================================================================ model
int bihip_init_samples(bihip_ctx* ctx, const bihip_model* m, uint64_t seed, uint32_t batch,
                       int P, void* X, int ld, const double* mean0_host) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if ((rc = check_model(m))) return rc;
  if (P < 0 || ld < P) return set_error(BIHIP_EINVAL, "need 0 <= P <= ld");
  double* mean0 = nullptr;
  if (m->model == BIHIP_MODEL_L96) {
    if (mean0_host == nullptr) return set_error(BIHIP_EINVAL, "L96 needs mean0");
    if ((rc = dalloc(&mean0, m->d))) return rc;
    BIHIP_HIP(hipMemcpyAsync(mean0, mean0_host, sizeof(double) * m->d, hipMemcpyHostToDevice,
                             ctx->stream));
  }
  rc = launch_init_samples(m, seed, batch, P, X, ld, mean0, ctx->stream);
  if (mean0) {
    BIHIP_HIP(hipStreamSynchronize(ctx->stream));
    dfree(mean0);
  }
  return rc;
}

int bihip_transition(bihip_ctx* ctx, const bihip_model* m, uint64_t seed, uint32_t batch,
                     uint32_t step, double t1, double t2, int P, const void* Xin, void* Xout,
                     int ld, const int* as, int* nsteps) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if ((rc = check_model(m))) return rc;
  if (P < 0 || ld < P) return set_error(BIHIP_EINVAL, "need 0 <= P <= ld");
  if (as != nullptr && Xin == Xout)
    return set_error(BIHIP_EINVAL, "a gathering transition cannot run in place");
  if ((rc = check_interval(m, t1, t2))) return rc;
  TransLaunch L{};
  L.model = m;
  L.Xin = Xin;
  L.Xout = Xout;
  L.ld = ld;
  L.P = P;
  L.as = as;
  L.t1 = t1;
  L.t2 = t2;
  L.seed = seed;
  L.step = step;
  L.batch = batch;
  L.nsteps_out = nsteps;
  return launch_transition(L, ctx->stream);
}

int bihip_flow(bihip_ctx* ctx, const bihip_model* m, double t1, double t2, int P, void* X,
               int ld, const double* aux, int* nsteps) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if ((rc = check_model(m))) return rc;
  if (P < 0 || ld < P) return set_error(BIHIP_EINVAL, "need 0 <= P <= ld");
  if ((rc = check_interval(m, t1, t2))) return rc;
  if (generated_info(m->model) != nullptr)
    return set_error(BIHIP_EINVAL, "the deterministic flow is defined for the built-in models");
  if (!(t2 - t1 > 0.0)) return BIHIP_OK;  // DOPRI5Integrator.hpp:51: nothing to do
  TransLaunch L{};
  L.model = m;
  L.Xin = X;
  L.Xout = X;
  L.ld = ld;
  L.P = P;
  L.t1 = t1;
  L.t2 = t2;
  L.nsteps_out = nsteps;
  L.flow_only = 1;
  L.aux = aux;
  return launch_transition(L, ctx->stream);
}

int bihip_obs_logdensity(bihip_ctx* ctx, const bihip_model* m, const double* y, int P,
                         const void* X, int ld, double* lws) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if ((rc = check_model(m))) return rc;
  if (P < 0 || ld < P) return set_error(BIHIP_EINVAL, "need 0 <= P <= ld");
  return launch_obs_logdensity(m, y, P, X, ld, lws, ctx->stream);
}

int bihip_obs_sample(bihip_ctx* ctx, const bihip_model* m, uint64_t seed, uint32_t batch,
                     uint32_t step, int P, const void* X, int ld, double* y) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if ((rc = check_model(m))) return rc;
  if (P < 0 || ld < P) return set_error(BIHIP_EINVAL, "need 0 <= P <= ld");
  return launch_obs_sample(m, seed, batch, step, P, X, ld, y, ctx->stream);
}

// ================================================================ reductions
static int reduce_to_out(bihip_ctx* ctx, const double* lws, int P, double host[5]) {
  if (P <= 0) return set_error(BIHIP_EINVAL, "reduction of an empty vector");
  const int nb = ceil_div(P, RED_TILE);
  int rc = ctx_partials(ctx, nb);
  if (rc) return rc;
  if ((rc = launch_lws_partials(lws, P, ctx->partials, ctx->stream))) return rc;
  if ((rc = launch_finalize(ctx->partials, nb, P, ctx->out, ctx->rw, ctx->stream))) return rc;
  if (host) {
    BIHIP_HIP(hipMemcpyAsync(host, ctx->out, 5 * sizeof(double), hipMemcpyDeviceToHost,
                             ctx->stream));
    BIHIP_HIP(hipStreamSynchronize(ctx->stream));
  }
  return BIHIP_OK;
}

int bihip_ess_reduce(bihip_ctx* ctx, const double* lws, int P, double* ess, double* lW) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  double h[5];
  if ((rc = reduce_to_out(ctx, lws, P, h))) return rc;
  if (ess) *ess = h[3];
  if (lW) *lW = h[4];
  return BIHIP_OK;
}

int bihip_logsumexp_reduce(bihip_ctx* ctx, const double* lws, int P, double* out) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  double h[5];
  if ((rc = reduce_to_out(ctx, lws, P, h))) return rc;
  if (out) *out = h[0] + log(h[1]);
  return BIHIP_OK;
}

int bihip_max_reduce(bihip_ctx* ctx, const double* x, int n, double* out) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  double h[5];
  if ((rc = reduce_to_out(ctx, x, n, h))) return rc;
  if (out) *out = h[0];
  return BIHIP_OK;
}

int bihip_sumexpu_inclusive_scan(bihip_ctx* ctx, const double* lws, int P, double* Ws,
                                 double* mx_host) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if (P <= 0) return set_error(BIHIP_EINVAL, "scan of an empty vector");
  if ((rc = reduce_to_out(ctx, lws, P, nullptr))) return rc;
  if ((rc = ctx_scan(ctx, P))) return rc;
  // a grid that is resident at once runs the single-pass look-back scan; a
  // larger one the tiled reduce-then-scan, whose tiles never wait on each
  // other (the look-back would number its tiles by a ticket, one device-scope
  // atomic per tile on one address, which serialises: 474 us at 2^24 against
  // 144 us for the tiled scan)
  if (scan_tiles(P) <= scan_resident_tiles())
    rc = launch_sumexp_scan(lws, P, ctx->out, nullptr, Ws, ctx->ws, ctx->stream);
  else
    rc = launch_sumexp_scan_tiles(lws, P, ctx->out, Ws, ctx->ws, ctx->stream);
  if (rc) return rc;
  if (mx_host) {
    BIHIP_HIP(hipMemcpyAsync(mx_host, ctx->out, sizeof(double), hipMemcpyDeviceToHost,
                             ctx->stream));
    BIHIP_HIP(hipStreamSynchronize(ctx->stream));
  }
  return BIHIP_OK;
}

// ================================================================ resamplers
// The host branch `if (pre.W > 0)` of the reference (SystematicResampler.hpp:118).
static int check_total(bihip_ctx* ctx, const double* Ws, int P) {
  if (P <= 0) return set_error(BIHIP_EINVAL, "no particles");
  double W = 0.0;
  BIHIP_HIP(hipMemcpyAsync(&W, Ws + (P - 1), sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  BIHIP_HIP(hipStreamSynchronize(ctx->stream));
  if (!(W > 0.0)) return set_error(BIHIP_DEGENERATE, "particle filter degenerated (W <= 0)");
  return BIHIP_OK;
}

int bihip_systematic_cumulative_offspring(bihip_ctx* ctx, const double* Ws, int P, int n,
                                          double a, int* Os) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if (n <= 0) return set_error(BIHIP_EINVAL, "n must be positive");
  if (!(a >= 0.0 && a <= 1.0)) return set_error(BIHIP_EINVAL, "a must lie in [0,1]");
  if ((rc = check_total(ctx, Ws, P))) return rc;
  OffspringLaunch L{};
  L.mode = BIHIP_SYSTEMATIC;
  L.Ws = Ws;
  L.P = P;
  L.n = n;
  L.a = a;
  L.Os = Os;
  return launch_cumulative_offspring(L, ctx->stream);
}

int bihip_stratified_cumulative_offspring(bihip_ctx* ctx, const double* Ws, int P, int n,
                                          const double* alphas, int* Os) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if (n <= 0) return set_error(BIHIP_EINVAL, "n must be positive");
  if (alphas == nullptr) return set_error(BIHIP_EINVAL, "alphas is NULL");
  if ((rc = check_total(ctx, Ws, P))) return rc;
  OffspringLaunch L{};
  L.mode = BIHIP_STRATIFIED;
  L.Ws = Ws;
  L.P = P;
  L.n = n;
  L.alphas = alphas;
  L.Os = Os;
  return launch_cumulative_offspring(L, ctx->stream);
}

int bihip_multinomial_ancestors(bihip_ctx* ctx, const double* Ws, int P, int n, const double* U,
                                int* as) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if (n < 0) return set_error(BIHIP_EINVAL, "n < 0");
  if (U == nullptr) return set_error(BIHIP_EINVAL, "U is NULL");
  if ((rc = check_total(ctx, Ws, P))) return rc;
  MultinomialLaunch L{};
  L.Ws = Ws;
  L.P = P;
  L.n = n;
  L.U = U;
  L.as = as;
  if (n > 0 && multinomial_chunks_fit(P)) {
    McWork w;
    if ((rc = ctx_mc(ctx, P, n, &w))) return rc;
    return launch_multinomial_chunks(L, w, ctx->stream);
  }
  return launch_multinomial(L, ctx->stream);
}

int bihip_ancestors_to_offspring(bihip_ctx* ctx, const int* as, int n, int P, int* os) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if (n < 0 || P <= 0) return set_error(BIHIP_EINVAL, "bad sizes");
  if ((rc = launch_zero_ints(os, P, nullptr, ctx->stream))) return rc;
  return launch_ancestors_to_offspring(as, n, P, os, ctx->stream);
}

int bihip_offspring_to_cumulative(bihip_ctx* ctx, const int* os, int P, int* Os) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if (P <= 0) return set_error(BIHIP_EINVAL, "P must be positive");
  if ((rc = ctx_scan(ctx, P))) return rc;
  return launch_count_scan(os, P, Os, ctx->ws, ctx->stream);
}

int bihip_cumulative_offspring_to_ancestors(bihip_ctx* ctx, const int* Os, int P, int* as) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if (P <= 0) return set_error(BIHIP_EINVAL, "P must be positive");
  return launch_cumulative_to_ancestors(Os, P, as, ctx->stream);
}

int bihip_cumulative_offspring_to_ancestors_permute(bihip_ctx* ctx, const int* Os, int P,
                                                    int* as) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if (P <= 0) return set_error(BIHIP_EINVAL, "P must be positive");
  if ((rc = ctx_scan(ctx, P))) return rc;
  if ((rc = ctx_EF(ctx, P))) return rc;
  return launch_ancestors_permute(Os, true, P, as, ctx->EF, ctx->ws, nullptr, ctx->stream);
}

int bihip_offspring_to_ancestors_permute(bihip_ctx* ctx, const int* os, int P, int* as) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if (P <= 0) return set_error(BIHIP_EINVAL, "P must be positive");
  if ((rc = ctx_scan(ctx, P))) return rc;
  if ((rc = ctx_EF(ctx, P))) return rc;
  return launch_ancestors_permute(os, false, P, as, ctx->EF, ctx->ws, nullptr, ctx->stream);
}

int bihip_gather_rows(bihip_ctx* ctx, const int* as, int P, int ncols, void* X, int ld,
                      int elem) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if (P < 0 || ncols < 0 || ld < P) return set_error(BIHIP_EINVAL, "bad sizes");
  return launch_gather_rows(as, P, ncols, X, ld, elem, ctx->stream);
}

}  // extern "C"

// ================================================================ filter
struct bihip_pf {
  bihip_ctx* ctx = nullptr;
  bihip_model model{};
  bihip_pf_config cfg{};
  int ld = 0, elem = 4, nobs = 1, nparts = 0;
  bool one_delta = false;  // gen_one_delta over the run's steps
  void* X[2] = {nullptr, nullptr};
  int cur = 0;
  double* lws = nullptr;
  double* Ws = nullptr;
  int* Os = nullptr;
  int* os = nullptr;
  int* as = nullptr;
  uint64_t* EF = nullptr;
  double* partials = nullptr;
  double* out = nullptr;
  ReduceWork rw{};
  PFDevice* dev = nullptr;
  double* y = nullptr;
  double* hist_li = nullptr;
  double* hist_ess = nullptr;
  int* hist_res = nullptr;
  int* hist_anc = nullptr;
  void* hist_X = nullptr;  // record = 2: the state at every time, (T + 1) x d x ld
  double* hist_lw = nullptr;  // record = 2: the log-weights at every time, (T + 1) x P
  void* hist_noise = nullptr;  // record = 2: the noise variables, (T + 1) x noise_width x ld
  int noise_width = 0;
  ScanWorkspace ws{};
  int k = 0;
  bool ready = false, have_obs = false;
  // timing: 0 off, BIHIP_TIMING_ALL, BIHIP_TIMING_TRANSITION
  int timing = 0;
  std::vector<hipEvent_t> ev[4];
  size_t ev_used[4] = {0, 0, 0, 0};
  // graphs (one per buffer parity)
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  // learned surrogate flow in place of the integrator (bihip_pf_set_surrogate)
  bool surr = false;
  SurrogateLaunch sl{};
  int nparts_exact = 0;
  // learned guided proposal (bihip_pf_set_guided)
  bool guided = false;
  GuidedLaunch gl{};
  // schedule (bihip_pf_set_schedule): T + 1 times and T observed flags
  double* times = nullptr;
  int* obs_mask = nullptr;
  std::vector<double> times_host;
  // inputs (bihip_pf_set_inputs): T x 16 model parameters
  double* inputs = nullptr;
  // imposed resampling (bihip_pf_impose_resampling): T flags, T x P ancestors
  int* imp_r = nullptr;
  int* imp_as = nullptr;
  // record = 3: the pruned ancestry tree (AncestryCache)
  AncestryCache cache{};
  bool use_cache = false;
  // multinomial resampling: the chunked search's scratch
  uint8_t* mc_buf = nullptr;
  McWork mc{};
};

// Timing events only: created without the system-scope fence, so recording one
// does not write back and invalidate the caches between the kernels it brackets
// (the stream is synchronized before the times are read).
static int pf_event_pair(bihip_pf* pf, int kind, hipEvent_t* a, hipEvent_t* b) {
  auto& v = pf->ev[kind];
  size_t& u = pf->ev_used[kind];
  while (v.size() < u + 2) {
    hipEvent_t e;
    BIHIP_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    v.push_back(e);
  }
  *a = v[u];
  *b = v[u + 1];
  u += 2;
  return BIHIP_OK;
}

// One observation interval: resample (device-gated) -> record -> predict+correct -> reduce.
// timing: 0, BIHIP_TIMING_ALL (events around every phase) or
// BIHIP_TIMING_TRANSITION (around the transition and proposal only).
// Room for one more generation in the ancestry tree: the host keeps an upper
// bound of its live nodes (P more per step) and reads the true count, and
// grows the store, only when that bound leaves no room.
static int pf_cache_room(bihip_pf* pf, hipStream_t s) {
  AncestryCache& c = pf->cache;
  const long P = pf->cfg.P;
  if (c.live_ub + P <= c.cap) return BIHIP_OK;
  int live = 0;
  BIHIP_HIP(hipMemcpyAsync(&live, c.live, sizeof(int), hipMemcpyDeviceToHost, s));
  BIHIP_HIP(hipStreamSynchronize(s));
  c.live_ub = live;
  if (live + P <= c.cap) return BIHIP_OK;
  const long want = std::max(2 * c.cap, (long)live + 2 * P);
  if (want >= (1L << 31))
    return set_error(BIHIP_EINVAL, "the ancestry tree would exceed 2^31 nodes");
  return ancestry_enlarge(&c, want, s);
}

static int pf_enqueue_step(bihip_pf* pf, hipStream_t s, int timing) {
  const bool timed = timing == BIHIP_TIMING_ALL;       // every phase
  const bool timed_trans = timing != 0;                // transition and proposal
  const int P = pf->cfg.P;
  int rc;
  if (pf->use_cache && (rc = pf_cache_room(pf, s))) return rc;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  // ---- resample (Resampler::resample, gated by dev->resample)
  if (timed) {
    if ((rc = pf_event_pair(pf, 1, &e0, &e1))) return rc;
    BIHIP_HIP(hipEventRecord(e0, s));
  }
  // an imposed history replaces the filter's own resampling: no chain, the
  // step gathers through the imposed ancestors of step k when flagged
  const int* step_as = pf->imp_r != nullptr ? pf->imp_as + (size_t)pf->k * P : pf->as;
  ResampleTilesLaunch R{};
  R.mode = pf->cfg.resampler;
  R.lws = pf->lws;
  R.mx = &pf->dev->mx;
  R.P = P;
  R.n = P;
  R.seed = pf->cfg.seed;
  R.batch = pf->cfg.batch;
  R.gate = pf->dev;
  R.status = pf->dev;
  R.Ws = pf->Ws;
  R.Os = pf->Os;
  R.os = pf->os;
  R.as = pf->as;
  R.EF = pf->EF;
  R.ws = pf->ws;
  R.mc = pf->mc_buf != nullptr ? &pf->mc : nullptr;
  // the previous correct step stored the tile weight sums (fixed per filter)
  const bool stored_sums =
      correct_stores_sums(pf->nparts) && ceil_div(P, SCAN_TILE) <= pf->ws.max_tiles;
  R.sums_ready = stored_sums ? 1 : 0;
  if (pf->imp_r == nullptr) {
    if ((rc = launch_resample_tiles(R, s))) return rc;
    if (pf->hist_anc) {
      if ((rc = launch_record_ancestors(pf->as, P, pf->dev, pf->hist_anc, s))) return rc;
    }
  }
  if (timed) BIHIP_HIP(hipEventRecord(e1, s));

  // ---- predict + correct (fused, ancestor gather on read)
  if (timed_trans) {
    if ((rc = pf_event_pair(pf, 0, &e0, &e1))) return rc;
    BIHIP_HIP(hipEventRecord(e0, s));
  }
  TransLaunch T{};
  T.model = &pf->model;
  T.Xin = pf->X[pf->cur];
  T.Xout = pf->X[1 - pf->cur];
  T.ld = pf->ld;
  T.P = P;
  T.as = step_as;
  T.imp_r = pf->imp_r;
  T.lws = pf->lws;
  T.seed = pf->cfg.seed;
  T.batch = pf->cfg.batch;
  T.partials = pf->partials;
  T.dev = pf->dev;
  T.t0 = pf->cfg.t0;
  T.dt = pf->cfg.dt;
  T.y_base = pf->y;
  T.nobs = pf->nobs;
  T.times = pf->times;
  T.mask = pf->obs_mask;
  T.inputs = pf->inputs;
  T.one_delta = pf->one_delta ? 1 : 0;
  T.noise_hist = pf->hist_noise;
  T.noise_width = pf->noise_width;
  if (pf->guided) {
    // exact flow of the gathered particles (no noise, no correction) into Xout,
    // then the proposal and the corrected weights in place
    TransLaunch F = T;
    F.flow_only = 1;
    F.y_base = nullptr;
    F.lws = nullptr;
    F.partials = nullptr;
    if ((rc = launch_transition(F, s))) return rc;
    GuidedLaunch G = pf->gl;
    G.X = T.Xout;
    G.ld = pf->ld;
    G.P = P;
    G.seed = pf->cfg.seed;
    G.batch = pf->cfg.batch;
    G.dev = pf->dev;
    G.y_base = pf->y;
    G.nobs = pf->nobs;
    G.lws = pf->lws;
    G.partials = pf->partials;
    G.imp_r = pf->imp_r;
    hipEvent_t g0 = nullptr, g1 = nullptr;
    if (timed_trans) {
      if ((rc = pf_event_pair(pf, 3, &g0, &g1))) return rc;
      BIHIP_HIP(hipEventRecord(g0, s));
    }
    if ((rc = launch_guided(G, s))) return rc;
    if (timed_trans) BIHIP_HIP(hipEventRecord(g1, s));
  } else if (pf->surr) {
    SurrogateLaunch S = pf->sl;
    S.Xin = T.Xin;
    S.Xout = T.Xout;
    S.ld = pf->ld;
    S.P = P;
    S.step = true;
    S.as = step_as;
    S.imp_r = pf->imp_r;
    S.lws = pf->lws;
    S.partials = pf->partials;
    S.dev = pf->dev;
    S.seed = pf->cfg.seed;
    S.batch = pf->cfg.batch;
    S.t0 = pf->cfg.t0;
    S.dt = pf->cfg.dt;
    S.y_base = pf->y;
    S.nobs = pf->nobs;
    if ((rc = launch_surrogate(S, s))) return rc;
  } else if ((rc = launch_transition(T, s))) {
    return rc;
  }
  if (timed_trans) BIHIP_HIP(hipEventRecord(e1, s));
  // the new generation into the ancestry tree (AncestryCache::writeState,
  // share/src/bi/cache/AncestryCache.hpp:400-436), before the reduction moves
  // the resampling decision on
  if (pf->use_cache) {
    if ((rc = ancestry_write(&pf->cache, T.Xout, pf->ld, step_as, pf->dev, pf->imp_r, s)))
      return rc;
    pf->cache.live_ub += P;
  }
  if (pf->hist_X != nullptr &&
      ((rc = launch_record_state(T.Xout, pf->ld, pf->model.d, P, pf->elem, pf->dev, pf->hist_X, 0,
                                 s)) ||
       (rc = launch_record_state(pf->lws, P, 1, P, 8, pf->dev, pf->hist_lw, 0, s))))
    return rc;

  // ---- BootstrapPF::correct bookkeeping
  if (timed) {
    if ((rc = pf_event_pair(pf, 2, &e0, &e1))) return rc;
    BIHIP_HIP(hipEventRecord(e0, s));
  }
  // with an imposed history the recorded flags are the imposed ones
  PFHistory H{pf->hist_li, pf->hist_ess, pf->imp_r != nullptr ? nullptr : pf->hist_res};
  if ((rc = launch_pf_correct_finalize(pf->partials, pf->nparts, P, pf->cfg.essRel, pf->dev, H,
                                       pf->rw, s, pf->obs_mask, pf->lws,
                                       stored_sums ? tile_sum_words(pf->ws) : nullptr)))
    return rc;
  if (timed) BIHIP_HIP(hipEventRecord(e1, s));
  return BIHIP_OK;
}

static int pf_capture(bihip_pf* pf, int parity) {
  hipStream_t cs;
  BIHIP_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipGraph_t g = nullptr;
  const int saved = pf->cur;
  pf->cur = parity;
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  int rc = BIHIP_OK;
  if (e == hipSuccess) {
    rc = pf_enqueue_step(pf, cs, 0);
    hipError_t e2 = hipStreamEndCapture(cs, &g);
    if (rc == BIHIP_OK && e2 != hipSuccess)
      rc = set_error(BIHIP_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e2));
  } else {
    rc = set_error(BIHIP_EHIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
  }
  pf->cur = saved;
  if (rc == BIHIP_OK) {
    hipError_t e3 = hipGraphInstantiate(&pf->graph[parity], g, nullptr, nullptr, 0);
    if (e3 != hipSuccess) rc = set_error(BIHIP_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e3));
  }
  if (g) (void)hipGraphDestroy(g);
  (void)hipStreamDestroy(cs);
  return rc;
}

extern "C" {

int bihip_pf_create(bihip_ctx* ctx, const bihip_model* m, const bihip_pf_config* cfg,
                    bihip_pf** out) {
  int rc = ctx_enter(ctx);
  if (rc) return rc;
  if ((rc = check_model(m))) return rc;
  if (cfg == nullptr || out == nullptr) return set_error(BIHIP_EINVAL, "NULL argument");
  if (cfg->P <= 0 || cfg->T <= 0) return set_error(BIHIP_EINVAL, "need P > 0 and T > 0");
  if (cfg->resampler < BIHIP_SYSTEMATIC || cfg->resampler > BIHIP_MULTINOMIAL)
    return set_error(BIHIP_EINVAL, "unknown resampler %d", cfg->resampler);
  if (!(cfg->essRel >= 0.0 && cfg->essRel <= 1.0))
    return set_error(BIHIP_EINVAL, "essRel must lie in [0,1]");
  if (!(cfg->dt > 0.0)) return set_error(BIHIP_EINVAL, "dt must be positive");
  if (cfg->record < 0 || cfg->record > 3) return set_error(BIHIP_EINVAL, "record must be 0..3");
  if (cfg->record == 3 && cfg->use_graph)
    return set_error(BIHIP_EINVAL, "the ancestry tree (record = 3) needs a filter without graph "
                                   "replay");
  if ((rc = check_interval(m, cfg->t0, cfg->t0 + cfg->dt))) return rc;
  bihip_pf* pf = new bihip_pf();
  pf->ctx = ctx;
  pf->model = *m;
  pf->cfg = *cfg;
  pf->one_delta = gen_one_delta(m, cfg->t0, cfg->dt, cfg->T);
  const int P = cfg->P;
  pf->ld = (P + 63) / 64 * 64;
  pf->elem = bihip_model_elem_size(m);
  pf->nobs = nobs_of(m);
  pf->nparts = transition_blocks(m, P);
  const size_t xbytes = (size_t)pf->ld * m->d * pf->elem;
#define PF_ALLOC(ptr, n)                       \
  if ((rc = dalloc(&(ptr), (n))) != BIHIP_OK) { \
    bihip_pf_destroy(pf);                      \
    return rc;                                 \
  }
  uint8_t* x0 = nullptr;
  uint8_t* x1 = nullptr;
  PF_ALLOC(x0, xbytes);
  PF_ALLOC(x1, xbytes);
  pf->X[0] = x0;
  pf->X[1] = x1;
  PF_ALLOC(pf->lws, P);
  PF_ALLOC(pf->Ws, P);
  PF_ALLOC(pf->Os, P);
  PF_ALLOC(pf->os, P);
  PF_ALLOC(pf->as, P);
  PF_ALLOC(pf->EF, P);
  PF_ALLOC(pf->partials, (size_t)3 * pf->nparts);
  PF_ALLOC(pf->out, 5);
  PF_ALLOC(pf->dev, 1);
  PF_ALLOC(pf->y, (size_t)cfg->T * pf->nobs);
  PF_ALLOC(pf->hist_li, cfg->T);
  PF_ALLOC(pf->hist_ess, cfg->T);
  PF_ALLOC(pf->hist_res, cfg->T);
  if (cfg->resampler == BIHIP_MULTINOMIAL && multinomial_chunks_fit(P)) {
    PF_ALLOC(pf->mc_buf, McWork::bytes(P, P, false));
    pf->mc.carve(pf->mc_buf, P, P, false);
  }
  if (cfg->record == 1 || cfg->record == 2) PF_ALLOC(pf->hist_anc, (size_t)cfg->T * P);
  if (cfg->record == 2) {
    uint8_t* hx = nullptr;
    PF_ALLOC(hx, (size_t)(cfg->T + 1) * m->d * pf->ld * pf->elem);
    pf->hist_X = hx;
    PF_ALLOC(pf->hist_lw, (size_t)(cfg->T + 1) * P);
    pf->noise_width = noise_width_of(m);
    if (pf->noise_width > 0) {
      uint8_t* hn = nullptr;
      PF_ALLOC(hn, (size_t)(cfg->T + 1) * pf->noise_width * pf->ld * pf->elem);
      pf->hist_noise = hn;
    }
  }
#undef PF_ALLOC
  if (cfg->record == 3) {
    // room for four generations to start with; grown when a step could overflow it
    if ((rc = ancestry_alloc(&pf->cache, P, m->d, pf->elem, (long)4 * P, ctx->stream))) {
      ancestry_free(&pf->cache);
      bihip_pf_destroy(pf);
      return rc;
    }
    pf->use_cache = true;
  }
  if ((rc = ws_alloc(&pf->ws, scan_tiles(P))) ||
      (rc = rw_alloc(&pf->rw, std::max(pf->nparts, ceil_div(P, RED_TILE))))) {
    bihip_pf_destroy(pf);
    return rc;
  }
  *out = pf;
  return BIHIP_OK;
}

int bihip_pf_destroy(bihip_pf* pf) {
  if (pf == nullptr) return BIHIP_OK;
  (void)hipSetDevice(pf->ctx->device);
  (void)hipStreamSynchronize(pf->ctx->stream);
  for (int i = 0; i < 2; ++i) {
    if (pf->graph[i]) (void)hipGraphExecDestroy(pf->graph[i]);
    dfree(pf->X[i]);
  }
  for (auto& v : pf->ev)
    for (auto e : v) (void)hipEventDestroy(e);
  dfree(pf->lws);
  dfree(pf->Ws);
  dfree(pf->Os);
  dfree(pf->os);
  dfree(pf->as);
  dfree(pf->EF);
  dfree(pf->partials);
  dfree(pf->hist_X);
  dfree(pf->hist_lw);
  dfree(pf->hist_noise);
  dfree(pf->mc_buf);
  dfree(pf->times);
  dfree(pf->obs_mask);
  dfree(pf->inputs);
  dfree(pf->imp_r);
  dfree(pf->imp_as);
  dfree(pf->out);
  dfree(pf->dev);
  dfree(pf->y);
  dfree(pf->hist_li);
  dfree(pf->hist_ess);
  dfree(pf->hist_res);
  dfree(pf->hist_anc);
  ancestry_free(&pf->cache);
  rw_free(&pf->rw);
  ws_free(&pf->ws);
  delete pf;
  return BIHIP_OK;
}

int bihip_pf_set_observations(bihip_pf* pf, const double* y_host) {
  if (pf == nullptr || y_host == nullptr) return set_error(BIHIP_EINVAL, "NULL argument");
  int rc = ctx_enter(pf->ctx);
  if (rc) return rc;
  BIHIP_HIP(hipMemcpyAsync(pf->y, y_host, sizeof(double) * pf->cfg.T * pf->nobs,
                           hipMemcpyHostToDevice, pf->ctx->stream));
  BIHIP_HIP(hipStreamSynchronize(pf->ctx->stream));
  pf->have_obs = true;
  return BIHIP_OK;
}

int bihip_pf_init(bihip_pf* pf, const double* mean0_host) {
  if (pf == nullptr) return set_error(BIHIP_EINVAL, "NULL filter");
  int rc = ctx_enter(pf->ctx);
  if (rc) return rc;
  hipStream_t s = pf->ctx->stream;
  pf->cur = 0;
  pf->k = 0;
  if ((rc = bihip_init_samples(pf->ctx, &pf->model, pf->cfg.seed, pf->cfg.batch, pf->cfg.P,
                               pf->X[0], pf->ld, mean0_host)))
    return rc;
  BIHIP_HIP(hipMemsetAsync(pf->lws, 0, sizeof(double) * pf->cfg.P, s));
  if (pf->hist_X != nullptr &&
      ((rc = launch_record_state(pf->X[0], pf->ld, pf->model.d, pf->cfg.P, pf->elem, nullptr,
                                 pf->hist_X, 0, s)) ||
       (rc = launch_record_state(pf->lws, pf->cfg.P, 1, pf->cfg.P, 8, nullptr, pf->hist_lw, 0,
                                 s))))
    return rc;
  if (pf->hist_noise != nullptr)
    BIHIP_HIP(hipMemsetAsync(pf->hist_noise, 0,
                             (size_t)(pf->cfg.T + 1) * pf->noise_width * pf->ld * pf->elem, s));
  if (pf->use_cache) {
    if ((rc = ancestry_init(&pf->cache, pf->X[0], pf->ld, s))) return rc;
    pf->cache.live_ub = pf->cfg.P;
  }
  PFDevice d{};
  d.status = BIHIP_OK;
  BIHIP_HIP(hipMemcpyAsync(pf->dev, &d, sizeof(d), hipMemcpyHostToDevice, s));
  BIHIP_HIP(hipMemsetAsync(pf->hist_li, 0, sizeof(double) * pf->cfg.T, s));
  BIHIP_HIP(hipMemsetAsync(pf->hist_ess, 0, sizeof(double) * pf->cfg.T, s));
  BIHIP_HIP(hipMemsetAsync(pf->hist_res, 0, sizeof(int) * pf->cfg.T, s));
  if (pf->imp_r != nullptr)  // the recorded flags of an imposed history are its own
    BIHIP_HIP(hipMemcpyAsync(pf->hist_res, pf->imp_r, sizeof(int) * pf->cfg.T,
                             hipMemcpyDeviceToDevice, s));
  BIHIP_HIP(hipStreamSynchronize(s));
  for (auto& u : pf->ev_used) u = 0;
  pf->ready = true;
  return BIHIP_OK;
}

int bihip_pf_step(bihip_pf* pf, int n) {
  if (pf == nullptr) return set_error(BIHIP_EINVAL, "NULL filter");
  if (!pf->ready) return set_error(BIHIP_ESTATE, "bihip_pf_init has not been called");
  if (!pf->have_obs) return set_error(BIHIP_ESTATE, "no observations set");
  if (n < 0 || pf->k + n > pf->cfg.T)
    return set_error(BIHIP_EINVAL, "steps beyond the schedule (k=%d, n=%d, T=%d)", pf->k, n,
                     pf->cfg.T);
  int rc = ctx_enter(pf->ctx);
  if (rc) return rc;
  hipStream_t s = pf->ctx->stream;
  for (int i = 0; i < n; ++i) {
    if (pf->cfg.use_graph && !pf->timing) {
      if (pf->graph[pf->cur] == nullptr) {
        if ((rc = pf_capture(pf, pf->cur))) return rc;
      }
      BIHIP_HIP(hipGraphLaunch(pf->graph[pf->cur], s));
    } else {
      if ((rc = pf_enqueue_step(pf, s, pf->timing))) return rc;
    }
    pf->cur = 1 - pf->cur;
    ++pf->k;
  }
  return BIHIP_OK;
}

int bihip_pf_term(bihip_pf* pf, double* logZ) {
  if (pf == nullptr) return set_error(BIHIP_EINVAL, "NULL filter");
  int rc = ctx_enter(pf->ctx);
  if (rc) return rc;
  double h[5];
  const int P = pf->cfg.P;
  const int nb = ceil_div(P, RED_TILE);
  int rc2 = ctx_partials(pf->ctx, nb);
  if (rc2) return rc2;
  hipStream_t s = pf->ctx->stream;
  if ((rc = launch_lws_partials(pf->lws, P, pf->ctx->partials, s))) return rc;
  if ((rc = launch_finalize(pf->ctx->partials, nb, P, pf->out, pf->ctx->rw, s))) return rc;
  BIHIP_HIP(hipMemcpyAsync(h, pf->out, sizeof(h), hipMemcpyDeviceToHost, s));
  BIHIP_HIP(hipStreamSynchronize(s));
  if (logZ) *logZ = h[0] + log(h[1]) - log((double)P);
  return bihip_pf_status(pf);
}

int bihip_pf_status(bihip_pf* pf) {
  if (pf == nullptr) return set_error(BIHIP_EINVAL, "NULL filter");
  int rc = ctx_enter(pf->ctx);
  if (rc) return rc;
  PFDevice d;
  ScanState st, rt;
  hipStream_t s = pf->ctx->stream;
  BIHIP_HIP(hipMemcpyAsync(&d, pf->dev, sizeof(d), hipMemcpyDeviceToHost, s));
  BIHIP_HIP(hipMemcpyAsync(&st, pf->ws.state, sizeof(st), hipMemcpyDeviceToHost, s));
  BIHIP_HIP(hipMemcpyAsync(&rt, pf->rw.state, sizeof(rt), hipMemcpyDeviceToHost, s));
  BIHIP_HIP(hipStreamSynchronize(s));
  if (st.error || rt.error)
    return set_error(BIHIP_EHIP, "cross-block wait exceeded its spin limit (scan %u, reduce %u)",
                     st.error, rt.error);
  if (d.status == BIHIP_DEGENERATE)
    return set_error(BIHIP_DEGENERATE, "particle filter degenerated (W <= 0)");
  return BIHIP_OK;
}

int bihip_pf_outputs(bihip_pf* pf, double* li, double* ess, int* res, double* ll) {
  if (pf == nullptr) return set_error(BIHIP_EINVAL, "NULL filter");
  int rc = ctx_enter(pf->ctx);
  if (rc) return rc;
  hipStream_t s = pf->ctx->stream;
  const int T = pf->cfg.T;
  if (li) BIHIP_HIP(hipMemcpyAsync(li, pf->hist_li, sizeof(double) * T, hipMemcpyDeviceToHost, s));
  if (ess)
    BIHIP_HIP(hipMemcpyAsync(ess, pf->hist_ess, sizeof(double) * T, hipMemcpyDeviceToHost, s));
  if (res) BIHIP_HIP(hipMemcpyAsync(res, pf->hist_res, sizeof(int) * T, hipMemcpyDeviceToHost, s));
  PFDevice d;
  BIHIP_HIP(hipMemcpyAsync(&d, pf->dev, sizeof(d), hipMemcpyDeviceToHost, s));
  BIHIP_HIP(hipStreamSynchronize(s));
  if (ll) *ll = d.logLikelihood;
  return BIHIP_OK;
}

int bihip_pf_ancestors(bihip_pf* pf, int k, int* as_host) {
  if (pf == nullptr || as_host == nullptr) return set_error(BIHIP_EINVAL, "NULL argument");
  int rc = ctx_enter(pf->ctx);
  if (rc) return rc;
  hipStream_t s = pf->ctx->stream;
  const int P = pf->cfg.P;
  if (k >= 0) {
    if (pf->hist_anc == nullptr) return set_error(BIHIP_ESTATE, "ancestors were not recorded");
    if (k >= pf->k) return set_error(BIHIP_EINVAL, "step %d not run yet", k);
    BIHIP_HIP(hipMemcpyAsync(as_host, pf->hist_anc + (size_t)k * P, sizeof(int) * P,
                             hipMemcpyDeviceToHost, s));
  } else {
    BIHIP_HIP(hipMemcpyAsync(as_host, pf->as, sizeof(int) * P, hipMemcpyDeviceToHost, s));
  }
  BIHIP_HIP(hipStreamSynchronize(s));
  return BIHIP_OK;
}

int bihip_pf_state(bihip_pf* pf, void** X, int* ld, double** lws) {
  if (pf == nullptr) return set_error(BIHIP_EINVAL, "NULL filter");
  if (X) *X = pf->X[pf->cur];
  if (ld) *ld = pf->ld;
  if (lws) *lws = pf->lws;
  return BIHIP_OK;
}

int bihip_pf_position(bihip_pf* pf) { return pf ? pf->k : -1; }

int bihip_pf_set_schedule(bihip_pf* pf, const double* times_host, const int* observed_host) {
  if (pf == nullptr || times_host == nullptr) return set_error(BIHIP_EINVAL, "NULL argument");
  if (pf->surr || pf->guided)
    return set_error(BIHIP_ESTATE, "detach the learned flow or proposal before setting a schedule");
  const int T = pf->cfg.T;
  int rc;
  for (int k = 0; k < T; ++k) {
    if (!(times_host[k + 1] > times_host[k]))
      return set_error(BIHIP_EINVAL, "schedule times must increase (times[%d] = %g, times[%d] = %g)",
                       k, times_host[k], k + 1, times_host[k + 1]);
    if ((rc = check_interval(&pf->model, times_host[k], times_host[k + 1]))) return rc;
  }
  if ((rc = ctx_enter(pf->ctx))) return rc;
  if (pf->times == nullptr && ((rc = dalloc(&pf->times, (size_t)T + 1)) ||
                               (rc = dalloc(&pf->obs_mask, (size_t)T))))
    return rc;
  std::vector<int> mask(T, 1);
  if (observed_host != nullptr)
    for (int k = 0; k < T; ++k) mask[k] = observed_host[k] != 0;
  pf->times_host.assign(times_host, times_host + T + 1);
  hipStream_t s = pf->ctx->stream;
  BIHIP_HIP(hipMemcpyAsync(pf->times, times_host, sizeof(double) * (T + 1), hipMemcpyHostToDevice,
                           s));
  BIHIP_HIP(hipMemcpyAsync(pf->obs_mask, mask.data(), sizeof(int) * T, hipMemcpyHostToDevice, s));
  BIHIP_HIP(hipStreamSynchronize(s));
  pf->cfg.t0 = times_host[0];
  for (int i = 0; i < 2; ++i) {  // captured steps hold the previous arguments
    if (pf->graph[i]) (void)hipGraphExecDestroy(pf->graph[i]);
    pf->graph[i] = nullptr;
  }
  return BIHIP_OK;
}

int bihip_pf_set_inputs(bihip_pf* pf, const double* inputs_host) {
  if (pf == nullptr) return set_error(BIHIP_EINVAL, "NULL filter");
  if (pf->surr || pf->guided)
    return set_error(BIHIP_ESTATE, "detach the learned flow or proposal before setting inputs");
  int rc = ctx_enter(pf->ctx);
  if (rc) return rc;
  if (inputs_host == nullptr) {
    dfree(pf->inputs);
    pf->inputs = nullptr;
  } else {
    const size_t n = (size_t)pf->cfg.T * 16;
    if (pf->inputs == nullptr && (rc = dalloc(&pf->inputs, n))) return rc;
    BIHIP_HIP(hipMemcpyAsync(pf->inputs, inputs_host, sizeof(double) * n, hipMemcpyHostToDevice,
                             pf->ctx->stream));
    BIHIP_HIP(hipStreamSynchronize(pf->ctx->stream));
  }
  for (int i = 0; i < 2; ++i) {  // captured steps hold the previous arguments
    if (pf->graph[i]) (void)hipGraphExecDestroy(pf->graph[i]);
    pf->graph[i] = nullptr;
  }
  return BIHIP_OK;
}

int bihip_pf_impose_resampling(bihip_pf* pf, const int* resampled_host,
                               const int* ancestors_host) {
  if (pf == nullptr) return set_error(BIHIP_EINVAL, "NULL filter");
  int rc = ctx_enter(pf->ctx);
  if (rc) return rc;
  if (resampled_host == nullptr) {
    BIHIP_HIP(hipStreamSynchronize(pf->ctx->stream));
    dfree(pf->imp_r);
    dfree(pf->imp_as);
    pf->imp_r = nullptr;
    pf->imp_as = nullptr;
    return BIHIP_OK;
  }
  if (ancestors_host == nullptr) return set_error(BIHIP_EINVAL, "NULL ancestors");
  if (pf->cfg.use_graph)
    return set_error(BIHIP_ESTATE, "an imposed history needs a filter without graph replay");
  const int T = pf->cfg.T, P = pf->cfg.P;
  // every ancestor of a flagged step is checked here, before any kernel reads it
  for (int k = 0; k < T; ++k) {
    if (!resampled_host[k]) continue;
    const int* a = ancestors_host + (size_t)k * P;
    for (int i = 0; i < P; ++i)
      if (a[i] < 0 || a[i] >= P)
        return set_error(BIHIP_EINVAL, "ancestor %d of step %d is %d, outside [0, %d)", i, k,
                         a[i], P);
  }
  BIHIP_HIP(hipStreamSynchronize(pf->ctx->stream));
  if (pf->imp_r == nullptr) {
    if ((rc = dalloc(&pf->imp_r, (size_t)T))) return rc;
    if ((rc = dalloc(&pf->imp_as, (size_t)T * P))) {
      dfree(pf->imp_r);
      pf->imp_r = nullptr;
      return rc;
    }
  }
  std::vector<int> flags(T);
  for (int k = 0; k < T; ++k) flags[k] = resampled_host[k] != 0;
  BIHIP_HIP(hipMemcpyAsync(pf->imp_r, flags.data(), sizeof(int) * T, hipMemcpyHostToDevice,
                           pf->ctx->stream));
  if (pf->hist_res != nullptr)
    BIHIP_HIP(hipMemcpyAsync(pf->hist_res, flags.data(), sizeof(int) * T,
                             hipMemcpyHostToDevice, pf->ctx->stream));
  BIHIP_HIP(hipMemcpyAsync(pf->imp_as, ancestors_host, sizeof(int) * (size_t)T * P,
                           hipMemcpyHostToDevice, pf->ctx->stream));
  BIHIP_HIP(hipStreamSynchronize(pf->ctx->stream));
  return BIHIP_OK;
}

int bihip_pf_set_surrogate(bihip_pf* pf, bihip_mlp* mlp, const bihip_surrogate_norm* norm) {
  if (pf == nullptr) return set_error(BIHIP_EINVAL, "NULL filter");
  int rc = ctx_enter(pf->ctx);
  if (rc) return rc;
  if (mlp != nullptr) {
    if (norm == nullptr) return set_error(BIHIP_EINVAL, "NULL normalisation");
    if (pf->model.model != BIHIP_MODEL_L96 || pf->model.d != 40 ||
        pf->model.precision != BIHIP_F32)
      return set_error(BIHIP_EINVAL, "the surrogate flow serves FP32 Lorenz-96 d=40 filters");
    if (pf->guided) return set_error(BIHIP_ESTATE, "detach the guided proposal first");
    if (pf->times != nullptr || pf->inputs != nullptr)
      return set_error(BIHIP_ESTATE,
                       "the surrogate flow serves the uniform schedule and parameters it was trained on");
    SurrogateLaunch S{};
    S.net = mlp_view(mlp);
    if (S.net.L != 4 || S.net.dims[0] != 41 || S.net.dims[4] != 40)
      return set_error(BIHIP_EINVAL, "the surrogate must be a 41-256-256-256-40 network");
    if ((rc = mlp_inference_weights(mlp, S.Winf))) return rc;
    S.x_shift = norm->x_shift;
    S.x_scale = norm->x_scale;
    S.dx_scale = norm->dx_scale;
    S.par = pf->model.par[0];
    S.par_shift = norm->par_shift;
    S.par_scale = norm->par_scale;
    S.noise_sd = (float)pf->model.par[1];
    S.obs_sd = pf->model.par[2];
    pf->sl = S;
  }
  BIHIP_HIP(hipStreamSynchronize(pf->ctx->stream));
  for (int i = 0; i < 2; ++i) {  // captured steps hold the previous transition kernel
    if (pf->graph[i]) (void)hipGraphExecDestroy(pf->graph[i]);
    pf->graph[i] = nullptr;
  }
  if (pf->nparts_exact == 0) pf->nparts_exact = pf->nparts;
  pf->surr = mlp != nullptr;
  pf->nparts = pf->surr ? surrogate_blocks(pf->cfg.P) : pf->nparts_exact;
  if (pf->nparts > pf->nparts_exact)
    return set_error(BIHIP_EINVAL, "partials buffer too small for the surrogate kernel");
  return BIHIP_OK;
}

int bihip_pf_set_guided(bihip_pf* pf, bihip_mlp* mlp, const bihip_guided_norm* norm) {
  if (pf == nullptr) return set_error(BIHIP_EINVAL, "NULL filter");
  int rc = ctx_enter(pf->ctx);
  if (rc) return rc;
  if (mlp != nullptr) {
    if (norm == nullptr) return set_error(BIHIP_EINVAL, "NULL normalisation");
    if (pf->model.model != BIHIP_MODEL_L96 || pf->model.d != 200 ||
        pf->model.precision != BIHIP_F32)
      return set_error(BIHIP_EINVAL, "the guided proposal serves FP32 Lorenz-96 d=200 filters");
    if (pf->surr) return set_error(BIHIP_ESTATE, "detach the surrogate flow first");
    if (pf->times != nullptr || pf->inputs != nullptr)
      return set_error(BIHIP_ESTATE,
                       "the guided proposal serves the uniform schedule and parameters it was trained on");
    GuidedLaunch G{};
    G.net = mlp_view(mlp);
    if (G.net.L != 4 || G.net.dims[0] != 301 || G.net.dims[4] != 400)
      return set_error(BIHIP_EINVAL, "the proposal must be a 301-256-256-256-400 network");
    if ((rc = mlp_inference_weights(mlp, G.Winf, G.binf, 16))) return rc;
    G.x_shift = norm->x_shift;
    G.x_scale = norm->x_scale;
    G.innov_scale = norm->innov_scale;
    G.par = pf->model.par[0];
    G.par_shift = norm->par_shift;
    G.par_scale = norm->par_scale;
    G.noise_sd = (float)pf->model.par[1];
    G.obs_sd = pf->model.par[2];
    pf->gl = G;
  }
  BIHIP_HIP(hipStreamSynchronize(pf->ctx->stream));
  for (int i = 0; i < 2; ++i) {  // captured steps hold the previous kernels
    if (pf->graph[i]) (void)hipGraphExecDestroy(pf->graph[i]);
    pf->graph[i] = nullptr;
  }
  if (pf->nparts_exact == 0) pf->nparts_exact = pf->nparts;
  pf->guided = mlp != nullptr;
  pf->nparts = pf->guided ? guided_blocks(pf->cfg.P) : pf->nparts_exact;
  if (pf->nparts > pf->nparts_exact)
    return set_error(BIHIP_EINVAL, "partials buffer too small for the guided kernel");
  return BIHIP_OK;
}

int bihip_pf_paths(bihip_pf* pf, const int* idx_host, int n, double* paths_host) {
  if (pf == nullptr || (n > 0 && (idx_host == nullptr || paths_host == nullptr)))
    return set_error(BIHIP_EINVAL, "NULL argument");
  if (!pf->use_cache && (pf->hist_X == nullptr || pf->hist_anc == nullptr))
    return set_error(BIHIP_ESTATE, "paths need record = 2 or 3");
  if (n < 0) return set_error(BIHIP_EINVAL, "n < 0");
  for (int q = 0; q < n; ++q)
    if (idx_host[q] < 0 || idx_host[q] >= pf->cfg.P)
      return set_error(BIHIP_EINVAL, "particle %d outside [0, %d)", idx_host[q], pf->cfg.P);
  int rc = ctx_enter(pf->ctx);
  if (rc) return rc;
  hipStream_t s = pf->ctx->stream;
  const int K = pf->k, D = pf->model.d;
  int* idx = nullptr;
  double* out = nullptr;
  if ((rc = dalloc(&idx, (size_t)n)) || (rc = dalloc(&out, (size_t)n * (K + 1) * D))) {
    dfree(idx);
    return rc;
  }
  hipError_t e = hipMemcpyAsync(idx, idx_host, sizeof(int) * n, hipMemcpyHostToDevice, s);
  if (e == hipSuccess &&
      (rc = pf->use_cache ? ancestry_paths(&pf->cache, K, idx, n, out, s)
                          : launch_paths(pf->hist_anc, pf->hist_X, pf->ld, D, pf->cfg.P, K,
                                         pf->elem, idx, n, out, s)) == BIHIP_OK)
    e = hipMemcpyAsync(paths_host, out, sizeof(double) * n * (K + 1) * D, hipMemcpyDeviceToHost,
                       s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  dfree(idx);
  dfree(out);
  if (rc) return rc;
  if (e != hipSuccess) return set_error(BIHIP_EHIP, "paths: %s", hipGetErrorString(e));
  return BIHIP_OK;
}

int bihip_pf_sample_path(bihip_pf* pf, double* path_host, int* index_host) {
  if (pf == nullptr || path_host == nullptr) return set_error(BIHIP_EINVAL, "NULL argument");
  int rc = ctx_enter(pf->ctx);
  if (rc) return rc;
  hipStream_t s = pf->ctx->stream;
  const int P = pf->cfg.P;
  std::vector<double> lw(P);
  BIHIP_HIP(hipMemcpyAsync(lw.data(), pf->lws, sizeof(double) * P, hipMemcpyDeviceToHost, s));
  BIHIP_HIP(hipStreamSynchronize(s));
  // one multinomial draw over the final log-weights: the first particle whose
  // cumulative weight exceeds u W (the multinomial resampler's upper bound)
  double mx = -INFINITY;
  for (double v : lw)
    if (!std::isnan(v) && v > mx) mx = v;
  std::vector<double> Ws(P);
  double acc = 0.0;
  for (int i = 0; i < P; ++i) {
    acc = acc + (std::isnan(lw[i]) ? 0.0 : std::exp(lw[i] - mx));
    Ws[i] = acc;
  }
  if (!(acc > 0.0)) return set_error(BIHIP_DEGENERATE, "all weights are zero");
  const double u = bihip_uniform_host(pf->cfg.seed, 0u, (uint32_t)pf->k, BIHIP_TAG_PATH,
                                      pf->cfg.batch);
  const double target = acc * u;
  const int p = (int)(std::upper_bound(Ws.begin(), Ws.end(), target) - Ws.begin());
  const int pick = p < P ? p : P - 1;
  if (index_host) *index_host = pick;
  return bihip_pf_paths(pf, &pick, 1, path_host);
}

int bihip_pf_cache_nodes(bihip_pf* pf, long* live, long* capacity) {
  if (pf == nullptr) return set_error(BIHIP_EINVAL, "NULL filter");
  if (!pf->use_cache) return set_error(BIHIP_ESTATE, "the filter keeps no ancestry tree (record != 3)");
  int rc = ctx_enter(pf->ctx);
  if (rc) return rc;
  int n = 0;
  BIHIP_HIP(hipMemcpyAsync(&n, pf->cache.live, sizeof(int), hipMemcpyDeviceToHost,
                           pf->ctx->stream));
  BIHIP_HIP(hipStreamSynchronize(pf->ctx->stream));
  if (live) *live = n;
  if (capacity) *capacity = pf->cache.cap;
  return BIHIP_OK;
}

int bihip_pf_write_netcdf(bihip_pf* pf, const char* path) {
  if (pf == nullptr || path == nullptr) return set_error(BIHIP_EINVAL, "NULL argument");
  if (pf->hist_X == nullptr || pf->hist_lw == nullptr || pf->hist_anc == nullptr)
    return set_error(BIHIP_ESTATE, "the NetCDF output needs a filter created with record = 2");
  if (!pf->ready) return set_error(BIHIP_ESTATE, "bihip_pf_init has not been called");
  int rc = ctx_enter(pf->ctx);
  if (rc) return rc;
  hipStream_t s = pf->ctx->stream;
  const int P = pf->cfg.P, nr = pf->k + 1;
  const size_t xbytes = (size_t)pf->model.d * pf->ld * pf->elem;
  std::vector<uint8_t> X(xbytes);
  const size_t zbytes = (size_t)pf->noise_width * pf->ld * pf->elem;
  std::vector<uint8_t> Z(zbytes);
  std::vector<double> lw(P);
  std::vector<int> as(P);
  bihip_pfout* o = nullptr;
  if ((rc = bihip_pfout_create(path, &pf->model, nr, P, &o))) return rc;
  for (int k = 0; k < nr && rc == BIHIP_OK; ++k) {
    const uint8_t* hx = static_cast<const uint8_t*>(pf->hist_X) + (size_t)k * xbytes;
    hipError_t e = hipMemcpyAsync(X.data(), hx, xbytes, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(lw.data(), pf->hist_lw + (size_t)k * P, sizeof(double) * P,
                         hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && k > 0)
      e = hipMemcpyAsync(as.data(), pf->hist_anc + (size_t)(k - 1) * P, sizeof(int) * P,
                         hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && pf->hist_noise != nullptr)
      e = hipMemcpyAsync(Z.data(), static_cast<const uint8_t*>(pf->hist_noise) + k * zbytes,
                         zbytes, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      rc = set_error(BIHIP_EHIP, "copy of output %d: %s", k, hipGetErrorString(e));
      break;
    }
    if (k == 0)  // the initial samples have themselves as ancestors
      for (int p = 0; p < P; ++p) as[p] = p;
    const double tk = pf->times != nullptr ? pf->times_host[k] : pf->cfg.t0 + k * pf->cfg.dt;
    if ((rc = bihip_pfout_write_time(o, k, tk)) ||
        (rc = bihip_pfout_write_state(o, k, X.data(), pf->ld, as.data())) ||
        (pf->hist_noise != nullptr && (rc = bihip_pfout_write_noise(o, k, Z.data(), pf->ld))) ||
        (rc = bihip_pfout_write_log_weights(o, k, lw.data())))
      break;
  }
  if (rc == BIHIP_OK) {
    PFDevice d{};
    BIHIP_HIP(hipMemcpy(&d, pf->dev, sizeof(d), hipMemcpyDeviceToHost));
    rc = bihip_pfout_write_log_likelihood(o, d.logLikelihood);
  }
  if (rc == BIHIP_OK) rc = bihip_pfout_write_clock(o, 0.0);
  const int rc2 = bihip_pfout_close(o);
  return rc != BIHIP_OK ? rc : rc2;
}

int bihip_pf_set_timing(bihip_pf* pf, int mode) {
  if (pf == nullptr) return set_error(BIHIP_EINVAL, "NULL filter");
  if (mode < 0 || mode > BIHIP_TIMING_TRANSITION)
    return set_error(BIHIP_EINVAL, "timing mode must be 0, BIHIP_TIMING_ALL or BIHIP_TIMING_TRANSITION");
  pf->timing = mode;
  for (auto& u : pf->ev_used) u = 0;
  return BIHIP_OK;
}

int bihip_pf_kernel_time(bihip_pf* pf, int kind, double* ms, long* count) {
  if (pf == nullptr || kind < 0 || kind > 3) return set_error(BIHIP_EINVAL, "bad argument");
  int rc = ctx_enter(pf->ctx);
  if (rc) return rc;
  BIHIP_HIP(hipStreamSynchronize(pf->ctx->stream));
  double total = 0.0;
  const size_t used = pf->ev_used[kind];
  for (size_t i = 0; i + 1 < used; i += 2) {
    float t = 0.f;
    BIHIP_HIP(hipEventElapsedTime(&t, pf->ev[kind][i], pf->ev[kind][i + 1]));
    total += t;
  }
  if (ms) *ms = total;
  if (count) *count = (long)(used / 2);
  return BIHIP_OK;
}

}  // extern "C"

namespace bihip {
void philox_host_block(uint64_t seed, uint32_t index, uint32_t step, uint32_t tag,
                       uint32_t batch, uint32_t block, uint32_t out[4]) {
  out[0] = index;
  out[1] = step;
  out[2] = tag | (batch << 8);
  out[3] = block;
  host_philox(out, (uint32_t)seed, (uint32_t)(seed >> 32));
}
}  // namespace bihip
