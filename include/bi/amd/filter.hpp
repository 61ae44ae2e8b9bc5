/**
 * @file
 *
 * C++ shims of LibBi's particle-filter step for the ON_DEVICE location:
 * models, State / BootstrapPFState, the primitives the resamplers use,
 * Resampler<R> with the systematic / stratified / multinomial bases,
 * RK4Integrator / DOPRI5Integrator, BootstrapPF and Filter<F>.  Each class
 * keeps the reference's name, template parameters and method signatures
 * (files cited per class) and forwards to the bihip C ABI.
 */
#ifndef BI_AMD_FILTER_HPP
#define BI_AMD_FILTER_HPP

#include "core.hpp"

namespace bi {

// ================================================================ models
namespace amd {

/**
 * Runtime description of a model on the device: the role of the generated
 * Model<name> (share/tt/cpp/model.hpp.tt) with its ode block settings.
 */
template <class Real>
class DeviceModel {
 public:
  typedef Real real_type;
  static const int precision = sizeof(Real) == 8 ? BIHIP_F64 : BIHIP_F32;

  explicit DeviceModel(const bihip_model& d) : desc(d) { desc.precision = precision; }
  int getDynSize() const { return desc.d; }
  int getNumObs() const { return bihip_model_nobs(&desc); }
  const bihip_model& description() const { return desc; }
  bihip_model& description() { return desc; }

 protected:
  bihip_model desc;
};

/** The PZ model of docs/src/users.tex:36-85 (RK4, h = 0.1 by default). */
template <class Real = double>
class PZ : public DeviceModel<Real> {
 public:
  PZ(double mu = 0.5, double sigma = 0.2, double obs_sd = 0.2)
      : DeviceModel<Real>(make(mu, sigma, obs_sd)) {}

 private:
  static bihip_model make(double mu, double sigma, double obs_sd) {
    bihip_model m{};
    m.model = BIHIP_MODEL_PZ;
    m.d = 2;
    m.integrator = BIHIP_RK4;
    m.nsteps = 1000;
    m.h0 = 0.1;
    m.atol = m.rtol = 1e-5;
    const double par[] = {mu, sigma, 0.25, 0.3, 0.1, 0.1, obs_sd, std::log(2.0), 0.2,
                          std::log(2.0), 0.1};
    for (int i = 0; i < 11; ++i) m.par[i] = par[i];
    return m;
  }
};

/** Lorenz-96 with additive noise (DOPRI5, rtol = atol = 1e-5, h0 = 0.05). */
template <int Dim = 40, class Real = float>
class Lorenz96 : public DeviceModel<Real> {
 public:
  Lorenz96(double F = 8.0, double noise_sd = 0.1, double obs_sd = 1.0, double init_sd = 1.0)
      : DeviceModel<Real>(make(F, noise_sd, obs_sd, init_sd)) {}

 private:
  static bihip_model make(double F, double noise_sd, double obs_sd, double init_sd) {
    bihip_model m{};
    m.model = BIHIP_MODEL_L96;
    m.d = Dim;
    m.integrator = BIHIP_DOPRI5;
    m.nsteps = 1000;
    m.h0 = 0.05;
    m.atol = m.rtol = 1e-5;
    m.par[0] = F;
    m.par[1] = noise_sd;
    m.par[2] = obs_sd;
    m.par[3] = init_sd;
    return m;
  }
};

/** Scalar linear-Gaussian model dx/dt = -lambda x (Kalman-checkable). */
template <class Real = double>
class LinearGaussian : public DeviceModel<Real> {
 public:
  LinearGaussian(double lambda = 0.5, double noise_sd = 0.5, double obs_sd = 0.5,
                 double m0 = 0.0, double s0 = 1.0)
      : DeviceModel<Real>(make(lambda, noise_sd, obs_sd, m0, s0)) {}

 private:
  static bihip_model make(double lambda, double noise_sd, double obs_sd, double m0, double s0) {
    bihip_model m{};
    m.model = BIHIP_MODEL_LG;
    m.d = 1;
    m.integrator = BIHIP_DOPRI5;
    m.nsteps = 1000;
    m.h0 = 0.1;
    m.atol = m.rtol = 1e-10;
    const double par[] = {lambda, noise_sd, obs_sd, m0, s0};
    for (int i = 0; i < 5; ++i) m.par[i] = par[i];
    return m;
  }
};

}  // namespace amd

// ================================================================ state
template <class B, Location L>
class State;

/**
 * Particle state on the device (share/src/bi/state/State.hpp:73-937): the
 * dynamic variables as a column-major P x d matrix whose leading dimension
 * is P rounded up to 64, plus a second buffer for out-of-place updates.
 */
template <class B>
class State<B, ON_DEVICE> {
 public:
  typedef typename B::real_type real_type;
  typedef device_vector<double> vector_type;
  typedef device_vector<int> int_vector_type;
  typedef device_vector<int> temp_int_vector_type;

  State(const B& m, int P)
      : P(P), ld((P + 63) / 64 * 64), d(m.getDynSize()), X(ld * d), Y(ld * d), clock(0),
        desc(m.description()) {}
  int size() const { return P; }
  /** The device description of the model the state belongs to (the
   * reference resolves it from the type B at compile time). */
  const bihip_model& model() const { return desc; }
  int lead() const { return ld; }
  real_type* buf() { return X.buf(); }
  /** Swap in the result of an out-of-place update. */
  real_type* next() { return Y.buf(); }
  void flip() { X.swap(Y); }
  /** In-place gather of the dynamic variables (State.hpp:905-909). */
  void gather(const int_vector_type& as) {
    amd::check(bihip_gather_rows(amd::ctx(), as.buf(), P, d, X.buf(), ld, sizeof(real_type)));
  }
  std::vector<real_type> to_host() const { return X.to_host(); }

 protected:
  int P, ld, d;
  device_vector<real_type> X, Y;

 public:
  long clock;

 private:
  bihip_model desc;
};

/**
 * Filter state (share/src/bi/state/FilterState.hpp:91-95 and
 * share/src/bi/state/BootstrapPFState.hpp:127-212): log-weights, ancestors,
 * ESS, the running log-likelihood and its increments.
 */
template <class B, Location L>
class BootstrapPFState;

template <class B>
class BootstrapPFState<B, ON_DEVICE> : public State<B, ON_DEVICE> {
 public:
  typedef State<B, ON_DEVICE> parent_type;
  typedef typename parent_type::vector_type vector_type;
  typedef typename parent_type::int_vector_type int_vector_type;

  BootstrapPFState(const B& m, int P, int Y)
      : parent_type(m, P), ess(0.0), logLikelihood(0.0), logIncrements(Y, 0.0), lws(P), as(P) {}
  vector_type& logWeights() { return lws; }
  const vector_type& logWeights() const { return lws; }
  int_vector_type& ancestors() { return as; }
  const int_vector_type& ancestors() const { return as; }
  /** BootstrapPFState::gather: state rows, then the ancestry output. */
  void gather(const ScheduleElement now, const int_vector_type& as1) {
    parent_type::gather(as1);
    if (now.hasOutput()) as = as1;
  }

  double ess;
  double logLikelihood;
  std::vector<double> logIncrements;

 private:
  vector_type lws;
  int_vector_type as;
};

// ================================================================ primitives
// share/src/bi/primitive/vector_primitive.hpp
inline double ess_reduce(const device_vector<double>& lws, double* lW) {
  double ess = 0.0, l = 0.0;
  amd::check(bihip_ess_reduce(amd::ctx(), lws.buf(), lws.size(), &ess, &l));
  if (lW) *lW = l;
  return ess;
}
inline double logsumexp_reduce(const device_vector<double>& x) {
  double r = 0.0;
  amd::check(bihip_logsumexp_reduce(amd::ctx(), x.buf(), x.size(), &r));
  return r;
}
inline double max_reduce(const device_vector<double>& x) {
  double r = 0.0;
  amd::check(bihip_max_reduce(amd::ctx(), x.buf(), x.size(), &r));
  return r;
}
inline double sumexpu_inclusive_scan(const device_vector<double>& x, device_vector<double>& y) {
  double mx = 0.0;
  y.resize(x.size());
  amd::check(bihip_sumexpu_inclusive_scan(amd::ctx(), x.buf(), x.size(), y.buf(), &mx));
  return mx;
}
inline void set_elements(device_vector<double>& x, double v) {
  amd::check(bihip_fill(amd::ctx(), x.buf(), x.size(), v));
}
inline void seq_elements(device_vector<int>& x, int) {
  amd::check(bihip_iota(amd::ctx(), x.buf(), x.size()));
}

// share/src/bi/resampler/misc.hpp:19-91
inline void ancestorsToOffspring(const device_vector<int>& as, device_vector<int>& os) {
  amd::check(bihip_ancestors_to_offspring(amd::ctx(), as.buf(), as.size(), os.size(), os.buf()));
}
inline void cumulativeOffspringToAncestors(const device_vector<int>& Os, device_vector<int>& as) {
  amd::check(bihip_cumulative_offspring_to_ancestors(amd::ctx(), Os.buf(), Os.size(), as.buf()));
}
inline void cumulativeOffspringToAncestorsPermute(const device_vector<int>& Os,
                                                  device_vector<int>& as) {
  amd::check(bihip_cumulative_offspring_to_ancestors_permute(amd::ctx(), Os.buf(), Os.size(),
                                                             as.buf()));
}
inline void offspringToAncestorsPermute(const device_vector<int>& os, device_vector<int>& as) {
  amd::check(bihip_offspring_to_ancestors_permute(amd::ctx(), os.buf(), os.size(), as.buf()));
}

// ================================================================ resamplers
/** share/src/bi/resampler/ScanResampler.hpp:13-17 */
template <Location L>
struct ScanResamplerPrecompute;
template <>
struct ScanResamplerPrecompute<ON_DEVICE> {
  device_vector<double> Ws;
  double W = 0.0;
};

/** share/src/bi/resampler/ScanResampler.hpp:24-40 */
class ScanResampler {
 public:
  template <Location L>
  void precompute(const device_vector<double>& lws, ScanResamplerPrecompute<L>& pre) {
    sumexpu_inclusive_scan(lws, pre.Ws);
    amd::check(bihip_memcpy_to_host(amd::ctx(), &pre.W, pre.Ws.buf() + (pre.Ws.size() - 1),
                                    sizeof(double)));
  }
};

/** share/src/bi/resampler/SystematicResampler.hpp:61-156 */
class SystematicResampler : public ScanResampler {
 public:
  static const uint32_t TAG = BIHIP_TAG_SYSTEMATIC;
  template <Location L>
  void cumulativeOffspring(Random& rng, const device_vector<double>& lws, const int P,
                           device_vector<int>& Os, ScanResamplerPrecompute<L>& pre) {
    if (pre.W > 0) {
      const double a = rng.uniform(0.0, 1.0);  // offset into strata
      Os.resize(lws.size());
      amd::check(bihip_systematic_cumulative_offspring(amd::ctx(), pre.Ws.buf(), lws.size(), P,
                                                       a, Os.buf()));
    } else {
      throw ParticleFilterDegeneratedException();
    }
  }
  template <Location L>
  void ancestors(Random& rng, const device_vector<double>& lws, device_vector<int>& as,
                 ScanResamplerPrecompute<L>& pre) {
    device_vector<int> Os(lws.size());
    cumulativeOffspring(rng, lws, as.size(), Os, pre);
    cumulativeOffspringToAncestors(Os, as);
  }
  template <Location L>
  void ancestorsPermute(Random& rng, const device_vector<double>& lws, device_vector<int>& as,
                        ScanResamplerPrecompute<L>& pre) {
    device_vector<int> Os(lws.size());
    cumulativeOffspring(rng, lws, as.size(), Os, pre);
    cumulativeOffspringToAncestorsPermute(Os, as);
  }
};

/** share/src/bi/resampler/StratifiedResampler.hpp:24-132 */
class StratifiedResampler : public ScanResampler {
 public:
  static const uint32_t TAG = BIHIP_TAG_STRATIFIED;
  template <Location L>
  void cumulativeOffspring(Random& rng, const device_vector<double>& lws, const int P,
                           device_vector<int>& Os, ScanResamplerPrecompute<L>& pre) {
    if (pre.W > 0) {
      device_vector<double> alphas(P);
      amd::check(bihip_uniforms(amd::ctx(), rng.seed(), rng.currentStep(), TAG, rng.batch(), P,
                                alphas.buf()));
      Os.resize(lws.size());
      amd::check(bihip_stratified_cumulative_offspring(amd::ctx(), pre.Ws.buf(), lws.size(), P,
                                                       alphas.buf(), Os.buf()));
    } else {
      throw ParticleFilterDegeneratedException();
    }
  }
  template <Location L>
  void ancestorsPermute(Random& rng, const device_vector<double>& lws, device_vector<int>& as,
                        ScanResamplerPrecompute<L>& pre) {
    device_vector<int> Os(lws.size());
    cumulativeOffspring(rng, lws, as.size(), Os, pre);
    cumulativeOffspringToAncestorsPermute(Os, as);
  }
};

/**
 * share/src/bi/resampler/MultinomialResampler.hpp:82-107.  ancestors() draws
 * a_j = #{i : Ws_i <= W U_j}; ancestorsPermute() lays the copies out as the
 * cumulative-offspring permutation does (the reference permutes the drawn
 * vector in place with ResamplerHost::permute, a different layout of the
 * same offspring counts).
 */
class MultinomialResampler : public ScanResampler {
 public:
  static const uint32_t TAG = BIHIP_TAG_MULTINOMIAL;
  template <Location L>
  void ancestors(Random& rng, const device_vector<double>& lws, device_vector<int>& as,
                 ScanResamplerPrecompute<L>& pre) {
    if (!(pre.W > 0)) throw ParticleFilterDegeneratedException();
    const int n = as.size();
    device_vector<double> U(n);
    amd::check(bihip_uniforms(amd::ctx(), rng.seed(), rng.currentStep(), TAG, rng.batch(), n,
                              U.buf()));
    amd::check(bihip_multinomial_ancestors(amd::ctx(), pre.Ws.buf(), lws.size(), n, U.buf(),
                                           as.buf()));
  }
  template <Location L>
  void ancestorsPermute(Random& rng, const device_vector<double>& lws, device_vector<int>& as,
                        ScanResamplerPrecompute<L>& pre) {
    device_vector<int> a0(as.size()), os(lws.size());
    ancestors(rng, lws, a0, pre);
    ancestorsToOffspring(a0, os);
    offspringToAncestorsPermute(os, as);
  }
};

/** share/src/bi/resampler/Resampler.hpp:33-182 */
template <class R>
class Resampler : public R {
 public:
  Resampler(const double essRel = 0.5, const bool anytime = false)
      : essRel(essRel), maxLogWeight(0.0), anytime(anytime) {
    if (!(essRel >= 0.0 && essRel <= 1.0)) throw std::invalid_argument("essRel in [0,1]");
  }
  double getEssRel() const { return essRel; }
  void setEssRel(const double e) { essRel = e; }
  double getMaxLogWeight() const { return maxLogWeight; }
  void setMaxLogWeight(const double m) { maxLogWeight = m; }

  double reduce(const device_vector<double>& lws, double* lW) {
    double ess = ess_reduce(lws, lW);
    if (anytime) {
      const int P = lws.size();
      *lW += std::log(P / (P - 1.0));
    }
    return ess;
  }

  template <class S1>
  bool resample(Random& rng, const ScheduleElement now, S1& s) {
    bool r = (now.isObserved() || now.hasBridge()) && s.ess < essRel * s.size();
    if (r) {
      ScanResamplerPrecompute<ON_DEVICE> pre;
      typename S1::temp_int_vector_type as1(s.size());
      rng.setStream(now.index(), R::TAG);
      R::precompute(s.logWeights(), pre);
      R::ancestorsPermute(rng, s.logWeights(), as1, pre);
      s.gather(now, as1);
      set_elements(s.logWeights(), s.logLikelihood);
    } else if (now.hasOutput()) {
      seq_elements(s.ancestors(), 0);
    }
    return r;
  }

 protected:
  double essRel;
  double maxLogWeight;
  bool anytime;
};

// ================================================================ integrators
namespace amd {
/** The ode block's settings (share/src/bi/ode/IntegratorConstants.hpp:29-44),
 * per host thread: the generated ode block sets them before it integrates
 * (share/tt/cpp/block/ode.hpp.tt:40-43). */
struct OdeSettings {
  bool set = false;
  double h0 = 0.0, atoler = 0.0, rtoler = 0.0;
};
inline OdeSettings& ode_settings() {
  thread_local OdeSettings o;
  return o;
}

/** The deterministic flow of the state s over [t1, t2] with integrator alg:
 * the state's model description, with the ode settings of this thread when
 * bi_ode_set has been called. */
template <class B>
inline void flow(int alg, double t1, double t2, State<B, ON_DEVICE>& s) {
  if (!(std::fabs(t2 - t1) > 0.0)) return;  // as the reference's update
  if (t1 > t2) throw DeviceError(BIHIP_EINVAL, "bihip: integrator update needs t1 <= t2");
  bihip_model d = s.model();
  d.integrator = alg;
  const OdeSettings& o = ode_settings();
  if (o.set) {
    d.h0 = o.h0;
    d.atol = o.atoler;
    d.rtol = o.rtoler;
  }
  check(bihip_flow(ctx(), &d, t1, t2, s.size(), s.buf(), s.lead(), nullptr, nullptr));
}
}  // namespace amd

/** share/src/bi/ode/RK4Integrator.hpp:21-75: the ode block alone, no noise,
 * called as the generated ode block calls it (share/tt/cpp/block/ode.hpp.tt:47-48). */
template <class B, class S = void>
class RK4Integrator {
 public:
  template <class T1>
  static void update(const T1 t1, const T1 t2, State<B, ON_DEVICE>& s) {
    amd::flow(BIHIP_RK4, (double)t1, (double)t2, s);
  }
};

/** share/src/bi/ode/DOPRI5Integrator.hpp:22-76 (share/tt/cpp/block/ode.hpp.tt:49-50). */
template <class B, class S = void>
class DOPRI5Integrator {
 public:
  template <class T1>
  static void update(const T1 t1, const T1 t2, State<B, ON_DEVICE>& s) {
    amd::flow(BIHIP_DOPRI5, (double)t1, (double)t2, s);
  }
};

}  // namespace bi

/** share/src/bi/ode/IntegratorConstants.hpp:29-44: initial step size and
 * tolerances of the next integrator updates on this host thread. */
inline void bi_ode_set(const double h0, const double atoler, const double rtoler) {
  bi::amd::OdeSettings& o = bi::amd::ode_settings();
  o.set = true;
  o.h0 = h0;
  o.atoler = atoler;
  o.rtoler = rtoler;
}

namespace bi {

// ================================================================ observer, forcer, output
/** Observations of T times, nobs values each, on the device (Observer.hpp:115-126). */
class Observer {
 public:
  Observer(const std::vector<double>& y, int nobs) : T((int)y.size() / nobs), nobs(nobs) {
    Y.from_host(y);
  }
  int size() const { return T; }
  /** Device pointer to the observations at observation index k. */
  const double* getMask(const int k) const { return Y.buf() + (size_t)k * nobs; }

 private:
  int T, nobs;
  device_vector<double> Y;
};

/** No time-varying inputs (the Forcer of share/src/bi/simulator/Forcer.hpp). */
class NullForcer {};

/** share/src/bi/null/ParticleFilterNullBuffer.hpp: outputs are discarded. */
class ParticleFilterNullBuffer {
 public:
  template <class S1>
  void write(const int, const double, const S1&) {}
  template <class S1>
  void writeT(const S1&) {}
};

/** ParticleFilterBuffer::write (share/src/bi/buffer/ParticleFilterBuffer.hpp:67-81)
 * into host memory: times, ancestors, log-weights and the final log-likelihood. */
class ParticleFilterMemoryBuffer {
 public:
  template <class S1>
  void write(const int k, const double t, const S1& s) {
    if ((int)times.size() <= k) {
      times.resize(k + 1);
      ancestors.resize(k + 1);
      logWeights.resize(k + 1);
    }
    times[k] = t;
    ancestors[k] = s.ancestors().to_host();
    logWeights[k] = s.logWeights().to_host();
  }
  template <class S1>
  void writeT(const S1& s) {
    logLikelihood = s.logLikelihood;
  }
  std::vector<double> times;
  std::vector<std::vector<int> > ancestors;
  std::vector<std::vector<double> > logWeights;
  double logLikelihood = 0.0;
};

/**
 * share/src/bi/netcdf/ParticleFilterNetCDFBuffer.hpp behind ParticleFilterBuffer::write /
 * writeT (share/src/bi/buffer/ParticleFilterBuffer.hpp:69-82): every output time's state,
 * ancestors and log-weights, then the log-likelihood, in LibBi's ParticleFilter NetCDF
 * schema (bihip_pfout_*).  T is the number of output times.
 */
class ParticleFilterNetCDFBuffer {
 public:
  template <class B>
  ParticleFilterNetCDFBuffer(const B& m, const int P, const int T, const std::string& file)
      : o(nullptr) {
    amd::check(bihip_pfout_create(file.c_str(), &m.description(), T, P, &o));
  }
  ~ParticleFilterNetCDFBuffer() { bihip_pfout_close(o); }
  ParticleFilterNetCDFBuffer(const ParticleFilterNetCDFBuffer&) = delete;
  ParticleFilterNetCDFBuffer& operator=(const ParticleFilterNetCDFBuffer&) = delete;

  template <class S1>
  void write(const int k, const double t, const S1& s) {
    amd::check(bihip_pfout_write_time(o, k, t));
    const auto X = s.to_host();
    const std::vector<int> as = s.ancestors().to_host();
    amd::check(bihip_pfout_write_state(o, k, X.data(), s.lead(), as.data()));
    const std::vector<double> lws = s.logWeights().to_host();
    amd::check(bihip_pfout_write_log_weights(o, k, lws.data()));
  }
  template <class S1>
  void writeT(const S1& s) {
    amd::check(bihip_pfout_write_log_likelihood(o, s.logLikelihood));
    amd::check(bihip_pfout_write_clock(o, (double)s.clock));
  }

 private:
  bihip_pfout* o;
};

// ================================================================ filter
/**
 * Bootstrap particle filter (share/src/bi/filter/BootstrapPF.hpp:26-196),
 * with Simulator::init / predict (share/src/bi/simulator/Simulator.hpp:253-484).
 */
template <class B, class F, class O, class R>
class BootstrapPF {
 public:
  BootstrapPF(B& m, F& in, O& obs, R& resam) : m(m), in(in), obs(obs), resam(resam) {}

  /** Initial state from the model's initial distribution; lws = 0. */
  template <class S1, class IO1>
  void init(Random& rng, S1& s, IO1& out, const std::vector<double>& mean0 = {}) {
    (void)out;
    amd::check(bihip_init_samples(amd::ctx(), &m.description(), rng.seed(), rng.batch(),
                                  s.size(), s.buf(), s.lead(),
                                  mean0.empty() ? nullptr : mean0.data()));
    set_elements(s.logWeights(), 0.0);
    seq_elements(s.ancestors(), 0);
    s.ess = 0.0;
    s.logLikelihood = 0.0;
    std::fill(s.logIncrements.begin(), s.logIncrements.end(), 0.0);
  }

  template <class S1, class IO1>
  void step(Random& rng, ScheduleIterator& iter, const ScheduleIterator last, S1& s, IO1& out) {
    do {
      this->resample(rng, *iter, s);
      ++iter;
      this->predict(rng, *iter, s);
      this->correct(rng, *iter, s);
      this->output(*iter, s, out);
    } while (iter + 1 != last && !iter->isObserved());
  }

  /** Simulator::predict: transitionSamples from the previous time to `next`. */
  template <class S1>
  void predict(Random& rng, const ScheduleElement next, S1& s) {
    amd::check(bihip_transition(amd::ctx(), &m.description(), rng.seed(), rng.batch(),
                                (uint32_t)next.index(), next.getFrom(), next.getTo(), s.size(),
                                s.buf(), s.next(), s.lead(), nullptr, nullptr));
    s.flip();
  }

  template <class S1>
  void correct(Random& rng, const ScheduleElement now, S1& s) {
    (void)rng;
    if (now.isObserved()) {
      amd::check(bihip_obs_logdensity(amd::ctx(), &m.description(), obs.getMask(now.indexObs()),
                                      s.size(), s.buf(), s.lead(), s.logWeights().buf()));
      double lW;
      s.ess = resam.reduce(s.logWeights(), &lW);
      s.logIncrements[now.indexObs()] = lW - s.logLikelihood;
      s.logLikelihood = lW;
    }
  }

  template <class S1>
  void resample(Random& rng, const ScheduleElement now, S1& s) {
    resam.resample(rng, now, s);
  }

  template <class S1>
  void term(S1& s) {
    s.logLikelihood = logsumexp_reduce(s.logWeights()) - std::log(double(s.size()));
  }

  template <class S1, class IO1>
  void output0(const S1&, IO1&) {}
  template <class S1, class IO1>
  void output(const ScheduleElement now, const S1& s, IO1& out) {
    if (now.hasOutput()) out.write(now.indexOutput(), now.getTime(), s);
  }
  template <class S1, class IO1>
  void outputT(const S1& s, IO1& out) {
    out.writeT(s);
  }

 protected:
  B& m;
  F& in;
  O& obs;
  R& resam;
};

/** share/src/bi/filter/Filter.hpp:26-92 */
template <class F>
class Filter : public F {
 public:
  template <class... Args>
  Filter(Args&&... args) : F(std::forward<Args>(args)...) {}

  template <class S1, class IO1>
  void filter(Random& rng, const ScheduleIterator first, const ScheduleIterator last, S1& s,
              IO1& out) {
    ScheduleIterator iter = first;
    this->output0(s, out);
    this->correct(rng, *iter, s);
    this->output(*iter, s, out);
    while (iter + 1 != last) {
      this->step(rng, iter, last, s, out);
    }
    this->term(s);
    this->outputT(s, out);
  }
};

}  // namespace bi

#endif
