/**
 * @file
 *
 * Umbrella header of the C++ shims: include this where the reference
 * includes bi/filter/Filter.hpp, bi/filter/BootstrapPF.hpp,
 * the bi/resampler headers and bi/state/BootstrapPFState.hpp, then link
 * libbihip.so.
 */
#ifndef BI_AMD_BI_HPP
#define BI_AMD_BI_HPP

#include "core.hpp"
#include "filter.hpp"

namespace bi {
namespace amd {

/**
 * The whole filter resident on the device (bihip_pf_*): the resampling
 * decision, ESS, logLikelihood and logIncrements stay on the GPU, so a step
 * never waits for the host.  Same results as Filter<BootstrapPF<...>>, which
 * synchronises at each resampling decision as the reference does.
 */
template <class B>
class DeviceBootstrapPF {
 public:
  DeviceBootstrapPF(const B& m, const bihip_pf_config& cfg) : pf(nullptr), P(cfg.P), T(cfg.T) {
    check(bihip_pf_create(ctx(), &m.description(), &cfg, &pf));
  }
  ~DeviceBootstrapPF() { bihip_pf_destroy(pf); }
  DeviceBootstrapPF(const DeviceBootstrapPF&) = delete;
  DeviceBootstrapPF& operator=(const DeviceBootstrapPF&) = delete;

  void setObservations(const std::vector<double>& y) {
    check(bihip_pf_set_observations(pf, y.data()));
  }
  void init(const std::vector<double>& mean0 = {}) {
    check(bihip_pf_init(pf, mean0.empty() ? nullptr : mean0.data()));
  }
  /** BootstrapPF::step for the next n observation intervals (asynchronous). */
  void step(int n = 1) { check(bihip_pf_step(pf, n)); }
  /** BootstrapPF::term: logsumexp(lws) - log P. */
  double term() {
    double z = 0.0;
    check(bihip_pf_term(pf, &z));
    return z;
  }
  std::vector<double> logIncrements() {
    std::vector<double> li(T);
    check(bihip_pf_outputs(pf, li.data(), nullptr, nullptr, nullptr));
    return li;
  }
  /** Ancestors of step k (record = 1), or of the last resampling for k < 0. */
  std::vector<int> ancestors(int k) {
    std::vector<int> a(P);
    check(bihip_pf_ancestors(pf, k, a.data()));
    return a;
  }
  /** The recorded history (record = 2) in LibBi's ParticleFilter NetCDF schema. */
  void writeNetCDF(const std::string& file) { check(bihip_pf_write_netcdf(pf, file.c_str())); }
  bihip_pf* handle() { return pf; }

 private:
  bihip_pf* pf;
  int P, T;
};

}  // namespace amd
}  // namespace bi

#endif
