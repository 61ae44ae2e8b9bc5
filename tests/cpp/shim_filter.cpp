// A filter client written against the C++ shims the way LibBi's generated
// client is written (share/tt/cpp/client/filter_cpu.cpp.tt:105-192):
// Random, Schedule, BootstrapPFState, Resampler, Filter<BootstrapPF<...>>.
//
// usage: shim_filter <pz|l96|lg> <resampler 0|1|2> <P> <dt> <seed> <essRel>
//                    <obs.bin> <out.bin> [mean0.bin] [fused]
// obs.bin: T x nobs doubles.  out.bin: logZ, T logIncrements, then the
// (T+1) x P ancestors (ints) output at every time.  With "fused" the run
// uses the device-resident filter instead and writes the same layout.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "bi/amd/bi.hpp"

template <class T>
static std::vector<T> read_file(const std::string& path) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) throw std::runtime_error("cannot read " + path);
  const size_t n = (size_t)f.tellg() / sizeof(T);
  std::vector<T> v(n);
  f.seekg(0);
  f.read(reinterpret_cast<char*>(v.data()), n * sizeof(T));
  return v;
}

template <class R, class B>
static int run(B& m, int P, double dt, uint64_t seed, double essRel,
               const std::vector<double>& y, const std::vector<double>& mean0,
               const std::string& out_path, bool fused) {
  const int nobs = m.getNumObs();
  const int T = (int)y.size() / nobs;
  double logZ;
  std::vector<double> li;
  std::vector<std::vector<int> > anc;
  if (!fused) {
    bi::Random rng(seed);
    bi::Schedule sched(0.0, dt, T);
    bi::NullForcer in;
    bi::Observer obs(y, nobs);
    bi::BootstrapPFState<B, bi::ON_DEVICE> s(m, P, T);
    bi::Resampler<R> resam(essRel);
    bi::Filter<bi::BootstrapPF<B, bi::NullForcer, bi::Observer, bi::Resampler<R> > > filter(
        m, in, obs, resam);
    bi::ParticleFilterMemoryBuffer out;
    filter.init(rng, s, out, mean0);
    filter.filter(rng, sched.begin(), sched.end(), s, out);
    logZ = s.logLikelihood;
    li = s.logIncrements;
    anc = out.ancestors;
  } else {
    bihip_pf_config cfg{};
    cfg.P = P;
    cfg.T = T;
    cfg.resampler = R::TAG == BIHIP_TAG_SYSTEMATIC   ? BIHIP_SYSTEMATIC
                    : R::TAG == BIHIP_TAG_STRATIFIED ? BIHIP_STRATIFIED
                                                     : BIHIP_MULTINOMIAL;
    cfg.essRel = essRel;
    cfg.t0 = 0.0;
    cfg.dt = dt;
    cfg.seed = seed;
    cfg.record = 1;
    bi::amd::DeviceBootstrapPF<B> pf(m, cfg);
    pf.setObservations(y);
    pf.init(mean0);
    pf.step(T);
    logZ = pf.term();
    li = pf.logIncrements();
    anc.resize(T + 1);
    anc[0].resize(P);
    for (int i = 0; i < P; ++i) anc[0][i] = i;
    for (int k = 0; k < T; ++k) anc[k + 1] = pf.ancestors(k);
  }
  std::ofstream f(out_path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(&logZ), sizeof(double));
  f.write(reinterpret_cast<const char*>(li.data()), sizeof(double) * li.size());
  for (auto& a : anc) f.write(reinterpret_cast<const char*>(a.data()), sizeof(int) * a.size());
  std::printf("logZ %.17g\n", logZ);
  return 0;
}

template <class B>
static int dispatch(B& m, int res, int P, double dt, uint64_t seed, double essRel,
                    const std::vector<double>& y, const std::vector<double>& mean0,
                    const std::string& out, bool fused) {
  if (res == 0) return run<bi::SystematicResampler>(m, P, dt, seed, essRel, y, mean0, out, fused);
  if (res == 1) return run<bi::StratifiedResampler>(m, P, dt, seed, essRel, y, mean0, out, fused);
  return run<bi::MultinomialResampler>(m, P, dt, seed, essRel, y, mean0, out, fused);
}

int main(int argc, char** argv) {
  if (argc < 9) {
    std::fprintf(stderr,
                 "usage: %s <pz|l96|lg> <resampler> <P> <dt> <seed> <essRel> <obs.bin> "
                 "<out.bin> [mean0.bin] [fused]\n",
                 argv[0]);
    return 2;
  }
  try {
    const std::string model = argv[1];
    const int res = std::atoi(argv[2]), P = std::atoi(argv[3]);
    const double dt = std::atof(argv[4]);
    const uint64_t seed = std::strtoull(argv[5], nullptr, 10);
    const double essRel = std::atof(argv[6]);
    const std::vector<double> y = read_file<double>(argv[7]);
    const std::string out = argv[8];
    bool fused = false;
    std::vector<double> mean0;
    for (int i = 9; i < argc; ++i) {
      if (std::string(argv[i]) == "fused") fused = true;
      else mean0 = read_file<double>(argv[i]);
    }
    if (model == "pz") {
      bi::amd::PZ<double> m;
      return dispatch(m, res, P, dt, seed, essRel, y, mean0, out, fused);
    } else if (model == "l96") {
      bi::amd::Lorenz96<40, double> m;
      return dispatch(m, res, P, dt, seed, essRel, y, mean0, out, fused);
    } else {
      bi::amd::LinearGaussian<double> m;
      return dispatch(m, res, P, dt, seed, essRel, y, mean0, out, fused);
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "shim_filter: %s\n", e.what());
    return 1;
  }
}
