// A filter client written against the C++ shims the way LibBi's generated
// client is written (share/tt/cpp/client/filter_cpu.cpp.tt:105-192):
// Random, Schedule, BootstrapPFState, Resampler, Filter<BootstrapPF<...>>.
//
// usage: shim_filter <pz|l96|lg> <resampler 0|1|2> <P> <dt> <seed> <essRel>
//                    <obs.bin> <out.bin> [mean0.bin] [fused] [nc=<file>]
// obs.bin: T x nobs doubles.  out.bin: logZ, T logIncrements, then the
// (T+1) x P ancestors (ints) output at every time.  With "fused" the run
// uses the device-resident filter instead and writes the same layout.  With
// nc=<file> the outputs also go to a NetCDF file in LibBi's ParticleFilter
// schema (ParticleFilterNetCDFBuffer; the device filter records its history).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <memory>
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

// Forwards the filter's outputs to the memory buffer and, when given, to the
// NetCDF buffer.
struct TeeBuffer {
  bi::ParticleFilterMemoryBuffer& mem;
  bi::ParticleFilterNetCDFBuffer* nc;
  template <class S1>
  void write(const int k, const double t, const S1& s) {
    mem.write(k, t, s);
    if (nc) nc->write(k, t, s);
  }
  template <class S1>
  void writeT(const S1& s) {
    mem.writeT(s);
    if (nc) nc->writeT(s);
  }
};

template <class R, class B>
static int run(B& m, int P, double dt, uint64_t seed, double essRel,
               const std::vector<double>& y, const std::vector<double>& mean0,
               const std::string& out_path, bool fused, const std::string& nc_path) {
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
    bi::ParticleFilterMemoryBuffer mem;
    std::unique_ptr<bi::ParticleFilterNetCDFBuffer> nc;
    if (!nc_path.empty()) nc.reset(new bi::ParticleFilterNetCDFBuffer(m, P, T + 1, nc_path));
    TeeBuffer out{mem, nc.get()};
    filter.init(rng, s, out, mean0);
    filter.filter(rng, sched.begin(), sched.end(), s, out);
    logZ = s.logLikelihood;
    li = s.logIncrements;
    anc = mem.ancestors;
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
    cfg.record = nc_path.empty() ? 1 : 2;
    bi::amd::DeviceBootstrapPF<B> pf(m, cfg);
    pf.setObservations(y);
    pf.init(mean0);
    pf.step(T);
    logZ = pf.term();
    if (!nc_path.empty()) pf.writeNetCDF(nc_path);
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
                    const std::string& out, bool fused, const std::string& nc) {
  if (res == 0)
    return run<bi::SystematicResampler>(m, P, dt, seed, essRel, y, mean0, out, fused, nc);
  if (res == 1)
    return run<bi::StratifiedResampler>(m, P, dt, seed, essRel, y, mean0, out, fused, nc);
  return run<bi::MultinomialResampler>(m, P, dt, seed, essRel, y, mean0, out, fused, nc);
}

int main(int argc, char** argv) {
  if (argc < 9) {
    std::fprintf(stderr,
                 "usage: %s <pz|l96|lg> <resampler> <P> <dt> <seed> <essRel> <obs.bin> "
                 "<out.bin> [mean0.bin] [fused] [nc=<file>]\n",
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
    std::string nc;
    std::vector<double> mean0;
    for (int i = 9; i < argc; ++i) {
      const std::string a = argv[i];
      if (a == "fused") fused = true;
      else if (a.compare(0, 3, "nc=") == 0) nc = a.substr(3);
      else mean0 = read_file<double>(a);
    }
    if (model == "pz") {
      bi::amd::PZ<double> m;
      return dispatch(m, res, P, dt, seed, essRel, y, mean0, out, fused, nc);
    } else if (model == "l96") {
      bi::amd::Lorenz96<40, double> m;
      return dispatch(m, res, P, dt, seed, essRel, y, mean0, out, fused, nc);
    } else {
      bi::amd::LinearGaussian<double> m;
      return dispatch(m, res, P, dt, seed, essRel, y, mean0, out, fused, nc);
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "shim_filter: %s\n", e.what());
    return 1;
  }
}
