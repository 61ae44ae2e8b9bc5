// The ode block of a generated LibBi model, written against the C++ shims the
// way share/tt/cpp/block/ode.hpp.tt:35-52 writes it: the block sets the
// integrator constants with bi_ode_set and calls the static
// Integrator<Model, action_typelist>::update(t1, t2, s).
//
// usage: shim_ode <rk4|dopri5> <pz|l96> <P> <t1> <t2> <h> <atoler> <rtoler> <in.bin> <out.bin>
// in.bin: d x P doubles (variable-major, as the state's columns); out.bin the same
// after the flow.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "bi/amd/bi.hpp"

// the action type list of the generated block (unused by the device path)
struct action_typelist {};

template <int ALG, class Model>
struct OdeBlock {
  // share/tt/cpp/block/ode.hpp.tt:35-52
  template <class T1>
  static void simulates(const T1 t1, const T1 t2, const double H, const double ATOLER,
                        const double RTOLER, bi::State<Model, bi::ON_DEVICE>& s) {
    bi_ode_set(H, ATOLER, RTOLER);
    if (ALG == BIHIP_RK4)
      bi::RK4Integrator<Model, action_typelist>::update(t1, t2, s);
    else
      bi::DOPRI5Integrator<Model, action_typelist>::update(t1, t2, s);
  }
};

template <int ALG, class Model>
static int run(const Model& m, int P, double t1, double t2, double h, double atol, double rtol,
               const std::string& in, const std::string& out) {
  std::ifstream f(in, std::ios::binary);
  const int d = m.getDynSize();
  std::vector<double> x((size_t)d * P);
  f.read(reinterpret_cast<char*>(x.data()), x.size() * sizeof(double));
  if (!f) {
    std::fprintf(stderr, "cannot read %s\n", in.c_str());
    return 1;
  }
  typedef typename Model::real_type real;
  bi::State<Model, bi::ON_DEVICE> s(m, P);
  std::vector<real> cols((size_t)s.lead() * d, real(0));
  for (int i = 0; i < d; ++i)
    for (int p = 0; p < P; ++p) cols[(size_t)i * s.lead() + p] = (real)x[(size_t)i * P + p];
  bi::amd::check(bihip_memcpy_to_device(bi::amd::ctx(), s.buf(), cols.data(),
                                        cols.size() * sizeof(real)));
  OdeBlock<ALG, Model>::simulates(t1, t2, h, atol, rtol, s);
  cols = s.to_host();
  for (int i = 0; i < d; ++i)
    for (int p = 0; p < P; ++p) x[(size_t)i * P + p] = (double)cols[(size_t)i * s.lead() + p];
  std::ofstream o(out, std::ios::binary);
  o.write(reinterpret_cast<const char*>(x.data()), x.size() * sizeof(double));
  return o ? 0 : 1;
}

int main(int argc, char** argv) {
  if (argc != 11) {
    std::fprintf(stderr,
                 "usage: shim_ode <rk4|dopri5> <pz|l96> <P> <t1> <t2> <h> <atoler> <rtoler> "
                 "<in.bin> <out.bin>\n");
    return 2;
  }
  const std::string alg = argv[1], model = argv[2];
  const int P = std::atoi(argv[3]);
  const double t1 = std::atof(argv[4]), t2 = std::atof(argv[5]), h = std::atof(argv[6]);
  const double atol = std::atof(argv[7]), rtol = std::atof(argv[8]);
  try {
    if (model == "pz") {
      bi::amd::PZ<double> m;
      return alg == "rk4" ? run<BIHIP_RK4>(m, P, t1, t2, h, atol, rtol, argv[9], argv[10])
                          : run<BIHIP_DOPRI5>(m, P, t1, t2, h, atol, rtol, argv[9], argv[10]);
    }
    bi::amd::Lorenz96<40, double> m;
    return alg == "rk4" ? run<BIHIP_RK4>(m, P, t1, t2, h, atol, rtol, argv[9], argv[10])
                        : run<BIHIP_DOPRI5>(m, P, t1, t2, h, atol, rtol, argv[9], argv[10]);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
