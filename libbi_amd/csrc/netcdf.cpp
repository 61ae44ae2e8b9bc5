// Particle-filter output in NetCDF, the layout of LibBi's DEFAULT schema:
//   ParticleFilterNetCDFBuffer::create  share/src/bi/netcdf/ParticleFilterNetCDFBuffer.cpp:19-38
//   SimulatorNetCDFBuffer::create       share/src/bi/netcdf/SimulatorNetCDFBuffer.cpp:23-83
//   SimulatorNetCDFBuffer::createVar    share/src/bi/netcdf/SimulatorNetCDFBuffer.cpp:167-215
//   ParticleFilterBuffer::write / writeT share/src/bi/buffer/ParticleFilterBuffer.hpp:69-82
// Dimensions nr (output times), the model's dimensions, np (particles);
// variables time(nr), each noise variable and each state variable (nr, [model
// dims reversed], np) -- the R_VAR and D_VAR types that create() writes --,
// ancestor(nr, np) int, logweight(nr, np), loglikelihood(), clock(); global
// attributes libbi_schema = "ParticleFilter", libbi_schema_version = 1,
// libbi_version.
//
// The reference writes NetCDF-4 through libnetcdf.  This writer has no
// dependency: it writes the classic format with 64-bit offsets (CDF-2), or,
// when a variable other than the last exceeds 4 GiB (the log-weights and
// noise of long runs of large filters), the 64-bit data format (CDF-5: every
// count, length, dimension id and size 64-bit); NetCDF libraries read both
// through the same nc_open / nc_get_var calls.  `clock` is NC_DOUBLE
// (microseconds; CDF-2 has no 64-bit integer).
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include <fcntl.h>
#include <unistd.h>

#include "internal.hpp"

namespace {

constexpr int NC_INT = 4, NC_FLOAT = 5, NC_DOUBLE = 6, NC_CHAR = 2;
constexpr uint32_t TAG_DIMENSION = 0x0A, TAG_VARIABLE = 0x0B, TAG_ATTRIBUTE = 0x0C;

int type_size(int t) { return t == NC_DOUBLE ? 8 : (t == NC_CHAR ? 1 : 4); }

struct Var {
  std::string name;
  std::vector<int> dims;  // dimension ids, outermost first
  int type = NC_DOUBLE;
  uint64_t bytes = 0;     // unpadded data size
  uint64_t begin = 0;     // file offset of the data
};

struct Bytes {
  std::vector<uint8_t> b;
  int version = 2;  // CDF-2 or CDF-5
  void u32(uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
  }
  void u64(uint64_t v) {
    for (int s = 56; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
  }
  void pad() {
    while (b.size() % 4) b.push_back(0);
  }
  // NON_NEG of the grammar: counts, lengths, dimension ids, sizes
  void nn(uint64_t v) {
    if (version == 5) u64(v);
    else u32((uint32_t)v);
  }
  void name(const std::string& n) {
    nn(n.size());
    b.insert(b.end(), n.begin(), n.end());
    pad();
  }
};

template <class T>
void put_be(uint8_t* dst, T v) {
  uint8_t raw[sizeof(T)];
  std::memcpy(raw, &v, sizeof(T));
  for (size_t i = 0; i < sizeof(T); ++i) dst[i] = raw[sizeof(T) - 1 - i];
}

}  // namespace

struct bihip_pfout {
  int fd = -1;
  std::string path;
  int nr = 0, np = 0, d = 0;
  int state_type = NC_DOUBLE;
  std::vector<std::pair<std::string, uint64_t>> dims;
  std::vector<Var> vars;
  int v_time = -1, v_clock = -1, v_ll = -1, v_anc = -1, v_lw = -1;
  std::vector<int> v_state;           // one per state variable
  std::vector<uint64_t> state_width;  // components per state variable and particle
  std::vector<int> state_first;       // first state component of each variable
  std::vector<int> v_noise;           // the same for the noise variables
  std::vector<uint64_t> noise_width;
  std::vector<int> noise_first;
  std::vector<uint8_t> scratch;
};

namespace {

using bihip::set_error;

int write_at(bihip_pfout* o, uint64_t off, const void* p, size_t n) {
  const uint8_t* c = static_cast<const uint8_t*>(p);
  while (n > 0) {
    const ssize_t w = pwrite(o->fd, c, n, (off_t)off);
    if (w <= 0) return set_error(BIHIP_EIO, "write to %s failed", o->path.c_str());
    c += w;
    off += (uint64_t)w;
    n -= (size_t)w;
  }
  return BIHIP_OK;
}

// The model's NetCDF variables: (name, dimension name or "" for a scalar,
// components), in state order.
struct StateVar {
  const char* name;
  const char* dim;
  int width;
};

std::vector<StateVar> state_vars(const bihip_model* m) {
  if (const bihip::GenInfo* g = bihip::generated_info(m->model)) {  // the model file's state variables
    std::vector<StateVar> v;
    for (int i = 0; g->names[BIHIP_VAR_STATE][i] != nullptr; ++i)
      v.push_back({g->names[BIHIP_VAR_STATE][i], g->state_dims[i], g->sizes[BIHIP_VAR_STATE][i]});
    return v;
  }
  switch (m->model) {
    case BIHIP_MODEL_PZ:  // docs/src/users.tex:36-85: state P, Z
      return {{"P", "", 1}, {"Z", "", 1}};
    case BIHIP_MODEL_L96:
      return {{"x", "n", m->d}};
    default:
      return {{"x", "", 1}};
  }
}

// The noise variables, in the order of the recorded noise rows: a generated
// model's from its file, a built-in model's as the model files declare them
// (PZ.bi's alpha, Lorenz96.bi's w[n], LinearGaussian.bi's w).
std::vector<StateVar> noise_vars(const bihip_model* m) {
  if (const bihip::GenInfo* g = bihip::generated_info(m->model)) {
    std::vector<StateVar> v;
    for (int i = 0; g->names[BIHIP_VAR_NOISE][i] != nullptr; ++i)
      v.push_back({g->names[BIHIP_VAR_NOISE][i], g->noise_dims[i], g->sizes[BIHIP_VAR_NOISE][i]});
    return v;
  }
  switch (m->model) {
    case BIHIP_MODEL_PZ:
      return {{"alpha", "", 1}};
    case BIHIP_MODEL_L96:
      return {{"w", "n", m->d}};
    default:
      return {{"w", "", 1}};
  }
}

// Variables v (each w components of ld-strided rows from `first` on) of
// output k, big-endian.
int write_rows(bihip_pfout* o, int k, const void* X_host, int ld, const std::vector<int>& vars,
               const std::vector<uint64_t>& width, const std::vector<int>& first) {
  const int esz = type_size(o->state_type);
  const uint64_t np = (uint64_t)o->np;
  for (size_t v = 0; v < vars.size(); ++v) {
    const uint64_t w = width[v];
    o->scratch.resize(w * np * esz);
    for (uint64_t c = 0; c < w; ++c) {
      const uint64_t row = (uint64_t)first[v] + c;
      uint8_t* dst = o->scratch.data() + c * np * esz;
      if (esz == 4) {
        const float* src = static_cast<const float*>(X_host) + row * (uint64_t)ld;
        for (uint64_t p = 0; p < np; ++p) put_be<float>(dst + 4 * p, src[p]);
      } else {
        const double* src = static_cast<const double*>(X_host) + row * (uint64_t)ld;
        for (uint64_t p = 0; p < np; ++p) put_be<double>(dst + 8 * p, src[p]);
      }
    }
    const Var& var = o->vars[vars[v]];
    int rc = write_at(o, var.begin + (uint64_t)k * w * np * esz, o->scratch.data(),
                      o->scratch.size());
    if (rc) return rc;
  }
  return BIHIP_OK;
}

}  // namespace

extern "C" {

int bihip_pfout_create(const char* path, const bihip_model* m, int nr, int np,
                       bihip_pfout** out) {
  return bihip_pfout_create_format(path, m, nr, np, 0, out);
}

int bihip_pfout_create_format(const char* path, const bihip_model* m, int nr, int np,
                              int format, bihip_pfout** out) {
  if (path == nullptr || m == nullptr || out == nullptr)
    return set_error(BIHIP_EINVAL, "NULL argument");
  if (nr <= 0 || np <= 0) return set_error(BIHIP_EINVAL, "need nr > 0 and np > 0");
  if (format != 0 && format != 2 && format != 5)
    return set_error(BIHIP_EINVAL, "format is 0 (the smallest that fits), 2 or 5");
  auto* o = new bihip_pfout();
  o->path = path;
  o->nr = nr;
  o->np = np;
  o->d = m->d;
  o->state_type = m->precision == BIHIP_F32 ? NC_FLOAT : NC_DOUBLE;
  const std::vector<StateVar> sv = state_vars(m);
  const std::vector<StateVar> nv = noise_vars(m);
  std::vector<StateVar> all = sv;
  all.insert(all.end(), nv.begin(), nv.end());

  // dimensions: nr, the model's, np
  o->dims.push_back({"nr", (uint64_t)nr});
  for (const StateVar& s : all) {  // each model dimension once (variables may share one)
    if (s.dim[0] == '\0') continue;
    bool seen = false;
    for (const auto& d : o->dims) seen = seen || d.first == s.dim;
    if (!seen) o->dims.push_back({s.dim, (uint64_t)s.width});
  }
  o->dims.push_back({"np", (uint64_t)np});
  const int nrDim = 0, npDim = (int)o->dims.size() - 1;
  auto dim_id = [&](const char* n) {
    for (size_t i = 0; i < o->dims.size(); ++i)
      if (o->dims[i].first == n) return (int)i;
    return -1;
  };

  auto add = [&](const char* name, std::vector<int> dims, int type) {
    Var v;
    v.name = name;
    v.dims = std::move(dims);
    v.type = type;
    uint64_t n = 1;
    for (int id : v.dims) n *= o->dims[id].second;
    v.bytes = n * (uint64_t)type_size(type);
    o->vars.push_back(v);
    return (int)o->vars.size() - 1;
  };
  o->v_time = add("time", {nrDim}, NC_DOUBLE);
  o->v_clock = add("clock", {}, NC_DOUBLE);
  o->v_ll = add("loglikelihood", {}, NC_DOUBLE);
  o->v_anc = add("ancestor", {nrDim, npDim}, NC_INT);
  o->v_lw = add("logweight", {nrDim, npDim}, NC_DOUBLE);
  int nfirst = 0;
  for (const StateVar& s : nv) {
    std::vector<int> dims{nrDim};
    if (s.dim[0] != '\0') dims.push_back(dim_id(s.dim));
    dims.push_back(npDim);
    o->v_noise.push_back(add(s.name, dims, o->state_type));
    o->noise_width.push_back((uint64_t)s.width);
    o->noise_first.push_back(nfirst);
    nfirst += s.width;
  }
  int first = 0;
  for (const StateVar& s : sv) {
    std::vector<int> dims{nrDim};
    if (s.dim[0] != '\0') dims.push_back(dim_id(s.dim));
    dims.push_back(npDim);
    o->v_state.push_back(add(s.name, dims, o->state_type));
    o->state_width.push_back((uint64_t)s.width);
    o->state_first.push_back(first);
    first += s.width;
  }
  if (first != m->d) {
    delete o;
    return set_error(BIHIP_EINVAL, "model dimension %d does not match its variables", m->d);
  }

  // header (the classic format's grammar with 64-bit begin offsets; CDF-5
  // when a variable before the last does not fit CDF-2's 32-bit vsize)
  Bytes h;
  bool big = false;
  for (size_t i = 0; i + 1 < o->vars.size(); ++i)
    big = big || (o->vars[i].bytes + 3) / 4 * 4 > 0xFFFFFFFCull;
  if (big && format == 2) {
    delete o;
    return set_error(BIHIP_EINVAL, "a variable before the last exceeds CDF-2's 4 GiB limit");
  }
  h.version = format == 5 || (format == 0 && big) ? 5 : 2;
  const char magic[4] = {'C', 'D', 'F', (char)h.version};
  h.b.insert(h.b.end(), magic, magic + 4);
  h.nn(0);  // numrecs: no record dimension
  h.u32(TAG_DIMENSION);
  h.nn(o->dims.size());
  for (auto& dm : o->dims) {
    h.name(dm.first);
    h.nn(dm.second);
  }
  const std::string schema = "ParticleFilter", version = "1.4.5";
  h.u32(TAG_ATTRIBUTE);
  h.nn(3);
  h.name("libbi_schema");
  h.u32(NC_CHAR);
  h.nn(schema.size());
  h.b.insert(h.b.end(), schema.begin(), schema.end());
  h.pad();
  h.name("libbi_schema_version");
  h.u32(NC_INT);
  h.nn(1);
  h.u32(1);
  h.name("libbi_version");
  h.u32(NC_CHAR);
  h.nn(version.size());
  h.b.insert(h.b.end(), version.begin(), version.end());
  h.pad();
  // variables: the header size does not depend on the offsets, so lay it
  // out once with zero offsets, then place the data after it
  std::vector<size_t> begin_pos;
  h.u32(TAG_VARIABLE);
  h.nn(o->vars.size());
  for (size_t i = 0; i < o->vars.size(); ++i) {
    const Var& v = o->vars[i];
    h.name(v.name);
    h.nn(v.dims.size());
    for (int id : v.dims) h.nn((uint64_t)id);
    h.u32(0);  // vatt_list ABSENT: a zero tag and a zero count
    h.nn(0);
    h.u32((uint32_t)v.type);
    const uint64_t padded = (v.bytes + 3) / 4 * 4;
    if (h.version == 5) h.u64(padded);
    else h.u32(padded > 0xFFFFFFFCull ? 0xFFFFFFFFu : (uint32_t)padded);  // only the last
    begin_pos.push_back(h.b.size());
    h.u64(0);
  }
  uint64_t off = (h.b.size() + 3) / 4 * 4;
  for (size_t i = 0; i < o->vars.size(); ++i) {
    o->vars[i].begin = off;
    put_be<uint64_t>(&h.b[begin_pos[i]], off);
    off += (o->vars[i].bytes + 3) / 4 * 4;
  }
  h.pad();

  o->fd = open(path, O_CREAT | O_TRUNC | O_WRONLY, 0644);
  if (o->fd < 0) {
    delete o;
    return set_error(BIHIP_EIO, "cannot create %s", path);
  }
  int rc = write_at(o, 0, h.b.data(), h.b.size());
  if (rc == BIHIP_OK && ftruncate(o->fd, (off_t)off) != 0)
    rc = set_error(BIHIP_EIO, "cannot size %s", path);
  if (rc) {
    close(o->fd);
    delete o;
    return rc;
  }
  *out = o;
  return BIHIP_OK;
}

// SimulatorNetCDFBuffer::writeTime
int bihip_pfout_write_time(bihip_pfout* o, int k, double t) {
  if (o == nullptr || k < 0 || k >= o->nr) return set_error(BIHIP_EINVAL, "bad output index");
  uint8_t b[8];
  put_be<double>(b, t);
  return write_at(o, o->vars[o->v_time].begin + 8ull * (uint64_t)k, b, 8);
}

// ParticleFilterNetCDFBuffer::writeState(k, X, as): the state at output k
// (column-major d x ld, FP32 or FP64 by the model's precision) and the
// ancestors that produced it.
int bihip_pfout_write_state(bihip_pfout* o, int k, const void* X_host, int ld,
                            const int* as_host) {
  if (o == nullptr || k < 0 || k >= o->nr) return set_error(BIHIP_EINVAL, "bad output index");
  if (X_host == nullptr || as_host == nullptr) return set_error(BIHIP_EINVAL, "NULL argument");
  if (ld < o->np) return set_error(BIHIP_EINVAL, "ld < np");
  const uint64_t np = (uint64_t)o->np;
  int rc = write_rows(o, k, X_host, ld, o->v_state, o->state_width, o->state_first);
  if (rc) return rc;
  o->scratch.resize(np * 4);
  for (uint64_t p = 0; p < np; ++p) put_be<int32_t>(o->scratch.data() + 4 * p, as_host[p]);
  return write_at(o, o->vars[o->v_anc].begin + (uint64_t)k * np * 4, o->scratch.data(), np * 4);
}

// SimulatorNetCDFBuffer::writeState's noise part (R_VAR, SimulatorNetCDFBuffer.hpp:341-352):
// the noise variables drawn in the step to output k, in the state's layout.
int bihip_pfout_write_noise(bihip_pfout* o, int k, const void* Z_host, int ld) {
  if (o == nullptr || k < 0 || k >= o->nr) return set_error(BIHIP_EINVAL, "bad output index");
  if (Z_host == nullptr) return set_error(BIHIP_EINVAL, "NULL argument");
  if (o->v_noise.empty()) return set_error(BIHIP_EINVAL, "the model has no recorded noise variables");
  if (ld < o->np) return set_error(BIHIP_EINVAL, "ld < np");
  return write_rows(o, k, Z_host, ld, o->v_noise, o->noise_width, o->noise_first);
}

// ParticleFilterNetCDFBuffer::writeLogWeights
int bihip_pfout_write_log_weights(bihip_pfout* o, int k, const double* lws_host) {
  if (o == nullptr || k < 0 || k >= o->nr) return set_error(BIHIP_EINVAL, "bad output index");
  if (lws_host == nullptr) return set_error(BIHIP_EINVAL, "NULL argument");
  const uint64_t np = (uint64_t)o->np;
  o->scratch.resize(np * 8);
  for (uint64_t p = 0; p < np; ++p) put_be<double>(o->scratch.data() + 8 * p, lws_host[p]);
  return write_at(o, o->vars[o->v_lw].begin + (uint64_t)k * np * 8, o->scratch.data(), np * 8);
}

// ParticleFilterNetCDFBuffer::writeLogLikelihood
int bihip_pfout_write_log_likelihood(bihip_pfout* o, double ll) {
  if (o == nullptr) return set_error(BIHIP_EINVAL, "NULL output");
  uint8_t b[8];
  put_be<double>(b, ll);
  return write_at(o, o->vars[o->v_ll].begin, b, 8);
}

// SimulatorNetCDFBuffer::writeClock (microseconds)
int bihip_pfout_write_clock(bihip_pfout* o, double clock_us) {
  if (o == nullptr) return set_error(BIHIP_EINVAL, "NULL output");
  uint8_t b[8];
  put_be<double>(b, clock_us);
  return write_at(o, o->vars[o->v_clock].begin, b, 8);
}

int bihip_pfout_close(bihip_pfout* o) {
  if (o == nullptr) return BIHIP_OK;
  int rc = BIHIP_OK;
  if (o->fd >= 0 && close(o->fd) != 0) rc = set_error(BIHIP_EIO, "close of %s failed", o->path.c_str());
  delete o;
  return rc;
}

}  // extern "C"
