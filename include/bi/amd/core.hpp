/**
 * @file
 *
 * Core of the C++ shims over the bihip C ABI: device context, error
 * mapping, device containers, Random and Schedule.
 *
 * These headers keep the class names and call signatures of LibBi's
 * ON_DEVICE path so that the reference's call sites (the generated
 * filter client, MarginalSIR, MarginalMH) compile against them.  They are
 * plain C++17 and reach the GPU only through include/bihip.h; compile with
 * any host compiler and link libbihip.so.
 */
#ifndef BI_AMD_CORE_HPP
#define BI_AMD_CORE_HPP

#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../bihip.h"

namespace bi {

/** Memory location of a container (share/src/bi/misc/location.hpp). */
enum Location { ON_HOST = 0, ON_DEVICE = 1 };

/** share/src/bi/misc/exception.hpp:47-49 */
struct ParticleFilterDegeneratedException : public std::runtime_error {
  ParticleFilterDegeneratedException()
      : std::runtime_error("particle filter degenerated (sum of weights <= 0)") {}
};

/** A bihip call failed for a reason other than degeneracy. */
struct DeviceError : public std::runtime_error {
  int code;
  DeviceError(int code, const std::string& msg) : std::runtime_error(msg), code(code) {}
};

namespace amd {

/** Throws the exception that matches a bihip status code. */
inline void check(int rc) {
  if (rc == BIHIP_OK) return;
  if (rc == BIHIP_DEGENERATE) throw ParticleFilterDegeneratedException();
  throw DeviceError(rc, std::string("bihip: ") + bihip_last_error());
}

/**
 * The device context: one per host thread and device, as the reference
 * keeps one CUDA context per OpenMP thread (share/src/bi/misc/omp.hpp:31-42).
 */
class Context {
 public:
  static Context& get(int device = 0) {
    thread_local Context c(device);
    return c;
  }
  bihip_ctx* handle() const { return ctx; }
  void synchronize() { check(bihip_ctx_synchronize(ctx)); }
  ~Context() { bihip_ctx_destroy(ctx); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;

 private:
  explicit Context(int device) : ctx(nullptr) { check(bihip_ctx_create(device, &ctx)); }
  bihip_ctx* ctx;
};

inline bihip_ctx* ctx() { return Context::get().handle(); }

}  // namespace amd

/**
 * Owning device vector (the role of bi::gpu_vector on the ON_DEVICE path).
 */
template <class T>
class device_vector {
 public:
  typedef T value_type;
  static const bool on_device = true;

  explicit device_vector(int n = 0) : n(0), p(nullptr) { resize(n); }
  device_vector(const device_vector& o) : n(0), p(nullptr) {
    resize(o.n);
    if (n) amd::check(bihip_memcpy_device(amd::ctx(), p, o.p, bytes()));
  }
  device_vector& operator=(const device_vector& o) {
    if (this != &o) {
      resize(o.n);
      if (n) amd::check(bihip_memcpy_device(amd::ctx(), p, o.p, bytes()));
    }
    return *this;
  }
  ~device_vector() {
    if (p) bihip_free(amd::ctx(), p);
  }
  void resize(int m) {
    if (m == n) return;
    if (p) amd::check(bihip_free(amd::ctx(), p));
    p = nullptr;
    n = m;
    if (n > 0) {
      void* q = nullptr;
      amd::check(bihip_malloc(amd::ctx(), bytes(), &q));
      p = static_cast<T*>(q);
    }
  }
  void swap(device_vector& o) {
    std::swap(n, o.n);
    std::swap(p, o.p);
  }
  int size() const { return n; }
  T* buf() { return p; }
  const T* buf() const { return p; }
  size_t bytes() const { return sizeof(T) * (size_t)n; }

  /** Copy to / from host memory (synchronous). */
  std::vector<T> to_host() const {
    std::vector<T> h(n);
    if (n) amd::check(bihip_memcpy_to_host(amd::ctx(), h.data(), p, bytes()));
    return h;
  }
  void from_host(const std::vector<T>& h) {
    resize((int)h.size());
    if (n) amd::check(bihip_memcpy_to_device(amd::ctx(), p, h.data(), bytes()));
  }

 private:
  int n;
  T* p;
};

/**
 * Random number generator (share/src/bi/random/Random.hpp:44-380).  The
 * reference keeps one boost::mt19937 per OpenMP thread; here every draw is
 * a pure function of (seed, stream, index) through Philox4x32-10, so host and
 * device draws are reproducible and independent of thread counts.  The
 * filter selects the stream (schedule index and purpose tag) before each
 * resampling; uniform() then draws index 0, 1, 2, ... of that stream.
 */
class Random {
 public:
  explicit Random(uint64_t seed = 0, uint32_t batch = 0)
      : s(seed), b(batch), step(0), tag(0), index(0) {}
  void seeds(uint64_t seed) { s = seed; }
  uint64_t seed() const { return s; }
  uint32_t batch() const { return b; }
  void setBatch(uint32_t batch) { b = batch; }
  /** Select the stream of the following host draws. */
  void setStream(uint32_t step_, uint32_t tag_) {
    step = step_;
    tag = tag_;
    index = 0;
  }
  uint32_t currentStep() const { return step; }
  /** Uniform on [lo, hi) (RngHost::uniform). */
  template <class T1>
  T1 uniform(const T1 lo = 0.0, const T1 hi = 1.0) {
    const double u = bihip_uniform_host(s, index++, step, tag, b);
    return (T1)(lo + (hi - lo) * u);
  }

 private:
  uint64_t s;
  uint32_t b, step, tag, index;
};

/**
 * One time of the schedule (share/src/bi/state/ScheduleElement.hpp:17-170),
 * restricted to what the filter step reads: t_k = t0 + k dt, observed for
 * k >= 1, output at every time, no bridges.
 */
class ScheduleElement {
 public:
  ScheduleElement(int k = 0, double t0 = 0.0, double dt = 1.0) : k(k), t0(t0), dt(dt) {}
  double getTime() const { return t0 + k * dt; }
  double getFrom() const { return t0 + (k > 0 ? k - 1 : 0) * dt; }
  double getTo() const { return getTime(); }
  bool isObserved() const { return k >= 1; }
  int indexObs() const { return k - 1; }
  bool hasBridge() const { return false; }
  bool hasOutput() const { return true; }
  int indexOutput() const { return k; }
  bool hasDelta() const { return k >= 1; }
  int index() const { return k; }

 private:
  int k;
  double t0, dt;
};

/** Random-access iterator over ScheduleElement values. */
class ScheduleIterator {
 public:
  ScheduleIterator(int k = 0, double t0 = 0.0, double dt = 1.0) : k(k), t0(t0), dt(dt) {}
  ScheduleElement operator*() const { return ScheduleElement(k, t0, dt); }
  struct Arrow {
    ScheduleElement e;
    const ScheduleElement* operator->() const { return &e; }
  };
  Arrow operator->() const { return Arrow{**this}; }
  ScheduleIterator& operator++() {
    ++k;
    return *this;
  }
  ScheduleIterator operator+(int n) const { return ScheduleIterator(k + n, t0, dt); }
  bool operator==(const ScheduleIterator& o) const { return k == o.k; }
  bool operator!=(const ScheduleIterator& o) const { return k != o.k; }

 private:
  int k;
  double t0, dt;
};

/** Schedule of T observations at t0 + k dt (share/src/bi/state/Schedule.hpp:214-337). */
class Schedule {
 public:
  Schedule(double t0, double dt, int T) : t0(t0), dt(dt), T(T) {}
  ScheduleIterator begin() const { return ScheduleIterator(0, t0, dt); }
  ScheduleIterator end() const { return ScheduleIterator(T + 1, t0, dt); }
  int numObs() const { return T; }
  int numOutputs() const { return T + 1; }

 private:
  double t0, dt;
  int T;
};

}  // namespace bi

#endif
