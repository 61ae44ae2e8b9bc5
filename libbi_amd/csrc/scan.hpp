// Single-pass tile scan with decoupled look-back over tile aggregates
// (after Merrill & Garland, "Single-pass parallel prefix scan with decoupled
// look-back", 2016).  Each tile publishes only its aggregate; a tile's
// prefix is the sum of all earlier aggregates, reduced by the whole block in
// a fixed order.  Tiles never wait on each other's prefixes.  Values are
// integers (counts, or weights in 128-bit fixed point), so every association
// order gives the same exact, monotone result.
//
// Tiles are taken from an atomic ticket, so every tile waited on has already
// started; a spin limit turns a lost tile into an error flag instead of a
// hang.  Flags carry a per-launch epoch kept on the device, so no memset is
// needed between launches (and a captured graph can replay the scan).
#pragma once

#include "blockreduce.hpp"
#include "internal.hpp"

namespace bihip {

constexpr uint32_t SCAN_SPIN_LIMIT = 1u << 22;

// Unsigned 128-bit integer for exact fixed-point prefix sums.
struct u128 {
  uint64_t lo, hi;
  u128() = default;
  __device__ __forceinline__ explicit u128(uint64_t v) : lo(v), hi(0) {}
  __device__ __forceinline__ u128(uint64_t l, uint64_t h) : lo(l), hi(h) {}
  __device__ __forceinline__ u128 operator+(const u128& o) const {
    const uint64_t l = lo + o.lo;
    return u128(l, hi + o.hi + (l < lo ? 1u : 0u));
  }
};

// Values travel between tiles as two 64-bit words.
template <class V>
struct Words {
  __device__ static void pack(const V& v, uint64_t& a, uint64_t& b) {
    a = (uint64_t)v;
    b = 0;
  }
  __device__ static V unpack(uint64_t a, uint64_t) { return (V)a; }
};
template <>
struct Words<u128> {
  __device__ static void pack(const u128& v, uint64_t& a, uint64_t& b) {
    a = v.lo;
    b = v.hi;
  }
  __device__ static u128 unpack(uint64_t a, uint64_t b) { return u128(a, b); }
};

template <class V>
__device__ __forceinline__ V shfl_up_v(const V& v, int o) {
  return __shfl_up(v, o, WAVE);
}
template <>
__device__ __forceinline__ u128 shfl_up_v<u128>(const u128& v, int o) {
  return u128(__shfl_up(v.lo, o, WAVE), __shfl_up(v.hi, o, WAVE));
}
template <class V>
__device__ __forceinline__ V shfl_xor_v(const V& v, int o) {
  return __shfl_xor(v, o, WAVE);
}
template <>
__device__ __forceinline__ u128 shfl_xor_v<u128>(const u128& v, int o) {
  return u128(__shfl_xor(v.lo, o, WAVE), __shfl_xor(v.hi, o, WAVE));
}

// Block sum in a fixed order for any V with operator+.
template <class V>
__device__ __forceinline__ V block_sum_v(V v, V* lds) {
  constexpr int NW = SCAN_BLOCK / WAVE;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) v = v + shfl_xor_v(v, o);
  const int w = threadIdx.x / WAVE, l = threadIdx.x & (WAVE - 1);
  __syncthreads();
  if (l == 0) lds[w] = v;
  __syncthreads();
  V r = lds[0];
#pragma unroll
  for (int i = 1; i < NW; ++i) r = r + lds[i];
  return r;
}

// Loader: V operator()(long i) const, valid for i < P.
// Writer: void operator()(long i, V inclusive, V exclusive) const.
// Gate:   bool operator()() const (false: the launch does nothing).
template <class V, class Loader, class Writer, class Gate>
__global__ __launch_bounds__(SCAN_BLOCK) void k_tile_scan(Loader load, Writer write, Gate gate,
                                                           int P, ScanWorkspace ws) {
  constexpr int NW = SCAN_BLOCK / WAVE;
  __shared__ uint32_t s_tile, s_epoch;
  __shared__ V lds[NW];
  if (!gate()) return;

  if (threadIdx.x == 0) {
    const uint32_t e =
        __hip_atomic_load(&ws.state->epoch, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) + 1u;
    const uint32_t t =
        __hip_atomic_fetch_add(&ws.state->ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    s_tile = t;
    s_epoch = e;
  }
  __syncthreads();
  const uint32_t tile = s_tile, epoch = s_epoch;
  const int ntiles = ceil_div(P, SCAN_TILE);
  if (tile >= (uint32_t)ntiles) {  // a stale ticket: report, touch nothing
    if (threadIdx.x == 0)
      __hip_atomic_store(&ws.state->error, 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }

  // thread-local inclusive scan of SCAN_ITEMS consecutive elements
  const long base = (long)tile * SCAN_TILE + (long)threadIdx.x * SCAN_ITEMS;
  V v[SCAN_ITEMS];
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) v[i] = (base + i < P) ? load(base + i) : V(0);
#pragma unroll
  for (int i = 1; i < SCAN_ITEMS; ++i) v[i] = v[i - 1] + v[i];
  const V total = v[SCAN_ITEMS - 1];

  // block exclusive scan of thread totals: wave scan, then wave prefixes
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  V incl = total;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const V u = shfl_up_v(incl, o);
    if (lane >= o) incl = u + incl;
  }
  V excl = shfl_up_v(incl, 1);
  if (lane == 0) excl = V(0);
  if (lane == WAVE - 1) lds[w] = incl;
  __syncthreads();
  V wpre = V(0), agg = V(0);
#pragma unroll
  for (int i = 0; i < NW; ++i) {
    if (i < w) wpre = wpre + lds[i];
    agg = agg + lds[i];
  }
  const V texcl = wpre + excl;

  // publish this tile's aggregate
  if (threadIdx.x == 0) {
    uint64_t a0, a1;
    Words<V>::pack(agg, a0, a1);
    __hip_atomic_store(&ws.aggregates[2 * tile], a0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&ws.aggregates[2 * tile + 1], a1, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&ws.flags[tile], epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }

  // prefix = sum of the aggregates of tiles [0, tile), fixed order.  Flags
  // are polled with relaxed loads (an acquire per poll would invalidate the
  // caches on every iteration); one acquire fence orders the aggregate reads.
  for (uint32_t j = threadIdx.x; j < tile; j += SCAN_BLOCK) {
    uint32_t spins = 0;
    while (__hip_atomic_load(&ws.flags[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
      if (++spins > SCAN_SPIN_LIMIT) {
        __hip_atomic_store(&ws.state->error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  V part = V(0);
  for (uint32_t j = threadIdx.x; j < tile; j += SCAN_BLOCK) {
    const uint64_t a0 =
        __hip_atomic_load(&ws.aggregates[2 * j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t a1 =
        __hip_atomic_load(&ws.aggregates[2 * j + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    part = part + Words<V>::unpack(a0, a1);
  }
  const V prefix = block_sum_v(part, lds);

#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    if (base + i < P) {
      const V before = (i == 0) ? V(0) : v[i - 1];
      write(base + i, prefix + (texcl + v[i]), prefix + (texcl + before));
    }
  }

  // the last ticket rearms the workspace for the next launch
  if (threadIdx.x == 0 && tile == (uint32_t)(ntiles - 1)) {
    __hip_atomic_store(&ws.state->ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_add(&ws.state->epoch, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
}

struct AlwaysGate {
  __device__ bool operator()() const { return true; }
};

struct PFGate {
  const PFDevice* dev;
  __device__ bool operator()() const {
    return dev == nullptr || (dev->resample != 0 && dev->status == BIHIP_OK);
  }
};

}  // namespace bihip
