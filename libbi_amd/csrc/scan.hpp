// Single-pass tile scan with decoupled look-back over tile aggregates
// (after Merrill & Garland, "Single-pass parallel prefix scan with decoupled
// look-back", 2016).  Each tile publishes only its aggregate; a tile's
// prefix is the sum of all earlier aggregates, reduced by the whole block in
// a fixed order.  Tiles never wait on each other's prefixes.  Values are
// integers (counts, or weights in 128-bit fixed point), so every association
// order gives the same exact, monotone result.
//
// Publication needs no memory fence: an aggregate is split into 32-bit
// chunks and each chunk is stored as one 64-bit word (epoch << 32 | chunk)
// with a relaxed device-scope atomic, so every word validates itself.  (On
// gfx950 a release at device scope writes back the whole L2, once per tile.)
// The epoch lives on the device and advances once per launch, so no memset
// is needed between launches and a captured graph can replay the scan.
// Tiles are taken from an atomic ticket, so every tile waited on has already
// started; a spin limit turns a lost tile into an error flag, never a hang.
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
  __device__ __forceinline__ u128 operator-(const u128& o) const {
    return u128(lo - o.lo, hi - o.hi - (lo < o.lo ? 1u : 0u));
  }
};

// A value as 32-bit chunks.
template <class V>
struct Chunks;
template <>
struct Chunks<uint64_t> {
  static constexpr int N = 2;
  __device__ static void split(uint64_t v, uint32_t (&c)[N]) {
    c[0] = (uint32_t)v;
    c[1] = (uint32_t)(v >> 32);
  }
  __device__ static uint64_t join(const uint32_t (&c)[N]) {
    return (uint64_t)c[0] | ((uint64_t)c[1] << 32);
  }
};
template <>
struct Chunks<u128> {
  static constexpr int N = 4;
  __device__ static void split(const u128& v, uint32_t (&c)[N]) {
    c[0] = (uint32_t)v.lo;
    c[1] = (uint32_t)(v.lo >> 32);
    c[2] = (uint32_t)v.hi;
    c[3] = (uint32_t)(v.hi >> 32);
  }
  __device__ static u128 join(const uint32_t (&c)[N]) {
    return u128((uint64_t)c[0] | ((uint64_t)c[1] << 32), (uint64_t)c[2] | ((uint64_t)c[3] << 32));
  }
};

template <>
struct Chunks<double> {
  static constexpr int N = 2;
  __device__ static void split(double v, uint32_t (&c)[N]) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    c[0] = (uint32_t)b;
    c[1] = (uint32_t)(b >> 32);
  }
  __device__ static double join(const uint32_t (&c)[N]) {
    return __longlong_as_double((long long)((uint64_t)c[0] | ((uint64_t)c[1] << 32)));
  }
};

// Publish / await a value in epoch-tagged words (see the file comment).
template <class V>
__device__ __forceinline__ void publish(uint64_t* words, uint32_t epoch, const V& v) {
  uint32_t c[Chunks<V>::N];
  Chunks<V>::split(v, c);
#pragma unroll
  for (int i = 0; i < Chunks<V>::N; ++i)
    __hip_atomic_store(&words[i], ((uint64_t)epoch << 32) | c[i], __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

template <class V>
__device__ __forceinline__ V await_value(uint64_t* words, uint32_t epoch, uint32_t* error) {
  uint32_t c[Chunks<V>::N];
#pragma unroll
  for (int i = 0; i < Chunks<V>::N; ++i) {
    uint32_t spins = 0;
    uint64_t w;
    while (((w = __hip_atomic_load(&words[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) >>
            32) != epoch) {
      if (++spins > SCAN_SPIN_LIMIT) {
        __hip_atomic_store(error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    c[i] = (uint32_t)w;
  }
  return Chunks<V>::join(c);
}

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

// Exclusive scan of one value per thread across the block, in a fixed order;
// agg receives the block total.  lds holds SCAN_BLOCK / WAVE values.
template <class V>
__device__ __forceinline__ V block_exclusive_scan(V v, V* lds, V& agg) {
  constexpr int NW = SCAN_BLOCK / WAVE;
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  V incl = v;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const V u = shfl_up_v(incl, o);
    if (lane >= o) incl = u + incl;
  }
  V excl = shfl_up_v(incl, 1);
  if (lane == 0) excl = V(0);
  __syncthreads();  // lds may still be read by an earlier reduction
  if (lane == WAVE - 1) lds[w] = incl;
  __syncthreads();
  V wpre = V(0), a = V(0);
#pragma unroll
  for (int i = 0; i < NW; ++i) {
    if (i < w) wpre = wpre + lds[i];
    a = a + lds[i];
  }
  agg = a;
  return wpre + excl;
}

// Weights w in [0, 1] as 128-bit integers with 62 fractional bits (w * 2^62 is
// exact in double and truncates to an integer below 2^62).  Sums of them are
// exact, so every association order gives the same monotone prefixes.
constexpr double FIX_SCALE = 0x1.0p62;
constexpr double FIX_UNIT = 0x1.0p-62;

__device__ __forceinline__ u128 fix_weight(double w) { return u128((uint64_t)(w * FIX_SCALE)); }

// hi * 2^64 * 2^-62 is exact; lo is rounded once; the sum rounds once
// (monotone in (hi, lo), including when lo wraps into hi).
__device__ __forceinline__ double fix_to_double(const u128& v) {
  return (double)v.hi * 4.0 + (double)v.lo * FIX_UNIT;
}

// Loader: V operator()(long i) const, valid for i < P.
// Writer: void operator()(long i, V inclusive, V exclusive) const.
// Gate:   bool operator()() const (false: the launch does nothing).
// ITEMS consecutive elements per thread; TICKET: tiles are numbered by an
// atomic ticket (any grid size) or by blockIdx.x (grids that fit on the
// device at once, so every awaited tile is running).
template <class V, int ITEMS, bool TICKET, class Loader, class Writer, class Gate>
__global__ __launch_bounds__(SCAN_BLOCK) void k_tile_scan(Loader load, Writer write, Gate gate,
                                                           int P, ScanWorkspace ws) {
  constexpr int NW = SCAN_BLOCK / WAVE;
  constexpr int NC = Chunks<V>::N;
  constexpr int TILE = SCAN_BLOCK * ITEMS;
  __shared__ uint32_t s_tile, s_epoch;
  __shared__ V lds[NW];
  if (!gate()) return;

  if (threadIdx.x == 0) {
    // the epoch is read before the tile number; the last tile advances it
    // only after every other tile has published (and so has read it)
    s_epoch =
        __hip_atomic_load(&ws.state->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u;
    s_tile = TICKET ? __hip_atomic_fetch_add(&ws.state->ticket, 1u, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT)
                    : blockIdx.x;
  }
  __syncthreads();
  const uint32_t tile = s_tile, epoch = s_epoch;
  const int ntiles = ceil_div(P, TILE);
  if (tile >= (uint32_t)ntiles) {  // a stale ticket: report, touch nothing
    if (threadIdx.x == 0)
      __hip_atomic_store(&ws.state->error, 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }

  // thread-local inclusive scan of ITEMS consecutive elements
  const long base = (long)tile * TILE + (long)threadIdx.x * ITEMS;
  V v[ITEMS];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) v[i] = (base + i < P) ? load(base + i) : V(0);
#pragma unroll
  for (int i = 1; i < ITEMS; ++i) v[i] = v[i - 1] + v[i];
  const V total = v[ITEMS - 1];

  // block exclusive scan of thread totals
  V agg;
  const V texcl = block_exclusive_scan(total, lds, agg);

  // publish the tile's aggregate at once; then decoupled look-back by wave
  // 0: lane l reads predecessor hi - 1 - l, taking its inclusive prefix when
  // that is published and its aggregate otherwise (waiting, bounded, for one
  // of the two); the window ends at the nearest inclusive prefix, else moves
  // 64 tiles back.  The sums are exact integers, so the order of the
  // additions cannot change the result.
  uint64_t* agg_words = ws.words;                              // NC words per tile
  uint64_t* incl_words = ws.words + 4 * (size_t)ws.max_tiles;  // NC words per tile
  if (threadIdx.x == 0) publish(agg_words + (size_t)NC * tile, epoch, agg);
  __shared__ V s_prefix;
  if (threadIdx.x < WAVE) {
    const int lane = threadIdx.x;
    V run = V(0);
    long hi = tile;
    while (hi > 0) {
      const long j = hi - 1 - lane;
      bool incl = j < 0;  // before tile 0: an inclusive prefix of zero
      V val = V(0);
      if (j >= 0) {
        const uint64_t* iw = incl_words + (size_t)NC * j;
        const uint64_t* aw = agg_words + (size_t)NC * j;
        uint32_t spins = 0;
        for (;;) {
          if ((__hip_atomic_load(iw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> 32) == epoch) {
            incl = true;
            break;
          }
          if ((__hip_atomic_load(aw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> 32) == epoch)
            break;
          if (++spins > SCAN_SPIN_LIMIT) {
            __hip_atomic_store(&ws.state->error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        val = await_value<V>(const_cast<uint64_t*>(incl ? iw : aw), epoch, &ws.state->error);
      }
      const uint64_t found = __ballot(incl);
      const int stop = found != 0 ? __ffsll((unsigned long long)found) - 1 : WAVE - 1;
      V mine = lane <= stop ? val : V(0);
#pragma unroll
      for (int o = 1; o < WAVE; o <<= 1) mine = mine + shfl_xor_v(mine, o);
      run = run + mine;
      if (found != 0) break;
      hi -= WAVE;
    }
    if (lane == 0) {
      s_prefix = run;
      publish(incl_words + (size_t)NC * tile, epoch, run + agg);
    }
  }
  __syncthreads();
  const V prefix = s_prefix;

#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    if (base + i < P) {
      const V before = (i == 0) ? V(0) : v[i - 1];
      write(base + i, prefix + (texcl + v[i]), prefix + (texcl + before));
    }
  }

  // the last tile rearms the workspace for the next launch
  if (threadIdx.x == 0 && tile == (uint32_t)(ntiles - 1)) {
    if (TICKET) __hip_atomic_store(&ws.state->ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&ws.state->epoch, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Tiles that are certainly resident together (two per CU); grids up to this
// size number their tiles by blockIdx.x, larger ones take tickets.
int scan_resident_tiles();

template <class V, class Loader, class Writer, class Gate>
int launch_tile_scan(Loader load, Writer write, Gate gate, int P, ScanWorkspace ws,
                     hipStream_t s) {
  const int nt = ceil_div(P, SCAN_TILE);
  if (nt > ws.max_tiles) return set_error(BIHIP_EINVAL, "scan workspace too small");
  if (nt <= scan_resident_tiles()) {
    hipLaunchKernelGGL((k_tile_scan<V, SCAN_ITEMS, false, Loader, Writer, Gate>), dim3(nt),
                       dim3(SCAN_BLOCK), 0, s, load, write, gate, P, ws);
  } else {
    hipLaunchKernelGGL((k_tile_scan<V, SCAN_ITEMS, true, Loader, Writer, Gate>), dim3(nt),
                       dim3(SCAN_BLOCK), 0, s, load, write, gate, P, ws);
  }
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
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
