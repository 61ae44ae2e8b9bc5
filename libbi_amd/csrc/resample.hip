// Resampling kernels: cumulative offspring (systematic, stratified),
// multinomial ancestors, the permuted ancestor layout and the ancestor
// gather.
//
// Replaces (ON_DEVICE):
//   SystematicResampler::cumulativeOffspring  share/src/bi/resampler/SystematicResampler.hpp:110-130
//   StratifiedResamplerGPU / Host::op         share/src/bi/host/resampler/StratifiedResamplerHost.hpp:25-74
//   MultinomialResamplerGPU::ancestors        share/src/bi/cuda/resampler/MultinomialResamplerGPU.hpp:26-51
//   cumulativeOffspringToAncestorsPermute     share/src/bi/host/resampler/ResamplerHost.hpp:133-168
//   cumulativeOffspringToAncestors            share/src/bi/host/resampler/ResamplerHost.hpp:109-131
//   ancestorsToOffspring                      share/src/bi/host/resampler/ResamplerHost.hpp:55-68
//   gather_rows                               share/src/bi/cuda/primitive/matrix_primitive.cuh:67-78
//
// The permuted layout is computed without the reference GPU's atomicMin
// claims (share/src/bi/cuda/resampler/ResamplerKernel.cuh:129-165): with
// e_i = max(o_i - 1, 0) and f_k = [o_k == 0], E and F their exclusive scans,
// slot k keeps particle k when o_k > 0, and otherwise receives the particle
// i whose extra copies cover rank F_k, i.e. the largest i with E_i <= F_k.
// This is the host's serial two-pointer layout, bit for bit.
#include <algorithm>

#include "scan.hpp"

namespace bihip {

constexpr int RS_BLOCK = 256;

// O_i = min(n, (int)(Ws_i / W * n + a)) with every operation rounded as
// written (SystematicResampler.hpp:47-50).
__device__ __forceinline__ int systematic_offspring(double Ws, double W, int n, double a) {
#pragma clang fp contract(off)
  const double v = Ws / W * (double)n + a;
  const int o = (int)v;
  return o < n ? o : n;
}

// StratifiedResamplerHost.hpp:68-71
__device__ __forceinline__ double stratified_reach(double Ws, double W, int n) {
#pragma clang fp contract(off)
  return Ws / W * (double)n;
}

__device__ __forceinline__ int stratified_offspring(double reach, double alpha, int n) {
#pragma clang fp contract(off)
  const int o = (int)(reach + alpha);
  return o < n ? o : n;
}

__device__ __forceinline__ bool gate_open(const PFDevice* g) {
  return g == nullptr || (g->resample != 0 && g->status == BIHIP_OK);
}

struct OffspringArgs {
  int mode;
  const double* Ws;
  int P, n;
  double a;
  const double* alphas;
  int philox;
  uint64_t seed;
  uint32_t step, batch;
  const PFDevice* gate;
  PFDevice* status;
  int* Os;
};

__global__ __launch_bounds__(RS_BLOCK) void k_cumulative_offspring(OffspringArgs a) {
  if (!gate_open(a.gate)) return;
  const double W = a.Ws[a.P - 1];
  if (!(W > 0.0)) {
    if (a.status != nullptr && blockIdx.x == 0 && threadIdx.x == 0)
      a.status->status = BIHIP_DEGENERATE;
    return;
  }
  const int i = blockIdx.x * RS_BLOCK + threadIdx.x;
  if (i >= a.P) return;
  const double Wi = a.Ws[i];
  int o;
  const uint32_t step = a.gate != nullptr ? (uint32_t)a.gate->k : a.step;
  if (a.mode == BIHIP_SYSTEMATIC) {
    double av = a.a;
    if (a.philox) {
      const u32x4 r = philox_block(a.seed, 0, step, TAG_SYSTEMATIC, a.batch, 0);
      av = u53(r.x, r.y);
    }
    o = systematic_offspring(Wi, W, a.n, av);
  } else {
    const double reach = stratified_reach(Wi, W, a.n);
    int k = (int)reach;
    if (k > a.n - 1) k = a.n - 1;
    double alpha;
    if (a.philox) {
      const u32x4 r = philox_block(a.seed, (uint32_t)k, step, TAG_STRATIFIED, a.batch, 0);
      alpha = u53(r.x, r.y);
    } else {
      alpha = a.alphas[k];
    }
    o = stratified_offspring(reach, alpha, a.n);
  }
  a.Os[i] = o;
}

int launch_cumulative_offspring(const OffspringLaunch& L, hipStream_t s) {
  if (L.P <= 0) return BIHIP_OK;
  OffspringArgs a{L.mode, L.Ws, L.P, L.n, L.a, L.alphas, L.philox ? 1 : 0, L.seed, L.step,
                  L.batch, L.gate, L.status, L.Os};
  hipLaunchKernelGGL(k_cumulative_offspring, dim3(ceil_div(L.P, RS_BLOCK)), dim3(RS_BLOCK), 0, s,
                     a);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

// ---------------------------------------------------------------- tiled chain
// The permuted layout and the filter's whole resampling chain run as
// reduce-then-scan over tiles of SCAN_TILE consecutive particles.  Every
// dependency between tiles crosses a kernel boundary; no tile waits for
// another inside a launch:
//   k_weight_sums      tile sums of the fixed-point weights
//   k_tiles_exclusive  exclusive scan of tile sums, in one block; only above
//                      SMALL_TILES tiles (below, every block sums the tile
//                      sums before its own, which stay in L2)
//   k_offspring_tiles  the tile's inclusive weight sums Ws_i -> cumulative
//                      offspring O_i (systematic, stratified) or Ws
//                      (multinomial); tile sums of (e, f)
//   k_ef_sums          tile sums of (e, f) for offspring given from outside
//   k_place_tiles      the tile's exclusive E and F: E_k stored, occupied
//                      slots keep their particle, free slot k listed at rank F_k
//   k_place_free       one thread per free slot of rank r: the largest i with
//                      E_i <= r, by binary search over E
// All sums are integers, so the results equal a sequential scan's exactly.
constexpr int SMALL_TILES = 1024;

__device__ __forceinline__ int offspring_at(const int* in, bool cumulative, long i) {
  return cumulative ? in[i] - (i > 0 ? in[i - 1] : 0) : in[i];
}

// (e, f) = (max(o - 1, 0), [o == 0]) packed as e | f << 32.
__device__ __forceinline__ uint64_t ef_pair(int o) {
  const uint64_t e = o > 1 ? (uint64_t)(o - 1) : 0u;
  const uint64_t f = o == 0 ? 1u : 0u;
  return e | (f << 32);
}

// Regions of the workspace's tile-chain words (the upper half of each tile's
// SCAN_WORDS; the totals follow the tile sums).
struct TileWork {
  u128* wsum;           // weight sum per tile
  uint64_t* efsum;      // (e, f) sum per tile
  u128* wtotal;         // after k_tiles_exclusive
  uint64_t* eftotal;
  uint32_t* free_slots;
};

uint64_t* tile_sum_words(const ScanWorkspace& ws) { return ws.words + 8 * (size_t)ws.max_tiles; }

static TileWork tile_work(const ScanWorkspace& ws) {
  const size_t mt = (size_t)ws.max_tiles;
  uint64_t* base = tile_sum_words(ws);
  TileWork w;
  w.wsum = reinterpret_cast<u128*>(base);  // 2 mt words
  w.efsum = base + 2 * mt;                 // mt words
  w.wtotal = reinterpret_cast<u128*>(base + 3 * mt);
  w.eftotal = base + 3 * mt + 2;           // used only with mt > SMALL_TILES
  w.free_slots = &ws.state->free_slots;
  return w;
}

// Tile exchanges through LDS between the striped order (element
// j * SCAN_BLOCK + tid of the tile: coalesced global access) and the blocked
// order (element tid * SCAN_ITEMS + j: a thread's consecutive elements, as
// the scans need them).  One padding element per 128 bytes spreads the
// blocked accesses over the banks.
template <class V>
__device__ __forceinline__ int xpos(int i) {
  return i + i / (128 / (int)sizeof(V));
}
template <class V>
constexpr int xsize() {
  return SCAN_TILE + SCAN_TILE / (128 / (int)sizeof(V));
}

// Sum of the tile sums before `tile` and of all of them, in two parts so that
// a kernel can issue the loads before its gate check and its other loads:
// tile_prefix_load reads this thread's share (at most SMALL_TILES / SCAN_BLOCK
// sums), tile_prefix_combine adds the shares across the block.
template <class V>
struct TilePrefix {
  V pre, tot;
};

template <class V>
__device__ __forceinline__ TilePrefix<V> tile_prefix_load(const V* sums, const V* total, int nt,
                                                          int tile) {
  TilePrefix<V> r{V(0), V(0)};
  if (nt <= SMALL_TILES) {
#pragma unroll
    for (int q = 0; q < SMALL_TILES / SCAN_BLOCK; ++q) {
      const int j = q * SCAN_BLOCK + threadIdx.x;
      if (j < nt) {
        const V v = sums[j];
        r.tot = r.tot + v;
        if (j < tile) r.pre = r.pre + v;
      }
    }
  } else {
    r.pre = sums[tile];
    r.tot = *total;
  }
  return r;
}

template <class V>
__device__ __forceinline__ void tile_prefix_combine(TilePrefix<V>& r, int nt, V* lds) {
  if (nt <= SMALL_TILES) {
    r.pre = block_sum_v(r.pre, lds);
    r.tot = block_sum_v(r.tot, lds);
  }
}

// The chain's kernels run filter blockIdx.y of a bank at the strides bk (a
// lone filter: y = 0).
__device__ __forceinline__ long bank_f() { return (long)blockIdx.y; }
constexpr long DEV_BYTES = (long)sizeof(PFDevice);

__global__ __launch_bounds__(SCAN_BLOCK) void k_weight_sums(const double* lws, const double* mx,
                                                            int P, u128* wsum,
                                                            const PFDevice* gate, BankStrides bk) {
  __shared__ u128 lds[SCAN_BLOCK / WAVE];
  const long f = bank_f();
  lws = bank_shift(lws, f * bk.p * 8);
  mx = bank_shift(mx, f * DEV_BYTES);
  wsum = bank_shift(wsum, f * bk.words);
  gate = bank_shift(gate, f * DEV_BYTES);
  // the loads are issued before the gate check (all in bounds) to overlap
  // their latency with the gate's
  const long t0 = (long)blockIdx.x * SCAN_TILE;
  double lw[SCAN_ITEMS];
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    const long i = t0 + (long)j * SCAN_BLOCK + threadIdx.x;
    lw[j] = i < P ? lws[i] : __builtin_nan("");  // NaN weighs zero
  }
  const double m = *mx;
  if (!gate_open(gate)) return;
  u128 s(0);
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) s = s + fix_weight(nanexp(lw[j] - m));
  s = block_sum_v(s, lds);
  if (threadIdx.x == 0) wsum[blockIdx.x] = s;
}

// In-place exclusive scan of nt tile sums by one block; total receives the sum.
template <class V>
__global__ __launch_bounds__(SCAN_BLOCK) void k_tiles_exclusive(V* sums, int nt, V* total,
                                                                const PFDevice* gate,
                                                                BankStrides bk) {
  __shared__ V lds[SCAN_BLOCK / WAVE];
  const long f = bank_f();
  sums = bank_shift(sums, f * bk.words);
  total = bank_shift(total, f * bk.words);
  gate = bank_shift(gate, f * DEV_BYTES);
  if (!gate_open(gate)) return;
  V carry(0);
  for (long c0 = 0; c0 < nt; c0 += SCAN_TILE) {
    const long base = c0 + (long)threadIdx.x * SCAN_ITEMS;
    V v[SCAN_ITEMS];
    V t(0);
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
      v[j] = base + j < nt ? sums[base + j] : V(0);
      t = t + v[j];
    }
    V agg;
    V run = carry + block_exclusive_scan(t, lds, agg);
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
      if (base + j < nt) sums[base + j] = run;
      run = run + v[j];
    }
    carry = carry + agg;
  }
  if (threadIdx.x == 0) *total = carry;
}

struct OffspringTileArgs {
  const double* lws;
  const double* mx;
  int P, n, nt;
  const u128* wsum;
  const u128* wtotal;
  uint64_t* efsum;
  uint64_t seed;
  uint32_t batch;
  const PFDevice* gate;
  PFDevice* status;
  double* Ws;
  int* Os;
  BankStrides bk;
};

// MODE: BIHIP_SYSTEMATIC or BIHIP_STRATIFIED (cumulative offspring, as
// k_cumulative_offspring from the Ws of launch_sumexp_scan, value for value),
// or BIHIP_MULTINOMIAL (the inclusive sums Ws only).
template <int MODE>
__global__ __launch_bounds__(SCAN_BLOCK) void k_offspring_tiles(OffspringTileArgs a) {
  __shared__ u128 lds[SCAN_BLOCK / WAVE];
  __shared__ uint64_t lds64[SCAN_BLOCK / WAVE];
  {
    const long f = bank_f();
    a.lws = bank_shift(a.lws, f * a.bk.p * 8);
    a.mx = bank_shift(a.mx, f * DEV_BYTES);
    a.wsum = bank_shift(a.wsum, f * a.bk.words);
    a.wtotal = bank_shift(a.wtotal, f * a.bk.words);
    a.efsum = bank_shift(a.efsum, f * a.bk.words);
    a.batch += (uint32_t)f;
    a.gate = bank_shift(a.gate, f * DEV_BYTES);
    a.status = bank_shift(a.status, f * DEV_BYTES);
    a.Ws = bank_shift(a.Ws, f * a.bk.p * 8);
    a.Os = bank_shift(a.Os, f * a.bk.p * 4);
  }
  const int tile = blockIdx.x;
  const long base = (long)tile * SCAN_TILE + (long)threadIdx.x * SCAN_ITEMS;
  // every load is issued before the gate check (all in bounds)
  double lw[SCAN_ITEMS];
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) lw[i] = base + i < a.P ? a.lws[base + i] : 0.0;
  TilePrefix<u128> tp = tile_prefix_load(a.wsum, a.wtotal, a.nt, tile);
  const double m = *a.mx;
  if (!gate_open(a.gate)) return;
  tile_prefix_combine(tp, a.nt, lds);
  const u128 pre = tp.pre;
  const double W = fix_to_double(tp.tot);
  if (MODE != BIHIP_MULTINOMIAL && !(W > 0.0)) {
    if (a.status != nullptr && tile == 0 && threadIdx.x == 0) a.status->status = BIHIP_DEGENERATE;
    return;
  }
  u128 v[SCAN_ITEMS];
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i)
    v[i] = base + i < a.P ? fix_weight(nanexp(lw[i] - m)) : u128(0);
#pragma unroll
  for (int i = 1; i < SCAN_ITEMS; ++i) v[i] = v[i - 1] + v[i];
  u128 agg;
  const u128 start = pre + block_exclusive_scan(v[SCAN_ITEMS - 1], lds, agg);  // Ws_{base-1}
  if constexpr (MODE == BIHIP_MULTINOMIAL) {
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i)
      if (base + i < a.P) a.Ws[base + i] = fix_to_double(start + v[i]);
    return;
  } else {
    const uint32_t step = (uint32_t)a.gate->k;
    double av = 0.0;
    if (MODE == BIHIP_SYSTEMATIC) {
      const u32x4 r = philox_block(a.seed, 0, step, TAG_SYSTEMATIC, a.batch, 0);
      av = u53(r.x, r.y);
    }
    auto cumulative = [&](const u128& incl) -> int {
      const double Wsi = fix_to_double(incl);
      if (MODE == BIHIP_SYSTEMATIC) return systematic_offspring(Wsi, W, a.n, av);
      const double reach = stratified_reach(Wsi, W, a.n);
      int k = (int)reach;
      if (k > a.n - 1) k = a.n - 1;
      const u32x4 r = philox_block(a.seed, (uint32_t)k, step, TAG_STRATIFIED, a.batch, 0);
      return stratified_offspring(reach, u53(r.x, r.y), a.n);
    };
    uint64_t ef = 0;
    if (base < a.P) {
      int prev = base > 0 ? cumulative(start) : 0;
#pragma unroll
      for (int i = 0; i < SCAN_ITEMS; ++i) {
        if (base + i >= a.P) break;
        const int O = cumulative(start + v[i]);
        a.Os[base + i] = O;
        ef += ef_pair(O - prev);
        prev = O;
      }
    }
    ef = block_sum_v(ef, lds64);
    if (threadIdx.x == 0) a.efsum[tile] = ef;
  }
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_ef_sums(const int* in, int cumulative, int P,
                                                        uint64_t* efsum, const PFDevice* gate) {
  __shared__ uint64_t lds[SCAN_BLOCK / WAVE];
  if (!gate_open(gate)) return;
  const long t0 = (long)blockIdx.x * SCAN_TILE;
  uint64_t s = 0;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    const long i = t0 + (long)j * SCAN_BLOCK + threadIdx.x;
    if (i < P) s += ef_pair(offspring_at(in, cumulative != 0, i));
  }
  s = block_sum_v(s, lds);
  if (threadIdx.x == 0) efsum[blockIdx.x] = s;
}

struct PlaceTileArgs {
  const int* in;
  int cumulative;
  int P, nt;
  const uint64_t* efsum;
  const uint64_t* eftotal;
  uint32_t* E;
  uint32_t* slots;
  int* as;
  uint32_t* free_slots;
  const PFDevice* gate;
  BankStrides bk;
};

constexpr uint32_t OCCUPIED = 0xFFFFFFFFu;

__global__ __launch_bounds__(SCAN_BLOCK) void k_place_tiles(PlaceTileArgs a) {
  __shared__ uint64_t lds[SCAN_BLOCK / WAVE];
  __shared__ int xi[xsize<int>()];       // offspring input, then E
  __shared__ uint32_t xf[xsize<int>()];  // free-slot rank, or OCCUPIED
  __shared__ int s_before;
  {
    const long f = bank_f();
    a.in = bank_shift(a.in, f * a.bk.p * 4);
    a.efsum = bank_shift(a.efsum, f * a.bk.words);
    a.eftotal = bank_shift(a.eftotal, f * a.bk.words);
    a.E = bank_shift(a.E, f * a.bk.p * 8);      // E and the slots share the filter's EF words
    a.slots = bank_shift(a.slots, f * a.bk.p * 8);
    a.as = bank_shift(a.as, f * a.bk.p * 4);
    a.free_slots = bank_shift(a.free_slots, f * a.bk.state);
    a.gate = bank_shift(a.gate, f * DEV_BYTES);
  }
  const int tile = blockIdx.x;
  const long t0 = (long)tile * SCAN_TILE;
  // every load is issued before the gate check (all in bounds)
  int c[SCAN_ITEMS];
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    const int e = j * SCAN_BLOCK + threadIdx.x;
    c[j] = t0 + e < a.P ? a.in[t0 + e] : 0;
  }
  const int before = (a.cumulative && t0 > 0 && threadIdx.x == 0) ? a.in[t0 - 1] : 0;
  TilePrefix<uint64_t> tp = tile_prefix_load(a.efsum, a.eftotal, a.nt, tile);
  if (!gate_open(a.gate)) return;
  tile_prefix_combine(tp, a.nt, lds);
  const uint64_t pre = tp.pre, tot = tp.tot;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) xi[xpos<int>(j * SCAN_BLOCK + threadIdx.x)] = c[j];
  if (threadIdx.x == 0) s_before = before;
  __syncthreads();
  const int b0 = threadIdx.x * SCAN_ITEMS;
  const long base = t0 + b0;
  int o[SCAN_ITEMS];
  uint64_t t = 0;
  int prev = b0 > 0 ? xi[xpos<int>(b0 - 1)] : s_before;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    o[i] = 1;  // past the end: no extras, not free
    if (base + i < a.P) {
      const int c = xi[xpos<int>(b0 + i)];
      o[i] = a.cumulative ? c - prev : c;
      prev = c;
    }
    t += ef_pair(o[i]);
  }
  uint64_t agg;
  uint64_t run = pre + block_exclusive_scan(t, lds, agg);  // its barriers end the reads of xi
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    xi[xpos<int>(b0 + i)] = (int)(uint32_t)run;
    xf[xpos<int>(b0 + i)] = o[i] > 0 ? OCCUPIED : (uint32_t)(run >> 32);
    run += ef_pair(o[i]);
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    const int e = j * SCAN_BLOCK + threadIdx.x;
    const long k = t0 + e;
    if (k < a.P) {
      a.E[k] = (uint32_t)xi[xpos<int>(e)];
      const uint32_t f = xf[xpos<int>(e)];
      if (f == OCCUPIED) a.as[k] = (int)k;
      else a.slots[f] = (uint32_t)k;
    }
  }
  if (tile == a.nt - 1 && threadIdx.x == 0) *a.free_slots = (uint32_t)(tot >> 32);
}

// Free slot of rank r receives the largest i with E_i <= r (the particle
// whose extra copies cover rank r).  A wave holds 64 consecutive ranks, whose
// owners are non-decreasing: it narrows one interval [lo, hi) around all of
// them by 64-ary search (each lane loads one pivot; a ballot counts the pivots
// <= the first and the last rank), so a search over 2^20 particles takes
// four dependent loads instead of twenty.  Once the interval holds at most 64
// particles every lane finds its own owner among them by shuffles; if the
// owners are spread too far apart to narrow further, each lane finishes with
// a binary search inside the interval.
__global__ __launch_bounds__(RS_BLOCK) void k_place_free(const uint32_t* E, const uint32_t* slots,
                                                         const uint32_t* free_slots, int P,
                                                         int* as, const PFDevice* gate,
                                                         BankStrides bk) {
  {
    const long f = bank_f();
    E = bank_shift(E, f * bk.p * 8);
    slots = bank_shift(slots, f * bk.p * 8);
    free_slots = bank_shift(free_slots, f * bk.state);
    as = bank_shift(as, f * bk.p * 4);
    gate = bank_shift(gate, f * DEV_BYTES);
  }
  const uint32_t r = blockIdx.x * RS_BLOCK + threadIdx.x;
  // slots[r] is loaded with the gate and the count (r < P: in bounds; used
  // only for r < Z)
  const uint32_t kr = r < (uint32_t)P ? slots[r] : 0u;
  const uint32_t Z = *free_slots;
  if (!gate_open(gate)) return;
  const uint32_t rA = r & ~(uint32_t)(WAVE - 1);  // the wave's first rank
  if (rA >= Z) return;                             // uniform across the wave
  const int lane = threadIdx.x & (WAVE - 1);
  const bool valid = r < Z;
  const uint32_t rB = min(rA + (uint32_t)(WAVE - 1), Z - 1);
  const uint32_t k = valid ? kr : 0u;
  int lo = 0, hi = P;  // E[lo] <= rA, and the owners of ranks rA..rB lie in [lo, hi)
  while (hi - lo > WAVE) {
    const int step = (hi - lo + WAVE - 1) / WAVE;
    const long idx = (long)lo + (long)lane * step;
    const uint32_t piv = idx < hi ? E[idx] : 0xFFFFFFFFu;
    const int cA = __popcll(__ballot(piv <= rA));  // >= 1: pivot 0 is E[lo]
    const int cB = __popcll(__ballot(piv <= rB));
    const int nlo = lo + (cA - 1) * step;
    const int nhi = min(hi, lo + cB * step);
    if (nhi - nlo >= hi - lo) break;  // spread over every pivot interval
    lo = nlo;
    hi = nhi;
  }
  int owner;
  if (hi - lo <= WAVE) {
    const uint32_t mine = lo + lane < hi ? E[lo + lane] : 0xFFFFFFFFu;
    int pos = 0;  // E[lo] <= r for every rank of the wave
#pragma unroll
    for (int b = WAVE / 2; b >= 1; b >>= 1) {
      const uint32_t v = (uint32_t)__shfl((int)mine, pos + b, WAVE);
      if (v <= r) pos += b;  // padding lanes hold 0xFFFFFFFF and never qualify
    }
    owner = lo + pos;
  } else {
    int a = lo, b = hi;  // first i in [lo, hi) with E_i > r
    while (a < b) {
      const int mid = a + ((b - a) >> 1);
      if (E[mid] <= r) a = mid + 1;
      else b = mid;
    }
    owner = a - 1;
  }
  if (valid) as[k] = owner;
}

// (e, f) tile sums in place -> E, the free-slot list and the ancestors.
// Ws = the inclusive scan of nanexp(lws - *mx) over P elements, by the tiled
// chain's kernels: exact 128-bit tile sums, (above SMALL_TILES tiles) their
// exclusive scan by one block, then each tile's inclusive prefixes.  The
// sums are exact integers, so the result equals the single-pass scan's bit
// for bit; no tile waits on another (no tickets), whatever the grid size.
int launch_sumexp_scan_tiles(const double* lws, int P, const double* mx, double* Ws,
                             ScanWorkspace ws, hipStream_t s) {
  if (P <= 0) return BIHIP_OK;
  const int nt = ceil_div(P, SCAN_TILE);
  if (nt > ws.max_tiles) return set_error(BIHIP_EINVAL, "scan workspace too small");
  const TileWork w = tile_work(ws);
  hipLaunchKernelGGL(k_weight_sums, dim3(nt), dim3(SCAN_BLOCK), 0, s, lws, mx, P, w.wsum,
                     (const PFDevice*)nullptr, BankStrides{});
  BIHIP_LAUNCH_CHECK();
  if (nt > SMALL_TILES) {
    hipLaunchKernelGGL(k_tiles_exclusive<u128>, dim3(1), dim3(SCAN_BLOCK), 0, s, w.wsum, nt,
                       w.wtotal, (const PFDevice*)nullptr, BankStrides{});
    BIHIP_LAUNCH_CHECK();
  }
  OffspringTileArgs a{lws, mx, P, P, nt, w.wsum, w.wtotal, w.efsum, 0, 0,
                      nullptr, nullptr, Ws, nullptr, BankStrides{}};
  hipLaunchKernelGGL(k_offspring_tiles<BIHIP_MULTINOMIAL>, dim3(nt), dim3(SCAN_BLOCK), 0, s, a);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

// banks filters at the strides bk (1 and zero strides: a lone filter).
static int place_from_sums(const int* in, bool cumulative, int P, int nt, int* as, uint64_t* EF,
                           const TileWork& w, const PFDevice* gate, hipStream_t s, int banks = 1,
                           BankStrides bk = {}) {
  if (nt > SMALL_TILES) {
    hipLaunchKernelGGL(k_tiles_exclusive<uint64_t>, dim3(1, banks), dim3(SCAN_BLOCK), 0, s,
                       w.efsum, nt, w.eftotal, gate, bk);
    BIHIP_LAUNCH_CHECK();
  }
  uint32_t* E = reinterpret_cast<uint32_t*>(EF);  // P words of E, then P of slots
  PlaceTileArgs a{in, cumulative ? 1 : 0, P, nt, w.efsum, w.eftotal, E, E + P, as,
                  w.free_slots, gate, bk};
  hipLaunchKernelGGL(k_place_tiles, dim3(nt, banks), dim3(SCAN_BLOCK), 0, s, a);
  BIHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_place_free, dim3(ceil_div(P, RS_BLOCK), banks), dim3(RS_BLOCK), 0, s, E,
                     E + P, w.free_slots, P, as, gate, bk);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

int launch_ancestors_permute(const int* in, bool cumulative, int P, int* as, uint64_t* EF,
                             ScanWorkspace ws, const PFDevice* gate, hipStream_t s) {
  if (P <= 0) return BIHIP_OK;
  const int nt = ceil_div(P, SCAN_TILE);
  if (nt > ws.max_tiles) return set_error(BIHIP_EINVAL, "scan workspace too small");
  const TileWork w = tile_work(ws);
  hipLaunchKernelGGL(k_ef_sums, dim3(nt), dim3(SCAN_BLOCK), 0, s, in, cumulative ? 1 : 0, P,
                     w.efsum, gate);
  BIHIP_LAUNCH_CHECK();
  return place_from_sums(in, cumulative, P, nt, as, EF, w, gate, s);
}

int launch_resample_tiles(const ResampleTilesLaunch& L, hipStream_t s) {
  if (L.P <= 0) return BIHIP_OK;
  if (L.gate == nullptr) return set_error(BIHIP_EINVAL, "the tiled resampling chain needs a gate");
  const int nt = ceil_div(L.P, SCAN_TILE);
  if (nt > L.ws.max_tiles) return set_error(BIHIP_EINVAL, "scan workspace too small");
  const TileWork w = tile_work(L.ws);
  const int banks = L.banks > 1 ? L.banks : 1;
  if (banks > 1 && L.mode == BIHIP_MULTINOMIAL)
    return set_error(BIHIP_EINVAL, "filter banks resample systematically or stratified");
  if (!L.sums_ready) {
    hipLaunchKernelGGL(k_weight_sums, dim3(nt, banks), dim3(SCAN_BLOCK), 0, s, L.lws, L.mx, L.P,
                       w.wsum, L.gate, L.bk);
    BIHIP_LAUNCH_CHECK();
  } else if (banks > 1) {
    return set_error(BIHIP_EINVAL, "stored tile sums serve a lone filter");
  }
  if (nt > SMALL_TILES) {
    hipLaunchKernelGGL(k_tiles_exclusive<u128>, dim3(1, banks), dim3(SCAN_BLOCK), 0, s, w.wsum,
                       nt, w.wtotal, L.gate, L.bk);
    BIHIP_LAUNCH_CHECK();
  }
  OffspringTileArgs a{L.lws, L.mx, L.P, L.n, nt, w.wsum, w.wtotal, w.efsum, L.seed, L.batch,
                      L.gate, L.status, L.Ws, L.Os, L.bk};
  switch (L.mode) {
    case BIHIP_SYSTEMATIC:
      hipLaunchKernelGGL(k_offspring_tiles<BIHIP_SYSTEMATIC>, dim3(nt, banks), dim3(SCAN_BLOCK), 0,
                         s, a);
      break;
    case BIHIP_STRATIFIED:
      hipLaunchKernelGGL(k_offspring_tiles<BIHIP_STRATIFIED>, dim3(nt, banks), dim3(SCAN_BLOCK), 0,
                         s, a);
      break;
    case BIHIP_MULTINOMIAL:
      hipLaunchKernelGGL(k_offspring_tiles<BIHIP_MULTINOMIAL>, dim3(nt), dim3(SCAN_BLOCK), 0, s,
                         a);
      break;
    default:
      return set_error(BIHIP_EINVAL, "unknown resampler %d", L.mode);
  }
  BIHIP_LAUNCH_CHECK();
  if (L.mode != BIHIP_MULTINOMIAL)
    return place_from_sums(L.Os, true, L.P, nt, L.as, L.EF, w, L.gate, s, banks, L.bk);
  int rc;
  MultinomialLaunch M{};
  M.Ws = L.Ws;
  M.P = L.P;
  M.n = L.n;
  M.seed = L.seed;
  M.batch = L.batch;
  M.gate = L.gate;
  M.status = L.status;
  M.os = L.os;
  if (L.mc != nullptr) {
    if ((rc = launch_multinomial_chunks(M, *L.mc, s))) return rc;  // zeroes os itself
  } else if ((rc = launch_zero_ints(L.os, L.P, L.gate, s)) || (rc = launch_multinomial(M, s))) {
    return rc;
  }
  return launch_ancestors_permute(L.os, false, L.P, L.as, L.EF, L.ws, L.gate, s);
}

// as_j = #{i : O_i <= j}: slot j of the unpermuted ancestor vector.
__global__ __launch_bounds__(RS_BLOCK) void k_cumulative_to_ancestors(const int* Os, int P,
                                                                      int* as) {
  const int n = Os[P - 1] < P ? Os[P - 1] : P;
  const int j = blockIdx.x * RS_BLOCK + threadIdx.x;
  if (j >= n) return;
  int lo = 0, hi = P;
  while (lo < hi) {
    const int mid = lo + ((hi - lo) >> 1);
    if (Os[mid] <= j) lo = mid + 1;
    else hi = mid;
  }
  as[j] = lo < P ? lo : P - 1;
}

int launch_cumulative_to_ancestors(const int* Os, int P, int* as, hipStream_t s) {
  if (P <= 0) return BIHIP_OK;
  // the number of ancestors is Os[P-1], bounded by the caller's buffer of P
  hipLaunchKernelGGL(k_cumulative_to_ancestors, dim3(ceil_div(P, RS_BLOCK)), dim3(RS_BLOCK), 0, s,
                     Os, P, as);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

// ---------------------------------------------------------------- multinomial
struct MultinomialArgs {
  const double* Ws;
  int P, n;
  const double* U;
  uint64_t seed;
  uint32_t step, batch;
  const PFDevice* gate;
  PFDevice* status;
  int* as;
  int* os;
};

// a_j = #{i : Ws_i <= W U_j}: upper bound, never a zero-weight particle.
// A workgroup of MN_BLOCK threads takes MN_DRAWS consecutive draws, MN_ITEMS
// per thread (draw j = block * MN_DRAWS + q * MN_BLOCK + thread, q < MN_ITEMS).
// The search runs in two levels: the block loads the last Ws of each of
// MN_FENCES equal chunks into LDS and finds the chunk of u there, then
// searches that chunk in global memory (11 dependent loads at 2^20 instead of
// 20), the thread's MN_ITEMS searches in lockstep so their loads overlap.
// The offspring counts are gathered per block in an LDS hash table: one
// device-wide atomic per distinct ancestor of the block's MN_DRAWS draws
// (plus one per draw whose slot another ancestor took), so a particle that
// takes thousands of draws is hit by one atomic per block, not one per draw.
// Counts are integers: the result does not depend on the order of the
// additions.
constexpr int MN_FENCES = 512;
constexpr int MN_BLOCK = 1024;
constexpr int MN_ITEMS = 4;
constexpr int MN_DRAWS = MN_BLOCK * MN_ITEMS;
constexpr int MN_HASH_BITS = 12;
constexpr int MN_HASH = 1 << MN_HASH_BITS;

__global__ __launch_bounds__(MN_BLOCK) void k_multinomial(MultinomialArgs a) {
  __shared__ double fence[MN_FENCES];
  __shared__ int hkey[MN_HASH];
  __shared__ int hcnt[MN_HASH];
  if (!gate_open(a.gate)) return;
  const double W = a.Ws[a.P - 1];
  if (!(W > 0.0)) {
    if (a.status != nullptr && blockIdx.x == 0 && threadIdx.x == 0)
      a.status->status = BIHIP_DEGENERATE;
    return;
  }
  const int stride = (a.P + MN_FENCES - 1) / MN_FENCES;
  const int nch = (a.P + stride - 1) / stride;  // chunks, at most MN_FENCES
  for (int c = threadIdx.x; c < MN_FENCES; c += MN_BLOCK) {
    const int last = (c + 1) * stride < a.P ? (c + 1) * stride - 1 : a.P - 1;
    fence[c] = c < nch ? a.Ws[last] : 0.0;
  }
  for (int h = threadIdx.x; h < MN_HASH; h += MN_BLOCK) {
    hkey[h] = -1;
    hcnt[h] = 0;
  }
  __syncthreads();
  const uint32_t step = a.gate != nullptr ? (uint32_t)a.gate->k : a.step;
  int anc[MN_ITEMS], lo[MN_ITEMS], top[MN_ITEMS];
  double u[MN_ITEMS];
#pragma unroll
  for (int q = 0; q < MN_ITEMS; ++q) {
    const int j = blockIdx.x * MN_DRAWS + q * MN_BLOCK + threadIdx.x;
    anc[q] = -1;
    lo[q] = 0;
    top[q] = 0;
    u[q] = 0.0;
    if (j < a.n) {
      double U;
      if (a.U != nullptr) {
        U = a.U[j];
      } else {
        const u32x4 r = philox_block(a.seed, (uint32_t)j, step, TAG_MULTINOMIAL, a.batch, 0);
        U = u53(r.x, r.y);
      }
      u[q] = W * U;
      // chunks lying wholly at or below u (Ws is non-decreasing)
      int f = 0, hi = nch;
      while (f < hi) {
        const int mid = f + ((hi - f) >> 1);
        if (fence[mid] <= u[q]) f = mid + 1;
        else hi = mid;
      }
      lo[q] = f * stride;
      top[q] = f < nch ? (lo[q] + stride < a.P ? lo[q] + stride : a.P) : lo[q];
      anc[q] = 0;  // valid draw
    }
  }
  // the searches inside the chunks, in lockstep
  for (bool more = true; more;) {
    more = false;
#pragma unroll
    for (int q = 0; q < MN_ITEMS; ++q) {
      if (lo[q] < top[q]) {
        const int mid = lo[q] + ((top[q] - lo[q]) >> 1);
        if (a.Ws[mid] <= u[q]) lo[q] = mid + 1;
        else top[q] = mid;
        more = more || lo[q] < top[q];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < MN_ITEMS; ++q) {
    if (anc[q] < 0) continue;
    anc[q] = lo[q] < a.P ? lo[q] : a.P - 1;
    if (a.as != nullptr) a.as[blockIdx.x * MN_DRAWS + q * MN_BLOCK + threadIdx.x] = anc[q];
  }
  if (a.os == nullptr) return;  // uniform: the launch either counts or it does not
#pragma unroll
  for (int q = 0; q < MN_ITEMS; ++q) {
    if (anc[q] < 0) continue;
    const int h = (int)(((uint32_t)anc[q] * 2654435761u) >> (32 - MN_HASH_BITS));
    const int prev = atomicCAS(&hkey[h], -1, anc[q]);
    if (prev == -1 || prev == anc[q]) atomicAdd(&hcnt[h], 1);
    else atomicAdd(&a.os[anc[q]], 1);  // its slot holds another ancestor
  }
  __syncthreads();
  for (int h = threadIdx.x; h < MN_HASH; h += MN_BLOCK)
    if (hkey[h] >= 0) atomicAdd(&a.os[hkey[h]], hcnt[h]);
}

int launch_multinomial(const MultinomialLaunch& L, hipStream_t s) {
  if (L.n <= 0) return BIHIP_OK;
  MultinomialArgs a{L.Ws, L.P, L.n, L.U, L.seed, L.step, L.batch, L.gate, L.status, L.as, L.os};
  hipLaunchKernelGGL(k_multinomial, dim3(ceil_div(L.n, MN_DRAWS)), dim3(MN_BLOCK), 0, s, a);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

// ---------------------------------------------------------------- multinomial by chunks
// The same draws and ancestors (a_j = #{i : Ws_i <= W U_j}), searched in LDS,
// with offspring counts written by the workgroup that owns their chunk:
//   k_mc_bin     each draw's chunk of MC_CHUNK particles (upper bound over the
//                chunks' last Ws, staged in LDS); per binning block, draws per
//                chunk (chunk-major matrix H); os zeroed
//   k_mc_offsets per chunk, the exclusive scan of its column of H over the
//                binning blocks, its total and its segments of at most
//                MC_SEG draws (one for an empty chunk), packed in one word;
//                k_tiles_exclusive over the words gives each chunk's first
//                sorted draw and first segment
//   k_mc_scatter the draws again, each target (and draw index) to its
//                chunk's next sorted position; the segment -> chunk map
//   k_mc_search  one workgroup per segment: the chunk's Ws in LDS (each
//                chunk read once), the segment's draws searched there and
//                counted in LDS, the counts stored (one-segment chunks: plain
//                stores of every particle's count) or added (a chunk of
//                several segments: one atomic per particle drawn).
// Counts are integers and each draw's ancestor depends on its target only, so
// the results equal k_multinomial's bit for bit whatever the order of the
// sorted draws within a chunk.
constexpr int MC_CHUNK = 2048;
constexpr int MC_MAX_CHUNKS = 8192;
constexpr int MC_BLOCK = 1024;
constexpr int MC_BIN_BLOCKS = 256;
constexpr int MC_SEG = 4096;

struct MCArgs {
  const double* Ws;
  int P, n;
  const double* U;  // nullable: Philox draws
  uint64_t seed;
  uint32_t step, batch;
  const PFDevice* gate;
  PFDevice* status;
  int nch, per_block, nb;  // chunks; draws per binning block; binning blocks
  uint32_t* H;             // nch x MC_BIN_BLOCKS, chunk-major
  uint64_t* cstart;        // nch + 1: (first segment << 32) | first sorted draw
  int* segmap;             // segments: their chunk
  double* st;              // n sorted targets
  int* sj;                 // n (nullable): their draws
  int* as;                 // nullable: the ancestor of each draw
  int* os;                 // nullable: offspring counts
};

__device__ __forceinline__ double mc_target(const MCArgs& a, double W, long j) {
  const uint32_t step = a.gate != nullptr ? (uint32_t)a.gate->k : a.step;
  double U;
  if (a.U != nullptr) {
    U = a.U[j];
  } else {
    const u32x4 r = philox_block(a.seed, (uint32_t)j, step, TAG_MULTINOMIAL, a.batch, 0);
    U = u53(r.x, r.y);
  }
  return W * U;
}

// the chunks' last Ws into LDS
__device__ __forceinline__ void mc_fences(const MCArgs& a, double* fence) {
  for (int c = threadIdx.x; c < a.nch; c += MC_BLOCK) {
    const long last = (long)(c + 1) * MC_CHUNK < a.P ? (long)(c + 1) * MC_CHUNK - 1 : a.P - 1;
    fence[c] = a.Ws[last];
  }
}

// chunk of target t: the number of fences <= t, the last chunk for t >= W
__device__ __forceinline__ int mc_chunk(const double* fence, int nch, double t) {
  int f = 0, hi = nch;
  while (f < hi) {
    const int mid = f + ((hi - f) >> 1);
    if (fence[mid] <= t) f = mid + 1;
    else hi = mid;
  }
  return f < nch ? f : nch - 1;
}

__global__ __launch_bounds__(MC_BLOCK) void k_mc_bin(MCArgs a) {
  __shared__ double fence[MC_MAX_CHUNKS];
  __shared__ uint32_t cnt[MC_MAX_CHUNKS];
  if (!gate_open(a.gate)) return;
  const double W = a.Ws[a.P - 1];
  if (!(W > 0.0)) {
    if (a.status != nullptr && blockIdx.x == 0 && threadIdx.x == 0)
      a.status->status = BIHIP_DEGENERATE;
    return;
  }
  mc_fences(a, fence);
  for (int c = threadIdx.x; c < a.nch; c += MC_BLOCK) cnt[c] = 0;
  if (a.os != nullptr)
    for (long i = (long)blockIdx.x * MC_BLOCK + threadIdx.x; i < a.P;
         i += (long)gridDim.x * MC_BLOCK)
      a.os[i] = 0;
  __syncthreads();
  const long j0 = (long)blockIdx.x * a.per_block;
  const long j1 = j0 + a.per_block < a.n ? j0 + a.per_block : a.n;
  for (long j = j0 + threadIdx.x; j < j1; j += MC_BLOCK)
    atomicAdd(&cnt[mc_chunk(fence, a.nch, mc_target(a, W, j))], 1u);
  __syncthreads();
  for (int c = threadIdx.x; c < a.nch; c += MC_BLOCK)
    a.H[(size_t)c * MC_BIN_BLOCKS + blockIdx.x] = cnt[c];
}

// one block of SCAN_BLOCK (>= MC_BIN_BLOCKS) threads per chunk
__global__ __launch_bounds__(SCAN_BLOCK) void k_mc_offsets(MCArgs a) {
  __shared__ uint64_t lds[SCAN_BLOCK / WAVE];
  if (!gate_open(a.gate) || !(a.Ws[a.P - 1] > 0.0)) return;
  const int c = blockIdx.x, b = threadIdx.x;
  uint32_t* col = a.H + (size_t)c * MC_BIN_BLOCKS;
  const uint64_t v = b < a.nb ? col[b] : 0u;
  uint64_t total;
  const uint64_t ex = block_exclusive_scan(v, lds, total);
  if (b < a.nb) col[b] = (uint32_t)ex;
  if (b == 0) {
    const uint64_t segs = total > 0 ? (total + MC_SEG - 1) / MC_SEG : 1;
    a.cstart[c] = (segs << 32) | total;
  }
}

__global__ __launch_bounds__(MC_BLOCK) void k_mc_scatter(MCArgs a) {
  __shared__ double fence[MC_MAX_CHUNKS];
  __shared__ uint32_t cur[MC_MAX_CHUNKS];
  if (!gate_open(a.gate)) return;
  const double W = a.Ws[a.P - 1];
  if (!(W > 0.0)) return;
  mc_fences(a, fence);
  for (int c = threadIdx.x; c < a.nch; c += MC_BLOCK)
    cur[c] = (uint32_t)a.cstart[c] + a.H[(size_t)c * MC_BIN_BLOCKS + blockIdx.x];
  // the segment -> chunk map
  for (long c = (long)blockIdx.x * MC_BLOCK + threadIdx.x; c < a.nch;
       c += (long)gridDim.x * MC_BLOCK) {
    const uint32_t s0 = (uint32_t)(a.cstart[c] >> 32), s1 = (uint32_t)(a.cstart[c + 1] >> 32);
    for (uint32_t q = s0; q < s1; ++q) a.segmap[q] = (int)c;
  }
  __syncthreads();
  const long j0 = (long)blockIdx.x * a.per_block;
  const long j1 = j0 + a.per_block < a.n ? j0 + a.per_block : a.n;
  for (long j = j0 + threadIdx.x; j < j1; j += MC_BLOCK) {
    const double t = mc_target(a, W, j);
    const uint32_t pos = atomicAdd(&cur[mc_chunk(fence, a.nch, t)], 1u);
    a.st[pos] = t;
    if (a.sj != nullptr) a.sj[pos] = (int)j;
  }
}

__global__ __launch_bounds__(MC_BLOCK) void k_mc_search(MCArgs a) {
  __shared__ double wsl[MC_CHUNK];
  __shared__ uint32_t cnt[MC_CHUNK];
  if (!gate_open(a.gate) || !(a.Ws[a.P - 1] > 0.0)) return;
  const uint32_t nseg = (uint32_t)(a.cstart[a.nch] >> 32);
  if (blockIdx.x >= nseg) return;
  const int c = a.segmap[blockIdx.x];
  const uint64_t w0 = a.cstart[c], w1 = a.cstart[c + 1];
  const uint32_t seg = blockIdx.x - (uint32_t)(w0 >> 32);
  const bool whole = (uint32_t)(w1 >> 32) - (uint32_t)(w0 >> 32) == 1u;
  const long c0 = (long)(uint32_t)w0, c1 = (long)(uint32_t)w1;
  const long lo = c0 + (long)seg * MC_SEG;
  const long hi = lo + MC_SEG < c1 ? lo + MC_SEG : c1;
  const long base = (long)c * MC_CHUNK;
  const int len = base + MC_CHUNK < a.P ? MC_CHUNK : (int)(a.P - base);
  for (int i = threadIdx.x; i < MC_CHUNK; i += MC_BLOCK) {
    wsl[i] = i < len ? a.Ws[base + i] : 0.0;
    cnt[i] = 0;
  }
  __syncthreads();
  for (long pos = lo + threadIdx.x; pos < hi; pos += MC_BLOCK) {
    const double t = a.st[pos];
    int f = 0, h = len;  // the first i with wsl[i] > t
    while (f < h) {
      const int mid = f + ((h - f) >> 1);
      if (wsl[mid] <= t) f = mid + 1;
      else h = mid;
    }
    const int k = base + f < a.P ? f : (int)(a.P - 1 - base);
    if (a.as != nullptr) a.as[a.sj[pos]] = (int)(base + k);
    if (a.os != nullptr) atomicAdd(&cnt[k], 1u);
  }
  if (a.os == nullptr) return;
  __syncthreads();
  for (int i = threadIdx.x; i < len; i += MC_BLOCK) {
    if (whole) a.os[base + i] = (int)cnt[i];
    else if (cnt[i] != 0) atomicAdd(&a.os[base + i], (int)cnt[i]);
  }
}

bool multinomial_chunks_fit(int P) { return P > 0 && ceil_div(P, MC_CHUNK) <= MC_MAX_CHUNKS; }

static long mc_segments_max(int P, int n) {
  return (long)ceil_div(P, MC_CHUNK) + ceil_div(n, MC_SEG) + 1;
}

size_t McWork::bytes(int P, int n, bool draws) {
  const size_t nch = (size_t)ceil_div(P, MC_CHUNK);
  return nch * MC_BIN_BLOCKS * 4 + (nch + 1) * 8 + (size_t)mc_segments_max(P, n) * 4 +
         (size_t)n * 8 + (draws ? (size_t)n * 4 : 0) + 64 * 6;
}

void McWork::carve(void* base, int P, int n, bool draws) {
  const size_t nch = (size_t)ceil_div(P, MC_CHUNK);
  uint8_t* b = static_cast<uint8_t*>(base);
  auto take = [&](size_t bytes) {
    uint8_t* r = b;
    b += (bytes + 63) / 64 * 64;
    return r;
  };
  cstart = reinterpret_cast<uint64_t*>(take((nch + 1) * 8));
  st = reinterpret_cast<double*>(take((size_t)n * 8));
  H = reinterpret_cast<uint32_t*>(take(nch * MC_BIN_BLOCKS * 4));
  segmap = reinterpret_cast<int*>(take((size_t)mc_segments_max(P, n) * 4));
  sj = draws ? reinterpret_cast<int*>(take((size_t)n * 4)) : nullptr;
}

int launch_multinomial_chunks(const MultinomialLaunch& L, const McWork& w, hipStream_t s) {
  if (L.n <= 0) return BIHIP_OK;
  if (!multinomial_chunks_fit(L.P)) return set_error(BIHIP_EINVAL, "too many particles");
  if (L.as != nullptr && w.sj == nullptr)
    return set_error(BIHIP_EINVAL, "the ancestors of the draws need the draw index buffer");
  MCArgs a{};
  a.Ws = L.Ws;
  a.P = L.P;
  a.n = L.n;
  a.U = L.U;
  a.seed = L.seed;
  a.step = L.step;
  a.batch = L.batch;
  a.gate = L.gate;
  a.status = L.status;
  a.nch = ceil_div(L.P, MC_CHUNK);
  a.nb = std::min(MC_BIN_BLOCKS, ceil_div(L.n, MC_BLOCK * 4));
  a.per_block = ceil_div(ceil_div(L.n, a.nb), MC_BLOCK) * MC_BLOCK;
  a.nb = ceil_div(L.n, a.per_block);
  a.H = w.H;
  a.cstart = w.cstart;
  a.segmap = w.segmap;
  a.st = w.st;
  a.sj = L.as != nullptr ? w.sj : nullptr;
  a.as = L.as;
  a.os = L.os;
  hipLaunchKernelGGL(k_mc_bin, dim3(a.nb), dim3(MC_BLOCK), 0, s, a);
  BIHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_mc_offsets, dim3(a.nch), dim3(SCAN_BLOCK), 0, s, a);
  BIHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_tiles_exclusive<uint64_t>, dim3(1), dim3(SCAN_BLOCK), 0, s, a.cstart,
                     a.nch, a.cstart + a.nch, L.gate, BankStrides{});
  BIHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_mc_scatter, dim3(a.nb), dim3(MC_BLOCK), 0, s, a);
  BIHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_mc_search, dim3((unsigned)mc_segments_max(L.P, L.n)), dim3(MC_BLOCK), 0,
                     s, a);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

__global__ void k_zero_ints(int* p, int n, const PFDevice* gate) {
  if (!gate_open(gate)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

int launch_zero_ints(int* p, int n, const PFDevice* gate, hipStream_t s) {
  if (n <= 0) return BIHIP_OK;
  hipLaunchKernelGGL(k_zero_ints, dim3(ceil_div(n, 256)), dim3(256), 0, s, p, n, gate);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

__global__ void k_histogram(const int* as, int n, int P, int* os) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int a = as[j];
  if (a >= 0 && a < P) atomicAdd(&os[a], 1);
}

int launch_ancestors_to_offspring(const int* as, int n, int P, int* os, hipStream_t s) {
  if (n <= 0) return BIHIP_OK;
  hipLaunchKernelGGL(k_histogram, dim3(ceil_div(n, 256)), dim3(256), 0, s, as, n, P, os);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

// ---------------------------------------------------------------- gather
template <class E>
__global__ void k_gather_rows(const int* as, int P, int ncols, E* X, int ld) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= P || j >= ncols) return;
  const int src = as[i];
  if (src != i && src >= 0 && src < P) X[(size_t)j * ld + i] = X[(size_t)j * ld + src];
}

int launch_gather_rows(const int* as, int P, int ncols, void* X, int ld, int elem,
                       hipStream_t s) {
  if (P <= 0 || ncols <= 0) return BIHIP_OK;
  const dim3 grid(ceil_div(P, 256), ncols);
  if (elem == 8) {
    hipLaunchKernelGGL(k_gather_rows<double>, grid, dim3(256), 0, s, as, P, ncols,
                       static_cast<double*>(X), ld);
  } else if (elem == 4) {
    hipLaunchKernelGGL(k_gather_rows<float>, grid, dim3(256), 0, s, as, P, ncols,
                       static_cast<float*>(X), ld);
  } else {
    return set_error(BIHIP_EINVAL, "gather_rows: element size must be 4 or 8");
  }
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

// Ancestry output of the step: as (resampled) or the identity.  Filter
// blockIdx.y of a bank: its ancestors at y * bp, PFDevice dev + y, history
// at out + y * hs.
__global__ void k_record_ancestors(const int* as, int P, const PFDevice* dev, int* out, long bp,
                                   long hs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const long f = blockIdx.y;
  as += f * bp;
  dev += f;
  out += f * hs;
  const bool r = dev->resample != 0 && dev->status == BIHIP_OK;
  out[(size_t)dev->k * P + i] = r ? as[i] : i;
}

int launch_record_ancestors(const int* as, int P, const PFDevice* gate, int* out,
                            hipStream_t s, int banks, long bp, long hs) {
  hipLaunchKernelGGL(k_record_ancestors, dim3(ceil_div(P, 256), banks > 1 ? banks : 1), dim3(256),
                     0, s, as, P, gate, out, bp, hs);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

// Whole-filter copies between the filters of banks: filter j of dst takes
// filter from[j] of src (from null: filter j) where mask is null or
// mask[j] != 0; n16 16-byte words per filter, filter j on grid row j.
__global__ void k_filter_copy(const uint4* src, uint4* dst, long n16, const int* from,
                              const int* mask) {
  const long j = blockIdx.y;
  if (mask != nullptr && mask[j] == 0) return;
  const long f = from != nullptr ? from[j] : j;
  const uint4* s = src + f * n16;
  uint4* d = dst + j * n16;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
       i += (long)gridDim.x * blockDim.x)
    d[i] = s[i];
}

int launch_filter_copy(const void* src, void* dst, long bytes, int banks, const int* from,
                       const int* mask, hipStream_t s) {
  if (bytes % 16 != 0) return set_error(BIHIP_EINVAL, "filter copies move 16-byte words");
  const long n16 = bytes / 16;
  const int gx = (int)std::min<long>(ceil_div(n16, 256L), 512L);
  hipLaunchKernelGGL(k_filter_copy, dim3(gx > 0 ? gx : 1, banks), dim3(256), 0, s,
                     static_cast<const uint4*>(src), static_cast<uint4*>(dst), n16, from, mask);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

// The state after the step into slot k + 1 of the history (slot 0 is the
// initial state): (T + 1) x D columns of leading dimension ld.
template <class T>
__global__ void k_record_state(const T* X, int ld, int D, int P, const PFDevice* dev, T* hist,
                               int slot0) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long)P * D) return;
  const int c = (int)(g / P), p = (int)(g % P);
  const int slot = dev != nullptr ? dev->k + 1 : slot0;
  hist[((size_t)slot * D + c) * ld + p] = X[(size_t)c * ld + p];
}

int launch_record_state(const void* X, int ld, int D, int P, int elem, const PFDevice* dev,
                        void* hist, int slot0, hipStream_t s) {
  const dim3 grid(ceil_div((long)P * D, 256)), block(256);
  if (elem == 8)
    hipLaunchKernelGGL(k_record_state<double>, grid, block, 0, s, static_cast<const double*>(X), ld,
                       D, P, dev, static_cast<double*>(hist), slot0);
  else
    hipLaunchKernelGGL(k_record_state<float>, grid, block, 0, s, static_cast<const float*>(X), ld,
                       D, P, dev, static_cast<float*>(hist), slot0);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

// Paths of particles idx[q] at the current time K, traced back through the
// recorded ancestors: out[q][k][c] = state c at time k of the lineage
// (AncestryCache::readPath, share/src/bi/cache/AncestryCache.hpp:328-352).
template <class T>
__global__ void k_paths(const int* anc, const T* hist, int ld, int D, int P, int K, const int* idx,
                        int n, double* out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  int p = idx[q];
  for (int k = K; k >= 0; --k) {
    for (int c = 0; c < D; ++c)
      out[((size_t)q * (K + 1) + k) * D + c] = (double)hist[((size_t)k * D + c) * ld + p];
    if (k > 0) p = anc[(size_t)(k - 1) * P + p];
  }
}

int launch_paths(const int* anc, const void* hist, int ld, int D, int P, int K, int elem,
                 const int* idx, int n, double* out, hipStream_t s) {
  if (n <= 0) return BIHIP_OK;
  const dim3 grid(ceil_div(n, 64)), block(64);
  if (elem == 8)
    hipLaunchKernelGGL(k_paths<double>, grid, block, 0, s, anc, static_cast<const double*>(hist), ld,
                       D, P, K, idx, n, out);
  else
    hipLaunchKernelGGL(k_paths<float>, grid, block, 0, s, anc, static_cast<const float*>(hist), ld,
                       D, P, K, idx, n, out);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

__global__ void k_fill_f64(double* x, int n, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

int launch_fill_f64(double* x, int n, double v, hipStream_t s) {
  if (n <= 0) return BIHIP_OK;
  hipLaunchKernelGGL(k_fill_f64, dim3(ceil_div(n, 256)), dim3(256), 0, s, x, n, v);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

__global__ void k_iota(int* x, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = i;
}

int launch_iota(int* x, int n, hipStream_t s) {
  if (n <= 0) return BIHIP_OK;
  hipLaunchKernelGGL(k_iota, dim3(ceil_div(n, 256)), dim3(256), 0, s, x, n);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

}  // namespace bihip
