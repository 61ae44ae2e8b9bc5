// Pruned ancestry tree on the device: the filter's record = 3.
//
// Replaces AncestryCache<ON_DEVICE> (share/src/bi/cache/AncestryCache.hpp:58-470)
// with its prune kernel (share/src/bi/cuda/cache/AncestryCacheKernel.cuh:29-52):
// every generation of particles becomes a generation of nodes in a tree of
// node slots, each node holding its state, the slot of its parent (-1 for the
// first generation) and its number of surviving children.  After a step,
// every leaf (a particle of the previous generation) learns its number of
// children from the step's ancestors; a leaf without children is freed, and
// so is every ancestor whose last surviving child that was, walking up the
// tree.  The new generation then takes free slots (those with no surviving
// children) in ascending slot order.  Only lineages that reach the current
// generation stay, so the cache holds a coalescent tree whose size is set by
// how far back the particles' lineages stay distinct, not by the number of
// steps: histories far beyond HBM as full records fit in it.
//
// Layout: node states Xs column-major with leading dimension cap (the node
// capacity), so one generation of P nodes is written with coalesced stores;
// node_as, node_os: int[cap]; leaves ls: int[P] (the slot of each current
// particle).  Slots of the new generation are found by a deterministic
// compaction of the free slots (tile counts, one-block scan of the tiles,
// then the first P free slots by rank), so the layout does not depend on
// thread timing.  Every kernel reads the step's resampling decision on the
// device (the filter's gate, or an imposed history).
#include "internal.hpp"

namespace bihip {

namespace {

constexpr int AC_BLOCK = 256;
constexpr int AC_TILE = AC_BLOCK * 8;  // slots per compaction tile

__device__ __forceinline__ bool step_resampled(const PFDevice* dev, const int* imp_r) {
  if (imp_r != nullptr) return imp_r[dev->k] != 0;
  return dev->resample != 0 && dev->status == BIHIP_OK;
}

// First generation: nodes 0..P-1 are the particles, roots of the tree.
template <class T>
__global__ void k_cache_roots(const T* X, int ld, int D, int P, T* Xs, long cap, int* node_as,
                              int* node_os, int* ls, int* live) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < (long)P * D) {
    const int c = (int)(g / P), p = (int)(g % P);
    Xs[(size_t)c * cap + p] = X[(size_t)c * ld + p];
  }
  if (g < P) {
    node_as[g] = -1;
    node_os[g] = 0;
    ls[g] = (int)g;
  }
  if (g == 0) *live = P;
}

// Children of every current leaf under the step's ancestors (one each
// without a resample).
__global__ void k_cache_children(const int* as, const PFDevice* dev, const int* imp_r, int P,
                                 int* cnt) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= P) return;
  if (step_resampled(dev, imp_r)) {
    atomicAdd(&cnt[as[j]], 1);
  } else {
    cnt[j] = 1;
  }
}

// Each leaf's children into its node; a leaf without children is freed, and
// the walk continues up through every node whose last surviving child it
// was (AncestryCacheKernel.cuh:29-52: the thread whose decrement reaches zero
// owns the parent).  *live drops by the nodes freed.
__global__ void k_cache_prune(const int* ls, const int* cnt, int P, const int* node_as,
                              int* node_os, int* live) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int removed = 0;
  if (i < P) {
    int node = ls[i];
    const int c = cnt[i];
    node_os[node] = c;
    if (c == 0) {
      while (true) {
        ++removed;
        const int parent = node_as[node];
        if (parent < 0) break;
        if (atomicSub(&node_os[parent], 1) != 1) break;  // other children survive
        node = parent;
      }
    }
  }
  // one atomic per wave
  for (int off = 32; off > 0; off >>= 1) removed += __shfl_xor(removed, off, 64);
  if ((threadIdx.x & 63) == 0 && removed != 0) atomicSub(live, removed);
}

// Free slots (no surviving children) per tile of AC_TILE slots.
__global__ __launch_bounds__(AC_BLOCK) void k_cache_free_tiles(const int* node_os, long cap,
                                                               int* tile_free) {
  __shared__ int part[AC_BLOCK / 64];
  const long t0 = (long)blockIdx.x * AC_TILE;
  int n = 0;
#pragma unroll
  for (int u = 0; u < AC_TILE / AC_BLOCK; ++u) {
    const long s = t0 + (long)u * AC_BLOCK + threadIdx.x;
    if (s < cap && node_os[s] == 0) ++n;
  }
  for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = n;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < AC_BLOCK / 64; ++w) t += part[w];
    tile_free[blockIdx.x] = t;
  }
}

// One block: exclusive scan of the tile counts in place; *total gets the sum.
__global__ __launch_bounds__(1024) void k_cache_scan_tiles(int* tile_free, int ntiles,
                                                           long* total) {
  __shared__ long sums[1024];
  const int per = (ntiles + 1023) / 1024;
  const int b = threadIdx.x * per, e = min(b + per, ntiles);
  long s = 0;
  for (int t = b; t < e; ++t) s += tile_free[t];
  sums[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const long v = threadIdx.x >= off ? sums[threadIdx.x - off] : 0;
    __syncthreads();
    sums[threadIdx.x] += v;
    __syncthreads();
  }
  long run = threadIdx.x > 0 ? sums[threadIdx.x - 1] : 0;
  for (int t = b; t < e; ++t) {
    const int c = tile_free[t];
    tile_free[t] = (int)run;
    run += c;
  }
  if (threadIdx.x == 1023) *total = sums[1023];
}

// The free slots of rank < P, in ascending slot order: slots[rank] = slot.
__global__ __launch_bounds__(AC_BLOCK) void k_cache_free_list(const int* node_os, long cap,
                                                              const int* tile_off, int P,
                                                              int* slots) {
  __shared__ int wave_off[AC_BLOCK / 64 + 1];
  const long t0 = (long)blockIdx.x * AC_TILE;
  int base = tile_off[blockIdx.x];
  if (base >= P) return;  // the whole block: no slot of this tile is needed
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int u = 0; u < AC_TILE / AC_BLOCK; ++u) {
    const long s = t0 + (long)u * AC_BLOCK + threadIdx.x;
    const bool f = s < cap && node_os[s] == 0;
    const uint64_t m = __ballot(f);
    const int below = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wave_off[w] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
      int acc = 0;
      for (int q = 0; q < AC_BLOCK / 64; ++q) {
        const int c = wave_off[q];
        wave_off[q] = acc;
        acc += c;
      }
      wave_off[AC_BLOCK / 64] = acc;
    }
    __syncthreads();
    const int r = base + wave_off[w] + below;
    if (f && r < P) slots[r] = (int)s;
    base += wave_off[AC_BLOCK / 64];
    __syncthreads();
  }
}

// The new generation into its slots: state, parent (the slot of the
// ancestor's leaf), no children yet; its slots become the leaves.
template <class T>
__global__ void k_cache_insert(const T* X, int ld, int D, int P, const int* as,
                               const PFDevice* dev, const int* imp_r, const int* ls_old,
                               const int* slots, const long* nfree, T* Xs, long cap,
                               int* node_as, int* node_os, int* ls_new, int* live) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (*nfree < P) return;  // the host keeps room for a generation; never taken
  if (g < (long)P * D) {
    const int c = (int)(g / P), j = (int)(g % P);
    Xs[(size_t)c * cap + slots[j]] = X[(size_t)c * ld + j];
  }
  if (g < P) {
    const int j = (int)g;
    const int a = step_resampled(dev, imp_r) ? as[j] : j;
    const int s = slots[j];
    node_as[s] = ls_old[a];
    node_os[s] = 0;
    ls_new[j] = s;
  }
  if (g == 0) atomicAdd(live, P);
}

// Paths of current particles idx[q]: the states of the nodes from its leaf up
// to the root (AncestryCache::readPath, AncestryCache.hpp:328-352).
template <class T>
__global__ void k_cache_paths(const int* ls, const T* Xs, long cap, const int* node_as, int D,
                              int K, const int* idx, int n, double* out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  int node = ls[idx[q]];
  for (int k = K; k >= 0 && node >= 0; --k) {
    for (int c = 0; c < D; ++c)
      out[((size_t)q * (K + 1) + k) * D + c] = (double)Xs[(size_t)c * cap + node];
    node = node_as[node];
  }
}

__global__ void k_fill_int(int* x, long n, int v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

int fill_int(int* x, long n, int v, hipStream_t s) {
  if (n <= 0) return BIHIP_OK;
  hipLaunchKernelGGL(k_fill_int, dim3(ceil_div(n, 256)), dim3(256), 0, s, x, n, v);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

}  // namespace

int ancestry_alloc(AncestryCache* c, int P, int D, int elem, long cap, hipStream_t s) {
  *c = AncestryCache{};
  c->P = P;
  c->D = D;
  c->elem = elem;
  c->cap = cap;
  int rc;
  uint8_t* xs = nullptr;
  if ((rc = dalloc(&xs, (size_t)cap * D * elem))) return rc;
  c->Xs = xs;
  if ((rc = dalloc(&c->node_as, (size_t)cap)) || (rc = dalloc(&c->node_os, (size_t)cap)) ||
      (rc = dalloc(&c->ls[0], (size_t)P)) || (rc = dalloc(&c->ls[1], (size_t)P)) ||
      (rc = dalloc(&c->cnt, (size_t)P)) || (rc = dalloc(&c->slots, (size_t)P)) ||
      (rc = dalloc(&c->tile_free, (size_t)ceil_div(cap, AC_TILE))) ||
      (rc = dalloc(&c->nfree, 1)) || (rc = dalloc(&c->live, 1)))
    return rc;
  return fill_int(c->node_os, cap, 0, s);
}

void ancestry_free(AncestryCache* c) {
  dfree(c->Xs);
  dfree(c->node_as);
  dfree(c->node_os);
  dfree(c->ls[0]);
  dfree(c->ls[1]);
  dfree(c->cnt);
  dfree(c->slots);
  dfree(c->tile_free);
  dfree(c->nfree);
  dfree(c->live);
  *c = AncestryCache{};
}

// A larger node store with the same contents (the reference's enlarge,
// AncestryCache.hpp:371-398); the caller has synchronised the stream.
int ancestry_enlarge(AncestryCache* c, long new_cap, hipStream_t s) {
  AncestryCache n{};
  int rc = ancestry_alloc(&n, c->P, c->D, c->elem, new_cap, s);
  if (rc) {
    ancestry_free(&n);
    return rc;
  }
  const size_t col = (size_t)c->cap * c->elem;
  for (int d = 0; d < c->D; ++d)
    BIHIP_HIP(hipMemcpyAsync(static_cast<uint8_t*>(n.Xs) + (size_t)d * new_cap * c->elem,
                             static_cast<uint8_t*>(c->Xs) + (size_t)d * col, col,
                             hipMemcpyDeviceToDevice, s));
  BIHIP_HIP(hipMemcpyAsync(n.node_as, c->node_as, sizeof(int) * c->cap, hipMemcpyDeviceToDevice, s));
  BIHIP_HIP(hipMemcpyAsync(n.node_os, c->node_os, sizeof(int) * c->cap, hipMemcpyDeviceToDevice, s));
  BIHIP_HIP(hipMemcpyAsync(n.ls[c->cur], c->ls[c->cur], sizeof(int) * c->P,
                           hipMemcpyDeviceToDevice, s));
  BIHIP_HIP(hipMemcpyAsync(n.live, c->live, sizeof(int), hipMemcpyDeviceToDevice, s));
  BIHIP_HIP(hipStreamSynchronize(s));
  n.cur = c->cur;
  n.live_ub = c->live_ub;
  ancestry_free(c);
  *c = n;
  return BIHIP_OK;
}

int ancestry_init(AncestryCache* c, const void* X, int ld, hipStream_t s) {
  const long n = (long)c->P * c->D;
  const dim3 grid(ceil_div(n > c->P ? n : c->P, 256)), block(256);
  int rc = fill_int(c->node_os, c->cap, 0, s);
  if (rc) return rc;
  c->cur = 0;
  if (c->elem == 8)
    hipLaunchKernelGGL(k_cache_roots<double>, grid, block, 0, s, static_cast<const double*>(X), ld,
                       c->D, c->P, static_cast<double*>(c->Xs), c->cap, c->node_as, c->node_os,
                       c->ls[0], c->live);
  else
    hipLaunchKernelGGL(k_cache_roots<float>, grid, block, 0, s, static_cast<const float*>(X), ld,
                       c->D, c->P, static_cast<float*>(c->Xs), c->cap, c->node_as, c->node_os,
                       c->ls[0], c->live);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

int ancestry_write(AncestryCache* c, const void* X, int ld, const int* as, const PFDevice* dev,
                   const int* imp_r, hipStream_t s) {
  const int P = c->P;
  int rc;
  if ((rc = fill_int(c->cnt, P, 0, s))) return rc;
  hipLaunchKernelGGL(k_cache_children, dim3(ceil_div(P, 256)), dim3(256), 0, s, as, dev, imp_r, P,
                     c->cnt);
  BIHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_cache_prune, dim3(ceil_div(P, 256)), dim3(256), 0, s, c->ls[c->cur], c->cnt,
                     P, c->node_as, c->node_os, c->live);
  BIHIP_LAUNCH_CHECK();
  const int nt = ceil_div(c->cap, AC_TILE);
  hipLaunchKernelGGL(k_cache_free_tiles, dim3(nt), dim3(AC_BLOCK), 0, s, c->node_os, c->cap,
                     c->tile_free);
  BIHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_cache_scan_tiles, dim3(1), dim3(1024), 0, s, c->tile_free, nt, c->nfree);
  BIHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_cache_free_list, dim3(nt), dim3(AC_BLOCK), 0, s, c->node_os, c->cap,
                     c->tile_free, P, c->slots);
  BIHIP_LAUNCH_CHECK();
  const long n = (long)P * c->D;
  const dim3 grid(ceil_div(n > P ? n : P, 256)), block(256);
  if (c->elem == 8)
    hipLaunchKernelGGL(k_cache_insert<double>, grid, block, 0, s, static_cast<const double*>(X), ld,
                       c->D, P, as, dev, imp_r, c->ls[c->cur], c->slots, c->nfree,
                       static_cast<double*>(c->Xs), c->cap, c->node_as, c->node_os,
                       c->ls[1 - c->cur], c->live);
  else
    hipLaunchKernelGGL(k_cache_insert<float>, grid, block, 0, s, static_cast<const float*>(X), ld,
                       c->D, P, as, dev, imp_r, c->ls[c->cur], c->slots, c->nfree,
                       static_cast<float*>(c->Xs), c->cap, c->node_as, c->node_os,
                       c->ls[1 - c->cur], c->live);
  BIHIP_LAUNCH_CHECK();
  c->cur = 1 - c->cur;
  return BIHIP_OK;
}

int ancestry_paths(const AncestryCache* c, int K, const int* idx, int n, double* out,
                   hipStream_t s) {
  if (n <= 0) return BIHIP_OK;
  const dim3 grid(ceil_div(n, 64)), block(64);
  if (c->elem == 8)
    hipLaunchKernelGGL(k_cache_paths<double>, grid, block, 0, s, c->ls[c->cur],
                       static_cast<const double*>(c->Xs), c->cap, c->node_as, c->D, K, idx, n, out);
  else
    hipLaunchKernelGGL(k_cache_paths<float>, grid, block, 0, s, c->ls[c->cur],
                       static_cast<const float*>(c->Xs), c->cap, c->node_as, c->D, K, idx, n, out);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

}  // namespace bihip
