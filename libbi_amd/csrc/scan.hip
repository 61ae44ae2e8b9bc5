// Scans used by the resamplers: the FP64 inclusive scan of exp(lw - max)
// (sumexpu_inclusive_scan, share/src/bi/primitive/vector_primitive.hpp:1075-1084)
// and the integer offspring scan.
#include "scan.hpp"

namespace bihip {

int scan_tiles(int P) { return ceil_div(P, SCAN_TILE); }

struct SumExpLoader {
  const double* lws;
  const double* mx;
  __device__ double operator()(long i) const { return nanexp(lws[i] - *mx); }
};

struct StoreInclusive {
  double* out;
  __device__ void operator()(long i, double incl, double) const { out[i] = incl; }
};

int launch_sumexp_scan(const double* lws, int P, const double* mx_dev, const PFDevice* gate,
                       double* Ws, ScanWorkspace ws, hipStream_t s) {
  const int nt = scan_tiles(P);
  if (nt > ws.max_tiles) return set_error(BIHIP_EINVAL, "scan workspace too small");
  hipLaunchKernelGGL((k_tile_scan<double, SumExpLoader, StoreInclusive, PFGate>), dim3(nt),
                     dim3(SCAN_BLOCK), 0, s, SumExpLoader{lws, mx_dev}, StoreInclusive{Ws},
                     PFGate{gate}, P, ws);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

struct CountLoader {
  const int* os;
  __device__ uint64_t operator()(long i) const { return (uint64_t)(uint32_t)os[i]; }
};

struct StoreCountInclusive {
  int* out;
  __device__ void operator()(long i, uint64_t incl, uint64_t) const { out[i] = (int)incl; }
};

int launch_count_scan(const int* os, int P, int* Os, ScanWorkspace ws, hipStream_t s) {
  const int nt = scan_tiles(P);
  if (nt > ws.max_tiles) return set_error(BIHIP_EINVAL, "scan workspace too small");
  hipLaunchKernelGGL((k_tile_scan<uint64_t, CountLoader, StoreCountInclusive, AlwaysGate>),
                     dim3(nt), dim3(SCAN_BLOCK), 0, s, CountLoader{os}, StoreCountInclusive{Os},
                     AlwaysGate{}, P, ws);
  BIHIP_LAUNCH_CHECK();
  return BIHIP_OK;
}

}  // namespace bihip
