// Scans used by the resamplers: the FP64 inclusive scan of exp(lw - max)
// (sumexpu_inclusive_scan, share/src/bi/primitive/vector_primitive.hpp:1075-1084)
// and the integer offspring scan.
#include "scan.hpp"

namespace bihip {

int scan_tiles(int P) { return ceil_div(P, SCAN_TILE); }

int scan_resident_tiles() {
  static int limit = -1;
  if (limit < 0) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      cus = 0;
    limit = 2 * cus;
  }
  return limit;
}

// Weights w = nanexp(lw - max) lie in [0, 1]; they are summed exactly in
// 128-bit fixed point (fix_weight) and each prefix is rounded to double once
// (fix_to_double).  Exact integer sums make the scan deterministic and
// monotone: a floating-point parallel scan rounds differently across tile
// boundaries and can step backwards, which would give a particle negative
// offspring.
struct SumExpLoader {
  const double* lws;
  const double* mx;
  __device__ u128 operator()(long i) const { return fix_weight(nanexp(lws[i] - *mx)); }
};

struct StoreInclusive {
  double* out;
  __device__ void operator()(long i, u128 incl, u128) const { out[i] = fix_to_double(incl); }
};

int launch_sumexp_scan(const double* lws, int P, const double* mx_dev, const PFDevice* gate,
                       double* Ws, ScanWorkspace ws, hipStream_t s) {
  return launch_tile_scan<u128>(SumExpLoader{lws, mx_dev}, StoreInclusive{Ws}, PFGate{gate}, P,
                                ws, s);
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
  return launch_tile_scan<uint64_t>(CountLoader{os}, StoreCountInclusive{Os}, AlwaysGate{}, P, ws,
                                    s);
}

}  // namespace bihip
