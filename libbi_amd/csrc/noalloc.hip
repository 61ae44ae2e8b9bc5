// Allocation-free variant of the permuted-ancestor entry point
// (bihip_cumulative_offspring_to_ancestors_permute_noalloc).
//
// The allocating entry point grows the context's scratch with hipMalloc on
// demand: the (e, f) tile sums, the free-slot count and the E / free-slot
// table of the placement (8 bytes per particle).  Callers that cannot
// allocate (code under HIP graph capture, serving loops, harnesses that
// forbid hipMalloc) use this one instead: its scratch is the file-scope
// device array `scratch` below, one copy per GPU, held until the process
// ends.  The kernels are the allocating entry point's (k_ef_sums,
// k_place_tiles, k_place_free), so the two give identical ancestors.
//
// The array's name is a host-side stand-in, not a device address: each
// device's address is looked up once with hipGetSymbolAddress, with that
// device current, when a context is created on it (noalloc_scratch_init),
// and only that address is ever handed to a kernel.
//
// Every call on one GPU shares that GPU's array and leaves its contents
// behind; every word a call reads was written earlier in the same call (the
// tile sums by k_ef_sums, the free-slot count by k_place_tiles, the E table
// and slot list by k_place_tiles), so nothing needs clearing between calls.
// Calls that overlap on one GPU (two streams, two host threads, two
// contexts on the same device) race on the array and must be serialised by
// the caller.
#include <mutex>

#include "internal.hpp"

namespace bihip {

namespace {

// Layout of the scratch array: the scan-state word (ScanState, 16 bytes,
// padded to 256), the tile words of a ScanWorkspace of NOALLOC_TILES tiles,
// then the placement table of 2 x 32 bits per particle.
constexpr int NOALLOC_MAX_P = BIHIP_NOALLOC_MAX_P;
constexpr int NOALLOC_TILES = (NOALLOC_MAX_P + SCAN_TILE - 1) / SCAN_TILE;
constexpr size_t NOALLOC_STATE_BYTES = 256;
constexpr size_t NOALLOC_WORDS_BYTES = sizeof(uint64_t) * SCAN_WORDS * NOALLOC_TILES;
constexpr size_t NOALLOC_EF_BYTES = sizeof(uint64_t) * (size_t)NOALLOC_MAX_P;
constexpr size_t NOALLOC_BYTES = NOALLOC_STATE_BYTES + NOALLOC_WORDS_BYTES + NOALLOC_EF_BYTES;
static_assert(sizeof(ScanState) <= NOALLOC_STATE_BYTES, "scan state does not fit");
static_assert(NOALLOC_WORDS_BYTES % 256 == 0, "tile words must keep the table aligned");

__device__ __align__(256) uint8_t scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_mutex;
void* g_scratch[NOALLOC_MAX_DEVICES] = {};  // device address of `scratch` per device

}  // namespace

// Looks up this device's copy of the scratch array once; the device must be
// current.  Called by bihip_ctx_create.
int noalloc_scratch_init(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES)
    return set_error(BIHIP_EINVAL, "device %d beyond the %d devices of the static scratch", device,
                     NOALLOC_MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_scratch[device] != nullptr) return BIHIP_OK;
  void* p = nullptr;
  BIHIP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (p == nullptr) return set_error(BIHIP_EHIP, "no device address for the static scratch");
  g_scratch[device] = p;
  return BIHIP_OK;
}

static void* noalloc_scratch(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(g_mutex);
  return g_scratch[device];
}

}  // namespace bihip

using namespace bihip;

extern "C" int bihip_cumulative_offspring_to_ancestors_permute_noalloc(bihip_ctx* ctx,
                                                                       const int* Os, int P,
                                                                       int* as) {
  int rc = ctx_enter_device(ctx);
  if (rc) return rc;
  if (P <= 0) return set_error(BIHIP_EINVAL, "P must be positive");
  if (P > NOALLOC_MAX_P)
    return set_error(BIHIP_EINVAL,
                     "P = %d exceeds the %d particles of the static scratch; use "
                     "bihip_cumulative_offspring_to_ancestors_permute",
                     P, NOALLOC_MAX_P);
  if (Os == nullptr || as == nullptr) return set_error(BIHIP_EINVAL, "NULL Os or as");
  int device = -1;
  BIHIP_HIP(hipGetDevice(&device));
  uint8_t* base = static_cast<uint8_t*>(noalloc_scratch(device));
  if (base == nullptr)
    return set_error(BIHIP_ESTATE, "no static scratch on device %d (no context created there)",
                     device);
  ScanWorkspace ws{};
  ws.state = reinterpret_cast<ScanState*>(base);
  ws.words = reinterpret_cast<uint64_t*>(base + NOALLOC_STATE_BYTES);
  ws.max_tiles = NOALLOC_TILES;
  uint64_t* EF = reinterpret_cast<uint64_t*>(base + NOALLOC_STATE_BYTES + NOALLOC_WORDS_BYTES);
  return launch_ancestors_permute(Os, true, P, as, EF, ws, nullptr, ctx_stream_of(ctx));
}
