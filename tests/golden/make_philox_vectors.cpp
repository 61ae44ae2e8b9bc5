// Writes tests/golden/philox4x32_10_rocrand.json: Philox4x32-10 blocks
// computed by rocRAND's own engine (rocrand_philox4x32_10.h, AMD open
// source), as an independent check of the oracle's and the kernels' Philox.
// Build and run on the CPU:  tests/golden/make_philox_vectors.sh
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
#include <cstdint>

// ten_rounds is a protected member of the engine: expose it.
struct Engine : rocrand_device::philox4x32_10_engine {
  using rocrand_device::philox4x32_10_engine::ten_rounds;
};

int main() {
  Engine eng;
  const uint32_t keys[][2] = {{0u, 0u}, {0xffffffffu, 0xffffffffu}, {0xa4093822u, 0x299f31d0u},
                              {12345u, 0u}, {0x9e3779b9u, 0x1u}};
  const uint32_t ctrs[][4] = {{0u, 0u, 0u, 0u},
                              {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
                              {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u},
                              {7u, 1u, 2u | (3u << 8), 0u},
                              {1048575u, 500u, 5u, 9u},
                              {42u, 0u, 1u, 2u}};
  printf("[\n");
  bool first = true;
  for (auto& k : keys) {
    for (auto& c : ctrs) {
      uint4 r = eng.ten_rounds(uint4{c[0], c[1], c[2], c[3]}, uint2{k[0], k[1]});
      printf("%s  {\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}",
             first ? "" : ",\n", c[0], c[1], c[2], c[3], k[0], k[1], r.x, r.y, r.z, r.w);
      first = false;
    }
  }
  printf("\n]\n");
  return 0;
}
