// Vector-ALU issue-rate probe (libbiprobe.so, used by bench.py --workload valu).
//
// Measures how many wave-level VALU instructions per second the chip issues
// when nothing else limits a kernel: every lane runs eight independent
// dependency chains (so latency never stalls issue), the grid fills every
// SIMD eight waves deep, and there are no memory accesses inside the loop.
// The result is the ceiling against which a VALU-bound kernel's
// SQ_INSTS_VALU rate is read; the same kernel, with its known instruction
// count, also calibrates that counter.
//
//   kind 0: v_fma_f32            (the DOPRI5 stages' arithmetic)
//   kind 1: v_xor_b32 + v_add_u32 (the Philox rounds' integer work)
//   kind 2: v_mul_f32 + v_add_f32 alternating (stage sums and norms)
// Each loop trip of a lane issues VALU_PER_TRIP instructions of its kind.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

constexpr int PROBE_BLOCK = 256;
constexpr int CHAINS = 8;
constexpr int UNROLL = 8;
constexpr int VALU_PER_TRIP = CHAINS * UNROLL;  // kinds 0 (one op per chain and unroll step)

template <int KIND>
__global__ __launch_bounds__(PROBE_BLOCK) void k_valu_probe(float* out, int trips, float b,
                                                            float c) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * PROBE_BLOCK + threadIdx.x;
  if constexpr (KIND == 0 || KIND == 2) {
    float a[CHAINS];
#pragma unroll
    for (int j = 0; j < CHAINS; ++j) a[j] = (float)t * 1e-7f + (float)j;
    for (int i = 0; i < trips; ++i) {
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
#pragma unroll
        for (int j = 0; j < CHAINS; ++j) {
          if constexpr (KIND == 0)
            a[j] = __builtin_fmaf(a[j], b, c);
          else if ((u & 1) == 0)
            a[j] = a[j] * b;
          else
            a[j] = a[j] + c;
        }
      }
    }
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < CHAINS; ++j) s += a[j];
    out[t] = s;
  } else {
    uint32_t a[CHAINS];
    const uint32_t k = __float_as_uint(b);
#pragma unroll
    for (int j = 0; j < CHAINS; ++j) a[j] = (uint32_t)t * 2654435761u + (uint32_t)j;
    for (int i = 0; i < trips; ++i) {
#pragma unroll
      for (int u = 0; u < UNROLL / 2; ++u) {
#pragma unroll
        for (int j = 0; j < CHAINS; ++j) a[j] = (a[j] ^ k) + 0x9E3779B9u;
      }
    }
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < CHAINS; ++j) s ^= a[j];
    out[t] = (float)s;
  }
}

}  // namespace

// Launches blocks x 256 lanes running `trips` loop trips of instruction kind
// `kind` on the null stream; *ms gets the kernel time by HIP events and
// *wave_instr the VALU instructions the loop issues (waves x trips x
// VALU_PER_TRIP).  out_dev holds blocks x 256 floats.  Returns 0 or a HIP
// error code.
extern "C" int biprobe_valu(int kind, int blocks, int trips, float* out_dev, float* ms,
                            double* wave_instr) {
  if (kind < 0 || kind > 2 || blocks <= 0 || trips <= 0 || out_dev == nullptr) return -1;
  hipEvent_t e0, e1;
  hipError_t e;
  if ((e = hipEventCreate(&e0)) != hipSuccess) return (int)e;
  if ((e = hipEventCreate(&e1)) != hipSuccess) return (int)e;
  if ((e = hipEventRecord(e0, nullptr)) != hipSuccess) return (int)e;
  switch (kind) {
    case 0:
      hipLaunchKernelGGL(k_valu_probe<0>, dim3(blocks), dim3(PROBE_BLOCK), 0, nullptr, out_dev,
                         trips, 0.999f, 0.001f);
      break;
    case 1:
      hipLaunchKernelGGL(k_valu_probe<1>, dim3(blocks), dim3(PROBE_BLOCK), 0, nullptr, out_dev,
                         trips, 0.999f, 0.001f);
      break;
    default:
      hipLaunchKernelGGL(k_valu_probe<2>, dim3(blocks), dim3(PROBE_BLOCK), 0, nullptr, out_dev,
                         trips, 0.999f, 0.001f);
      break;
  }
  if ((e = hipGetLastError()) != hipSuccess) return (int)e;
  if ((e = hipEventRecord(e1, nullptr)) != hipSuccess) return (int)e;
  if ((e = hipEventSynchronize(e1)) != hipSuccess) return (int)e;
  float t = 0.f;
  if ((e = hipEventElapsedTime(&t, e0, e1)) != hipSuccess) return (int)e;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = t;
  *wave_instr = (double)blocks * (PROBE_BLOCK / 64) * (double)trips * VALU_PER_TRIP;
  return 0;
}
