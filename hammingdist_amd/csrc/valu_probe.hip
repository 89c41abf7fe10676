// VALU issue-rate probe for the roofline of the all-pairs kernels: long
// independent chains of the tile kernel's instruction mix (v_bitop3 with a
// scalar operand, popcount-accumulate) and no memory traffic in the loop.
// Prints one JSON line: lane-ops per second and the shader clock seen by
// s_memtime against the 100 MHz s_memrealtime counter.
//
//   hd_valu_probe [workgroups] [iterations]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                                          \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) {                                                               \
      std::fprintf(stderr, "hd_valu_probe: %s: %s\n", #x, hipGetErrorString(e_));        \
      return 1;                                                                           \
    }                                                                                     \
  } while (0)

constexpr int kChains = 8;
constexpr int kOpsPerStep = 3; // two v_bitop3 and one v_bcnt per chain step

__global__ __launch_bounds__(256) void k_valu(uint32_t seed, uint32_t iters,
                                              uint32_t *__restrict__ out,
                                              unsigned long long *__restrict__ clocks) {
  uint32_t a[kChains], acc[kChains];
  const uint32_t b = __builtin_amdgcn_readfirstlane(seed * 7u + 1u);
  const uint32_t c = __builtin_amdgcn_readfirstlane(seed * 13u + 5u);
#pragma unroll
  for (int k = 0; k < kChains; ++k) {
    a[k] = threadIdx.x * (k + 3) + seed;
    acc[k] = 0;
  }
  const bool timer = blockIdx.x == 0 && threadIdx.x == 0;
  unsigned long long t0 = 0, r0 = 0;
  if (timer) {
    t0 = __builtin_amdgcn_s_memtime();
    r0 = __builtin_amdgcn_s_memrealtime();
  }
  for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
    for (int k = 0; k < kChains; ++k) {
      const uint32_t s = __builtin_amdgcn_bitop3_b32(a[k], b, acc[k], 0x71);
      acc[k] += __builtin_popcount(s);
      a[k] = __builtin_amdgcn_bitop3_b32(a[k], c, s, 0xBE);
    }
  }
  uint32_t r = 0;
#pragma unroll
  for (int k = 0; k < kChains; ++k) r ^= acc[k] + a[k];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = r;
  if (timer) {
    clocks[0] = __builtin_amdgcn_s_memtime() - t0;
    clocks[1] = __builtin_amdgcn_s_memrealtime() - r0;
  }
}

int main(int argc, char **argv) {
  const uint32_t blocks = argc > 1 ? (uint32_t)std::atoi(argv[1]) : 8192;
  const uint32_t iters = argc > 2 ? (uint32_t)std::atoi(argv[2]) : 20000;
  if (blocks == 0 || iters == 0) return 2;
  uint32_t *out = nullptr;
  unsigned long long *clocks = nullptr;
  CHECK(hipMalloc(&out, (size_t)blocks * 256 * sizeof(uint32_t)));
  CHECK(hipMalloc(&clocks, 2 * sizeof(unsigned long long)));
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  double best = 0, ghz = 0;
  for (int rep = 0; rep < 3; ++rep) { // the first launch also warms the clock up
    CHECK(hipEventRecord(e0, nullptr));
    hipLaunchKernelGGL(k_valu, dim3(blocks), dim3(256), 0, nullptr, 1u, iters, out, clocks);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e1, nullptr));
    CHECK(hipEventSynchronize(e1));
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    unsigned long long h[2] = {0, 0};
    CHECK(hipMemcpy(h, clocks, sizeof h, hipMemcpyDeviceToHost));
    const double ops = (double)blocks * 256 * iters * kChains * kOpsPerStep;
    const double rate = ops / (ms * 1e-3);
    if (rate > best) {
      best = rate;
      ghz = h[1] ? (double)h[0] / (double)h[1] * 0.1 : 0.0;
    }
  }
  std::printf("{\"lane_ops_per_s\": %.6e, \"shader_clock_ghz\": %.4f, \"workgroups\": %u, "
              "\"iterations\": %u}\n",
              best, ghz, blocks, iters);
  CHECK(hipFree(out));
  CHECK(hipFree(clocks));
  return 0;
}
