// VALU issue-rate probe for the roofline of the all-pairs kernels.
//
// Each mode runs long independent chains (8 per lane, 256-thread workgroups,
// far more workgroups than the GPU holds) of one instruction mix with no
// memory operation in the loop, and reports lane-ops per second, the shader
// clock (s_memtime against the 100 MHz s_memrealtime counter) and wavefront
// VALU instructions issued per CU per clock:
//
//   bitop3_s   v_bitop3_b32 with one SGPR operand (the tile loop's column word)
//   bitop3_v   v_bitop3_b32 with three VGPR operands
//   bcnt       v_bcnt_u32_b32 (popcount-accumulate)
//   add        v_add_u32 (a plain 2-input VOP2 op, control)
//   fma        v_fma_f32 (unpacked; control against the published FP32 rate)
//   mix31      3 x v_bitop3_b32 (SGPR operand) + 1 x v_bcnt_u32_b32: the dense
//              3-op and gap-3 pair mix (4 ops per pair and 32-site word) with
//              the column word as an SGPR operand
//   mix21      2 x v_bitop3_b32 + 1 x v_bcnt_u32_b32: the gap-2 mix
//   mix41      4 x v_bitop3_b32 + 1 x v_bcnt_u32_b32: the nibble / class-3 mix
//   mix31v     3 x v_bitop3_b32 (VGPR operands only) + 1 x v_bcnt_u32_b32
//   bcast      per column word: 3 x v_mov_b32 from SGPRs, then for 8 rows
//              3 x v_bitop3_b32 (VGPR operands only) + 1 x v_bcnt_u32_b32
//              (35 instructions for 32 algorithmic ops; the rate counts the
//              32): the form of the tile loops, which copy each column word
//              into a VGPR before its rows use it
//   bcast2     per block of 4 column words: 2 x v_mov_b32 from SGPRs per
//              word, then for 2 rows of each word v_xor_b32 with the third
//              word as an SGPR operand, 2 x v_bitop3_b32 (VGPR operands) and
//              v_bcnt_u32_b32, the block's 8 chains independent (40
//              instructions for 32 algorithmic ops): the form of the dense
//              and gap tile loops with 2 rows per lane
//
//   hd_valu_probe [mode|all] [workgroups] [iterations]
//
// Prints one JSON line per mode.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

#define CHECK(x)                                                                          \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) {                                                               \
      std::fprintf(stderr, "hd_valu_probe: %s: %s\n", #x, hipGetErrorString(e_));        \
      return 1;                                                                           \
    }                                                                                     \
  } while (0)

constexpr int kChains = 8;

enum Mode {
  BITOP3_S, BITOP3_V, BCNT, ADD, FMA, MIX31, MIX21, MIX41, MIX31V, BCAST, BCAST2,
  N_MODES
};

struct ModeInfo {
  const char *name;
  int ops; // VALU instructions per chain step
};

constexpr ModeInfo kModes[N_MODES] = {
    {"bitop3_s", 4}, {"bitop3_v", 4}, {"bcnt", 4}, {"add", 4},
    {"fma", 4},      {"mix31", 4},    {"mix21", 3}, {"mix41", 5},
    {"mix31v", 4},   {"bcast", 4},    {"bcast2", 4},
};

__device__ __forceinline__ uint32_t bcnt_acc(uint32_t x, uint32_t acc) {
  return __builtin_popcount(x) + acc;
}

template <int M>
__global__ __launch_bounds__(256) void k_valu(uint32_t seed, uint32_t iters,
                                              uint32_t *__restrict__ out,
                                              unsigned long long *__restrict__ clocks) {
  uint32_t a[kChains], acc[kChains];
  float f[kChains];
  // wave-uniform words held in SGPRs, as the tile loop's column planes are
  const uint32_t b0 = __builtin_amdgcn_readfirstlane(seed * 7u + 1u);
  const uint32_t b1 = __builtin_amdgcn_readfirstlane(seed * 13u + 5u);
  const uint32_t b2 = __builtin_amdgcn_readfirstlane(seed * 29u + 3u);
  const uint32_t b3 = __builtin_amdgcn_readfirstlane(seed * 31u + 9u);
  // per-lane words in VGPRs
  const uint32_t v0 = threadIdx.x * 0x9E3779B9u + seed;
  const uint32_t v1 = threadIdx.x * 0x85EBCA6Bu ^ seed;
  const float fm = 1.0f + 1e-7f * (float)threadIdx.x, fa = 1e-9f * (float)seed;
#pragma unroll
  for (int k = 0; k < kChains; ++k) {
    a[k] = threadIdx.x * (k + 3) + seed;
    acc[k] = k;
    f[k] = (float)(threadIdx.x + k);
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
      if constexpr (M == BITOP3_S) {
        a[k] = __builtin_amdgcn_bitop3_b32(a[k], b0, acc[k], 0x71);
        acc[k] = __builtin_amdgcn_bitop3_b32(acc[k], b1, a[k], 0xBE);
        a[k] = __builtin_amdgcn_bitop3_b32(a[k], b2, acc[k], 0x96);
        acc[k] = __builtin_amdgcn_bitop3_b32(acc[k], b3, a[k], 0x6C);
      } else if constexpr (M == BITOP3_V) {
        a[k] = __builtin_amdgcn_bitop3_b32(a[k], v0, acc[k], 0x71);
        acc[k] = __builtin_amdgcn_bitop3_b32(acc[k], v1, a[k], 0xBE);
        a[k] = __builtin_amdgcn_bitop3_b32(a[k], v0, acc[k], 0x96);
        acc[k] = __builtin_amdgcn_bitop3_b32(acc[k], v1, a[k], 0x6C);
      } else if constexpr (M == BCNT) {
        a[k] = bcnt_acc(a[k], acc[k]);
        acc[k] = bcnt_acc(acc[k], a[k]);
        a[k] = bcnt_acc(a[k], acc[k]);
        acc[k] = bcnt_acc(acc[k], a[k]);
      } else if constexpr (M == ADD) {
        a[k] += acc[k];
        acc[k] += a[k];
        a[k] += acc[k];
        acc[k] += a[k];
      } else if constexpr (M == FMA) {
        f[k] = __builtin_fmaf(f[k], fm, fa);
        f[k] = __builtin_fmaf(f[k], fm, fa);
        f[k] = __builtin_fmaf(f[k], fm, fa);
        f[k] = __builtin_fmaf(f[k], fm, fa);
      } else if constexpr (M == MIX31) {
        uint32_t s = __builtin_amdgcn_bitop3_b32(a[k], b0, acc[k], 0x71);
        s = __builtin_amdgcn_bitop3_b32(s, b1, a[k], 0xBE);
        a[k] = __builtin_amdgcn_bitop3_b32(s, b2, a[k], 0x96);
        acc[k] = bcnt_acc(a[k], acc[k]);
      } else if constexpr (M == MIX21) {
        uint32_t s = __builtin_amdgcn_bitop3_b32(a[k], b0, acc[k], 0x71);
        a[k] = __builtin_amdgcn_bitop3_b32(s, b1, a[k], 0x96);
        acc[k] = bcnt_acc(a[k], acc[k]);
      } else if constexpr (M == MIX41) {
        uint32_t s = __builtin_amdgcn_bitop3_b32(a[k], b0, acc[k], 0x71);
        s = __builtin_amdgcn_bitop3_b32(s, b1, a[k], 0xBE);
        s = __builtin_amdgcn_bitop3_b32(s, b2, a[k], 0x96);
        a[k] = __builtin_amdgcn_bitop3_b32(s, b3, a[k], 0x6C);
        acc[k] = bcnt_acc(a[k], acc[k]);
      } else if constexpr (M == MIX31V) {
        uint32_t s = __builtin_amdgcn_bitop3_b32(a[k], v0, acc[k], 0x71);
        s = __builtin_amdgcn_bitop3_b32(s, v1, a[k], 0xBE);
        a[k] = __builtin_amdgcn_bitop3_b32(s, v0, a[k], 0x96);
        acc[k] = bcnt_acc(a[k], acc[k]);
      } else if constexpr (M == BCAST) {
        if (k == 0) {
          uint32_t m1, m2, m3;
          asm volatile("v_mov_b32 %0, %1" : "=v"(m1) : "s"(b0));
          asm volatile("v_mov_b32 %0, %1" : "=v"(m2) : "s"(b1));
          asm volatile("v_mov_b32 %0, %1" : "=v"(m3) : "s"(b2));
#pragma unroll
          for (int r = 0; r < kChains; ++r) {
            uint32_t t = __builtin_amdgcn_bitop3_b32(a[r], m1, acc[r], 0x71);
            t = __builtin_amdgcn_bitop3_b32(t, m2, a[r], 0xBE);
            a[r] = __builtin_amdgcn_bitop3_b32(t, m3, a[r], 0x96);
            acc[r] = bcnt_acc(a[r], acc[r]);
          }
        }
      } else if constexpr (M == BCAST2) {
        if (k == 0) {
          // a block of 4 column words x 2 rows, its 8 chains independent
          uint32_t m2[kChains / 2], m3[kChains / 2];
#pragma unroll
          for (int c = 0; c < kChains / 2; ++c) {
            asm volatile("v_mov_b32 %0, %1" : "=v"(m2[c]) : "s"(b1));
            asm volatile("v_mov_b32 %0, %1" : "=v"(m3[c]) : "s"(b2));
          }
#pragma unroll
          for (int q = 0; q < kChains; ++q) {
            uint32_t t = a[q] ^ b0; // VOP2 with the SGPR operand
            t = __builtin_amdgcn_bitop3_b32(a[q], m3[q / 2], t, 0x14);
            a[q] = __builtin_amdgcn_bitop3_b32(t, m2[q / 2], a[q], 0x4d);
            acc[q] = bcnt_acc(a[q], acc[q]);
          }
        }
      }
    }
  }
  uint32_t r = 0;
#pragma unroll
  for (int k = 0; k < kChains; ++k) r ^= acc[k] + a[k] + __float_as_uint(f[k]);
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = r;
  if (timer) {
    clocks[0] = __builtin_amdgcn_s_memtime() - t0;
    clocks[1] = __builtin_amdgcn_s_memrealtime() - r0;
  }
}

using KernelFn = void (*)(uint32_t, uint32_t, uint32_t *, unsigned long long *);

static KernelFn kernel_of(int m) {
  switch (m) {
  case BITOP3_S: return k_valu<BITOP3_S>;
  case BITOP3_V: return k_valu<BITOP3_V>;
  case BCNT: return k_valu<BCNT>;
  case ADD: return k_valu<ADD>;
  case FMA: return k_valu<FMA>;
  case MIX31: return k_valu<MIX31>;
  case MIX21: return k_valu<MIX21>;
  case MIX41: return k_valu<MIX41>;
  case MIX31V: return k_valu<MIX31V>;
  case BCAST2: return k_valu<BCAST2>;
  default: return k_valu<BCAST>;
  }
}

static int run_mode(int m, uint32_t blocks, uint32_t iters, int cus, uint32_t *out,
                    unsigned long long *clocks) {
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  double best = 0, ghz = 0;
  for (int rep = 0; rep < 3; ++rep) { // the first launch also warms the clock up
    CHECK(hipEventRecord(e0, nullptr));
    hipLaunchKernelGGL(kernel_of(m), dim3(blocks), dim3(256), 0, nullptr, 1u, iters, out,
                       clocks);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e1, nullptr));
    CHECK(hipEventSynchronize(e1));
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    unsigned long long h[2] = {0, 0};
    CHECK(hipMemcpy(h, clocks, sizeof h, hipMemcpyDeviceToHost));
    const double ops = (double)blocks * 256 * iters * kChains * kModes[m].ops;
    const double rate = ops / (ms * 1e-3);
    if (rate > best) {
      best = rate;
      ghz = h[1] ? (double)h[0] / (double)h[1] * 0.1 : 0.0;
    }
  }
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
  const double per_cu_clock = ghz > 0 ? best / 64.0 / cus / (ghz * 1e9) : 0.0;
  std::printf("{\"mode\": \"%s\", \"lane_ops_per_s\": %.6e, \"shader_clock_ghz\": %.4f, "
              "\"valu_instr_per_cu_per_clock\": %.4f, \"cus\": %d, \"workgroups\": %u, "
              "\"iterations\": %u}\n",
              kModes[m].name, best, ghz, per_cu_clock, cus, blocks, iters);
  std::fflush(stdout);
  return 0;
}

int main(int argc, char **argv) {
  const char *which = argc > 1 ? argv[1] : "mix31";
  const uint32_t blocks = argc > 2 ? (uint32_t)std::atoi(argv[2]) : 8192;
  const uint32_t iters = argc > 3 ? (uint32_t)std::atoi(argv[3]) : 20000;
  if (blocks == 0 || iters == 0) return 2;
  int first = -1;
  for (int m = 0; m < N_MODES; ++m)
    if (std::strcmp(which, kModes[m].name) == 0) first = m;
  const bool all = std::strcmp(which, "all") == 0;
  if (!all && first < 0) {
    std::fprintf(stderr, "hd_valu_probe: unknown mode %s\n", which);
    return 2;
  }
  hipDeviceProp_t prop;
  CHECK(hipGetDeviceProperties(&prop, 0));
  uint32_t *out = nullptr;
  unsigned long long *clocks = nullptr;
  CHECK(hipMalloc(&out, (size_t)blocks * 256 * sizeof(uint32_t)));
  CHECK(hipMalloc(&clocks, 2 * sizeof(unsigned long long)));
  int rc = 0;
  for (int m = all ? 0 : first; m < (all ? N_MODES : first + 1) && rc == 0; ++m)
    rc = run_mode(m, blocks, iters, prop.multiProcessorCount, out, clocks);
  CHECK(hipFree(out));
  CHECK(hipFree(clocks));
  return rc;
}
