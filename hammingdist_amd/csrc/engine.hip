// HIP engine of the all-pairs Hamming distance path: kernels other than the
// all-pairs tile kernel (kernels.h), their launchers, and the C ABI declared
// in include/hammingdist_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "engine_common.h"
#include "kernels.h"

namespace hd {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_last_error;

void set_last_error(const std::string &msg) { g_last_error = msg; }

// ---------------------------------------------------------------------------
// device block pool (DeviceBuffer)
// ---------------------------------------------------------------------------
size_t pool_class(size_t bytes) {
  if (bytes <= (1u << 20)) {
    size_t c = 256;
    while (c < bytes) c <<= 1;
    return c;
  }
  return (bytes + (2u << 20) - 1) & ~(size_t)((2u << 20) - 1);
}

namespace {
struct DevicePool {
  std::mutex mu;
  std::map<int, std::multimap<size_t, void *>> free; // per device: class -> blocks
  std::map<int, size_t> cached;
};
// Never destroyed: no HIP call runs during static destruction.
DevicePool &device_pool() {
  static DevicePool *p = new DevicePool;
  return *p;
}
int current_device() {
  int d = 0;
  (void)hipGetDevice(&d);
  return d;
}
} // namespace

void *pool_take(size_t cls, int *dev_out) {
  const int dev = current_device();
  *dev_out = dev;
  {
    DevicePool &pool = device_pool();
    std::lock_guard<std::mutex> lock(pool.mu);
    auto &blocks = pool.free[dev];
    auto it = blocks.find(cls);
    if (it != blocks.end()) {
      void *p = it->second;
      blocks.erase(it);
      pool.cached[dev] -= cls;
      return p;
    }
  }
  void *p = nullptr;
  HD_HIP(hipMalloc(&p, cls));
  return p;
}

void pool_give(void *p, size_t cls, int dev) {
  // The device synchronisation hipFree would perform: work still reading or
  // writing the block ends before anyone can take it.
  const int cur = current_device();
  if (cur != dev) (void)hipSetDevice(dev);
  (void)hipDeviceSynchronize();
  if (cur != dev) (void)hipSetDevice(cur);
  DevicePool &pool = device_pool();
  {
    std::lock_guard<std::mutex> lock(pool.mu);
    if (pool.cached[dev] + cls <= kPoolCachedMax) {
      pool.free[dev].emplace(cls, p);
      pool.cached[dev] += cls;
      return;
    }
  }
  (void)hipFree(p);
}

void ensure_device() {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    throw Error(HD_ERR_NO_DEVICE, "no HIP device available");
}

// ---------------------------------------------------------------------------
// plane codes
// ---------------------------------------------------------------------------
// Nibble codes of the packed input (src/hamming_impl.cc:31-44): A=1 C=2 G=4
// T=8 '-'=15, X=3 when packed with the include_x table, every other char 0.
static constexpr uint32_t kLetters = (1u << 1) | (1u << 2) | (1u << 4) | (1u << 8);
static constexpr uint32_t kGapBit = 1u << 15;
static constexpr uint32_t kXBit = 1u << 3;
static constexpr uint32_t kInvalidBit = 1u << 0;

PlaneCode plane_code(int semantics, uint32_t mask, uint32_t words) {
  PlaneCode pc;
  const uint8_t one_hot[4] = {1, 2, 4, 8};
  switch (semantics) {
  case HD_SEM_DENSE:
    pc.planes = 4;
    if ((mask & ~(kLetters | kGapBit | kInvalidBit) & 0xFFFFu) == 0 && words < kBufWords) {
      // Only 0, A, C, G, T and '-': the 3-op code, planes (u, v, w, y)
      // (the encoder derives them from the nibble bits; the table states
      // the same code).
      pc.op = kDense3;
      pc.table[0] = 0x7;  // invalid: u v w
      pc.table[1] = 0x1;  // A: u
      pc.table[2] = 0x3;  // C: u v
      pc.table[4] = 0x8;  // G: y
      pc.table[8] = 0xA;  // T: v y
      pc.table[15] = 0x4; // '-': w
      return pc;
    }
    // The nibble itself: f = ((a & b) == 0) for every pair of codes.
    for (int v = 0; v < 16; ++v) pc.table[v] = (uint8_t)v;
    return pc;
  case HD_SEM_SPARSE:
  case HD_SEM_SPARSE_X: {
    // f = classes differ and neither is '-'.  Classes: the letters, X
    // (SPARSE_X only) and one invalid class for every other code but '-'.
    // Each class present gets a code 0, 1, ...; with the kGap codes '-'
    // gets the next one, and 2 code bits serve up to 4 codes, else 3; the
    // plane after the code bits marks '-'.  Rows of kBufWords words or more
    // take kClass3, whose '-' plane masks each pair.
    const bool x_class = semantics == HD_SEM_SPARSE_X;
    const uint32_t known = kLetters | kGapBit | (x_class ? kXBit : 0u);
    int next = 0;
    uint8_t code[16];
    int inv_code = -1;
    if ((mask & ~known & 0xFFFFu) != 0) inv_code = next++;
    for (int l = 0; l < 4; ++l)
      code[one_hot[l]] = (mask >> one_hot[l]) & 1u ? (uint8_t)next++ : 0;
    const int x_code = x_class && (mask & kXBit) ? next++ : 0;
    // HD_SPARSE_CLASS_CODE=1 selects kClass3 for any row length (it lets
    // the long-row code be exercised on small data).
    const char *force_class = std::getenv("HD_SPARSE_CLASS_CODE");
    const bool gap_code = words < kBufWords && !(force_class && *force_class == '1');
    const int dash = gap_code ? next++ : 0;
    const int bits = gap_code && next <= 4 ? 2 : 3;
    pc.op = gap_code ? (bits == 2 ? kGap2 : kGap3) : kClass3;
    pc.planes = bits + 1;
    for (int v = 0; v < 16; ++v) {
      if (v == 15) pc.table[v] = (uint8_t)(dash | (1u << bits));
      else if ((kLetters >> v) & 1u) pc.table[v] = code[v];
      else if (x_class && v == 3) pc.table[v] = (uint8_t)x_code;
      else pc.table[v] = (uint8_t)(inv_code < 0 ? 0 : inv_code);
    }
    return pc;
  }
  case HD_SEM_CHAR_X: {
    // f = (a & b) == 0 || (a != b && no '-'): one-hot {A,C,G,T,X}, '-' all
    // ones, every invalid code all zeros (distance 1 against anything).
    const bool has_x = (mask & kXBit) != 0;
    pc.planes = 4 + (has_x ? 1 : 0);
    pc.op = pc.planes;
    for (int v = 0; v < 16; ++v) pc.table[v] = 0;
    for (int l = 0; l < 4; ++l) pc.table[one_hot[l]] = one_hot[l];
    pc.table[3] = has_x ? (uint8_t)(1u << 4) : 0;
    pc.table[15] = (uint8_t)((1u << pc.planes) - 1);
    return pc;
  }
  default:
    throw Error(HD_ERR_INVALID, "unknown semantics " + std::to_string(semantics));
  }
}

// ---------------------------------------------------------------------------
// tile variants of the all-pairs kernel
// ---------------------------------------------------------------------------
// Rows per lane, columns per wave, waves per workgroup.
static const Variant kVariants[] = {
    {4, 16, 4}, // 256 x 64 tiles: large n
    {8, 8, 4},  // 512 x 32
    {4, 8, 4},  // 256 x 32: the gap codes with votes at large n
    {2, 16, 4}, // 128 x 64: mid-size n
    {2, 8, 4},  // 128 x 32: small datasets, early exit under small caps
};
static constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

int variant_count() { return kNumVariants; }
const Variant &variant_of(int v) { return kVariants[v]; }

int default_variant(const Geometry &g, bool early_exit, uint32_t max_dist) {
  const uint64_t n = g.n;
  if (n < 512) return 4;
  const bool votes = check_interval(early_exit, max_dist, g.W) < g.W;
  // Below a few thousand rows the launch holds few tiles per CU slot and
  // each tile walks the whole row in turn, so smaller tiles finish sooner.
  // Every code copies its column words into VGPRs; the times below are with
  // those copies (v_bitop3 with VGPR operands only issues twice per CU and
  // clock, with an SGPR operand once).
  if (g.code.op == kDense3) {
    // A cap below kSmallCap is usually passed within the first chunks: a
    // tile stops at the vote its slowest pair passes, and 128 x 32 tiles
    // stop sooner than larger ones: n = 50,000, L = 29,903 with every 'N'
    // replaced by 'A' (no row settled by the invalid-site bound), cap 2:
    // 15.9 ms against 18.5-18.6 for 256 x 32 and 128 x 64 and 26.5 for 256 x
    // 64; half the row blocks so, 5.1 against 5.8-8.3 ms.  Where the bound
    // settles every row the launch is a fill of the triangle, which 128 x 32
    // tiles split finest (0.49-0.51 ms against 0.24-0.29).
    constexpr uint32_t kSmallCap = 64;
    if (votes && max_dist < kSmallCap) return 4;
    // Larger caps, and no votes: 128 x 64 tiles from 8,192 rows (256 x 64
    // with votes from 24,576), 128 x 32 below (L = 29,903 without votes:
    // n = 6,000 / 20,000 / 50,000 128 x 64 1.65 / 14.7 / 90.4 ms, 128 x 32
    // 1.61 / 16.1 / 105.3, 256 x 64 1.92 / 15.6 / 92.3; with votes at cap
    // 255, n = 20,000 / 50,000 128 x 64 15.9 / 99.0 ms, 256 x 64 16.3 /
    // 97.6).  128 x 64 tiles fetch twice the bytes from HBM that 256 x 64
    // tiles do (240 GB against 124 GB per launch at n = 50,000: each tile
    // streams its columns over the whole row, for half as many rows), 2.7
    // TB/s at that time, a third of the HBM rate.
    if (votes) return n < 8192 ? 4 : n < 24576 ? 3 : 0;
    return n < 8192 ? 4 : 3;
  }
  if (g.code.op == kGap2 || g.code.op == kGap3) {
    // The wide data (L = 10,000, gap-3 code) without votes: 256 x 64 tiles
    // from 32,768 rows (n = 100,000: 123.6-124.3 ms against 125.0 for 128 x
    // 64 and 131.7-133.1 for 256 x 32; n = 50,000: 31.2 / 31.8 / 33.5 ms),
    // 128 x 64 below (n = 6,000 / 10,000 / 20,000: 0.59 / 1.43 / 5.20 ms
    // against 0.69 / 1.52 / 5.27 for 256 x 64), 128 x 32 below 4,096 (n =
    // 3,000: 0.21 against 0.25 ms).  With votes, 256 x 32 tiles from 32,768
    // rows (n = 100,000, cap 255: 51.4-51.7 ms against 51.2-53.5 for 128 x
    // 32; cap 3: 16.0 against 16.9), 128 x 32 below (n = 20,000, cap 255:
    // 2.07-2.12 against 2.16).
    if (votes) return n < 32768 ? 4 : 2;
    return n < 4096 ? 4 : n < 32768 ? 3 : 0;
  }
  // The pointer-addressed codes (nibble, one-hot, class-3) on the wide data
  // at L = 10,000: 256 x 64 tiles from 8,192 rows (n = 30,000: nibble code
  // 15.1 ms against 15.9 for 128 x 64, one-hot 17.9 against 19.1 for 256 x
  // 32; n = 10,000: 1.92 / 2.25 ms, tied with or ahead of the others), 128 x
  // 64 below (n = 6,000: 0.76 against 0.78-0.84 ms), 128 x 32 below 4,096
  // (n = 3,000: 0.29 against 0.33 ms).
  return n < 4096 ? 4 : n < 8192 ? 3 : 0;
}

template <int P, int R, int C, int NW, typename OutT>
static void launch_allpairs_t(const uint32_t *planes, uint32_t n,
                              uint32_t n_pad, uint32_t W, uint64_t rb0,
                              uint64_t rb1, OutT *out, uint64_t cell_offset,
                              uint32_t max_dist, uint32_t check_every,
                              hipStream_t s) {
  const uint32_t *corr =
      has_corr(P) ? planes + (size_t)W * stored_planes(P) * n_pad : nullptr;
  constexpr uint64_t M = (64 * R) / (NW * C);
  const uint64_t t0 = tiles_before(rb0, M), t1 = tiles_before(rb1, M);
  const uint64_t nt = t1 - t0;
  if (nt == 0) return;
  if (nt > 0x7fffffffull) throw Error(HD_ERR_INVALID, "too many tiles in one launch");
  using Kernel = void (*)(const uint32_t *, uint32_t, uint32_t, uint64_t, uint32_t, OutT *,
                         uint64_t, uint32_t, uint32_t, const uint32_t *, uint32_t);
  Kernel kernel;
  if constexpr (P == kGap2 || P == kGap3)
    kernel = check_every < W ? k_allpairs_gap<P, R, C, NW, OutT, true>
                             : k_allpairs_gap<P, R, C, NW, OutT, false>;
  else if constexpr (P == kDense3)
    kernel = check_every < W ? k_allpairs<P, R, C, NW, OutT, true>
                             : k_allpairs<P, R, C, NW, OutT, false>;
  else
    // The pointer-addressed codes take the vote-compiled kernel also when
    // there is no vote (check_every = W: one chunk, the vote skipped): their
    // vote-free builds spill 18-128 SGPRs into VGPR lanes.
    kernel = k_allpairs<P, R, C, NW, OutT, true>;
  hipLaunchKernelGGL(kernel, dim3((uint32_t)nt), dim3(64 * NW), 0, s, planes, n, W, t0,
                     (uint32_t)nt, out, cell_offset, max_dist, check_every, corr, n_pad);
  HD_HIP(hipGetLastError());
}

template <int P, typename OutT>
static void launch_allpairs_p(int v, const uint32_t *planes, uint32_t n,
                              uint32_t n_pad, uint32_t W, uint64_t rb0,
                              uint64_t rb1, OutT *out, uint64_t cell_offset,
                              uint32_t max_dist, uint32_t check_every,
                              hipStream_t s) {
#define HD_V(idx, R, C, NW)                                                    \
  case idx:                                                                    \
    launch_allpairs_t<P, R, C, NW, OutT>(planes, n, n_pad, W, rb0, rb1, out,   \
                                         cell_offset, max_dist, check_every, s); \
    return;
  switch (v) {
    HD_V(0, 4, 16, 4)
    HD_V(1, 8, 8, 4)
    HD_V(2, 4, 8, 4)
    HD_V(3, 2, 16, 4)
    HD_V(4, 2, 8, 4)
  default:
    throw Error(HD_ERR_INVALID, "unknown tile variant");
  }
#undef HD_V
}

template <typename OutT>
static void launch_allpairs_o(int P, int v, const uint32_t *planes, uint32_t n,
                              uint32_t n_pad, uint32_t W, uint64_t rb0,
                              uint64_t rb1, OutT *out, uint64_t cell_offset,
                              uint32_t max_dist, uint32_t check_every,
                              hipStream_t s) {
  switch (P) {
  case kDense3:
    return launch_allpairs_p<kDense3, OutT>(v, planes, n, n_pad, W, rb0, rb1, out,
                                            cell_offset, max_dist, check_every, s);
  case 4:
    return launch_allpairs_p<4, OutT>(v, planes, n, n_pad, W, rb0, rb1, out,
                                      cell_offset, max_dist, check_every, s);
  case 5:
    return launch_allpairs_p<5, OutT>(v, planes, n, n_pad, W, rb0, rb1, out,
                                      cell_offset, max_dist, check_every, s);
  case kGap2:
    return launch_allpairs_p<kGap2, OutT>(v, planes, n, n_pad, W, rb0, rb1, out,
                                          cell_offset, max_dist, check_every, s);
  case kGap3:
    return launch_allpairs_p<kGap3, OutT>(v, planes, n, n_pad, W, rb0, rb1, out,
                                          cell_offset, max_dist, check_every, s);
  case kClass3:
    return launch_allpairs_p<kClass3, OutT>(v, planes, n, n_pad, W, rb0, rb1, out,
                                            cell_offset, max_dist, check_every, s);
  default:
    throw Error(HD_ERR_INVALID, "unsupported plane count");
  }
}

static constexpr uint64_t kRowPad = 512; // multiple of every variant's TI

Geometry geometry(uint64_t n, uint64_t bpr, int semantics, uint32_t mask) {
  HD_REQUIRE(bpr > 0 && bpr % 8 == 0,
             "bytes_per_seq must be a positive multiple of 8");
  HD_REQUIRE(n < (1ull << 31), "too many sequences");
  Geometry g;
  g.n = n;
  g.bpr = bpr;
  g.n_pad = std::max<uint64_t>(kRowPad, (n + kRowPad - 1) / kRowPad * kRowPad);
  g.W = (uint32_t)((2 * bpr + 31) / 32);
  g.code = plane_code(semantics, mask, g.W);
  return g;
}

static uint32_t tile_rows(int v) { return 64u * kVariants[v].R; }

static_assert(kCheckWords == 32, "Geometry::total_words assumes 32-word chunks");

uint32_t check_interval(bool early_exit, uint32_t max_dist, uint32_t W) {
  // The early-exit vote runs every kCheckWords words (1024 sites); it is
  // skipped when max_dist is out of reach.
  return (early_exit && (uint64_t)max_dist < 32ull * W) ? kCheckWords : W;
}

void run_allpairs(const Geometry &g, const uint32_t *planes, int out_bytes,
                  uint32_t max_dist, bool early_exit, int variant, uint64_t rb0,
                  uint64_t rb1, void *out, uint64_t cell_offset, hipStream_t s) {
  const uint32_t W = g.W;
  const uint32_t check_every = check_interval(early_exit, max_dist, W);
  const int P = g.code.op;
  switch (out_bytes) {
  case 1:
    launch_allpairs_o<uint8_t>(P, variant, planes, (uint32_t)g.n,
                               (uint32_t)g.n_pad, W, rb0, rb1, (uint8_t *)out,
                               cell_offset, max_dist, check_every, s);
    break;
  case 2:
    launch_allpairs_o<uint16_t>(P, variant, planes, (uint32_t)g.n,
                                (uint32_t)g.n_pad, W, rb0, rb1, (uint16_t *)out,
                                cell_offset, max_dist, check_every, s);
    break;
  case 4:
    launch_allpairs_o<uint32_t>(P, variant, planes, (uint32_t)g.n,
                                (uint32_t)g.n_pad, W, rb0, rb1, (uint32_t *)out,
                                cell_offset, max_dist, check_every, s);
    break;
  default:
    throw Error(HD_ERR_INVALID, "out_bytes must be 1, 2 or 4");
  }
}

static void check_max_dist(int out_bytes, uint32_t max_dist) {
  HD_REQUIRE(out_bytes == 1 || out_bytes == 2 || out_bytes == 4,
             "out_bytes must be 1, 2 or 4");
  const uint64_t cap = out_bytes == 4 ? 0xFFFFFFFFull : ((1ull << (8 * out_bytes)) - 1);
  HD_REQUIRE(max_dist <= cap, "max_dist does not fit the output cell type");
}

// ---------------------------------------------------------------------------
// encode: packed nibble rows -> bit-sliced planes
// ---------------------------------------------------------------------------
// Workgroup b encodes tile (row group b % groups, words (b / groups) *
// kEncodeWords ...) through LDS (kernels.h, encode_tile).
__global__ __launch_bounds__(256) void k_encode(const uint8_t *__restrict__ packed,
                                                uint32_t n, uint64_t bpr, uint32_t groups,
                                                uint32_t W, int P, int op, Table16 tbl,
                                                int mode, uint32_t *__restrict__ planes,
                                                uint32_t *__restrict__ corr) {
  __shared__ uint8_t s_tbl[16];
  __shared__ __attribute__((aligned(16))) uint8_t s_rows[kEncodeLdsBytes];
  if (threadIdx.x < 16) s_tbl[threadIdx.x] = tbl.v[threadIdx.x];
  const uint32_t g = blockIdx.x % groups, kt = blockIdx.x / groups;
  encode_tile(packed, n, bpr, W, P, op, s_tbl, mode, planes, corr, groups * kWave, g,
              kt * kEncodeWords, s_rows);
}

// Whole-row correction counts of the kDense3 and kGap codes from the chunk
// counts encode wrote, one thread per row.
__global__ __launch_bounds__(256) void k_corr_totals(uint32_t *__restrict__ corr, uint32_t W,
                                                     uint32_t n_pad, int op) {
  const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row < n_pad) corr_totals_row(corr, W, n_pad, op, row);
}

int encode_mode(const PlaneCode &c) {
  return is_dense3(c) ? kEncDense3 : is_identity(c) ? kEncIdentity : kEncTable;
}

bool is_identity(const PlaneCode &c) {
  if (c.planes != 4 || c.op != 4) return false;
  for (int v = 0; v < 16; ++v)
    if (c.table[v] != v) return false;
  return true;
}

void launch_encode(const Geometry &g, const uint8_t *d_packed, uint32_t *d_planes,
                   hipStream_t s) {
  Table16 t;
  std::memcpy(t.v, g.code.table, 16);
  const uint64_t groups = g.n_pad / kWave;
  const uint64_t blocks = groups * ((g.W + kEncodeWords - 1) / kEncodeWords);
  HD_REQUIRE(blocks <= 0x7fffffffull, "dataset too large for the encode grid");
  uint32_t *corr = has_corr(g.code.op) ? d_planes + g.plane_words() : nullptr;
  hipLaunchKernelGGL(k_encode, dim3((uint32_t)blocks), dim3(256), 0, s, d_packed,
                     (uint32_t)g.n, g.bpr, (uint32_t)groups, g.W, g.code.planes, g.code.op,
                     t, encode_mode(g.code), d_planes, corr);
  HD_HIP(hipGetLastError());
  if (corr) {
    hipLaunchKernelGGL(k_corr_totals, dim3((uint32_t)(g.n_pad / 256)), dim3(256), 0, s, corr,
                       g.W, (uint32_t)g.n_pad, g.code.op);
    HD_HIP(hipGetLastError());
  }
}

// ---------------------------------------------------------------------------
// nibble presence mask
// ---------------------------------------------------------------------------
// Units of 8 bytes per thread, loaded before any is used.
static constexpr uint32_t kMaskLoads = 8;

__global__ __launch_bounds__(256) void k_nibble_mask(const uint2 *__restrict__ data,
                                                     uint64_t n_units,
                                                     uint32_t *__restrict__ mask) {
  // Each thread takes kMaskLoads units of its workgroup's 256 x kMaskLoads
  // (coalesced) and issues every load before using one, so the whole grid's
  // bytes stream (a grid-stride loop over 1,024 workgroups with four, then
  // eight, loads per step ran 0.143-0.163 ms for 0.75 GB).  A unit past the
  // end repeats the last one, which adds no new code.
  using u32x2 = uint32_t __attribute__((ext_vector_type(2)));
  const uint64_t u0 = (uint64_t)blockIdx.x * (256 * kMaskLoads) + threadIdx.x;
  u32x2 v[kMaskLoads];
#pragma unroll
  for (uint32_t k = 0; k < kMaskLoads; ++k) {
    const uint64_t x = min(u0 + 256 * k, n_units - 1);
    v[k] = __builtin_nontemporal_load(reinterpret_cast<const u32x2 *>(data) + x);
  }
  uint32_t m = 0;
#pragma unroll
  for (uint32_t k = 0; k < kMaskLoads; ++k)
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      m |= 1u << ((v[k].x >> (4 * s)) & 15u);
      m |= 1u << ((v[k].y >> (4 * s)) & 15u);
    }
  // OR across the wavefront, then across the workgroup; then an atomic only
  // when the workgroup holds a code the mask lacks (atomics on one address
  // serialise, and after the first workgroups almost every one holds only
  // codes already set).  Bits are only ever set, so one seen set stays set.
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m |= __shfl_xor(m, off);
  __shared__ uint32_t s_m[4];
  if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t w = s_m[0] | s_m[1] | s_m[2] | s_m[3];
    if (w & ~__hip_atomic_load(mask, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicOr(mask, w);
  }
}

void launch_nibble_mask(const uint8_t *d_packed, uint64_t bytes, uint32_t *d_mask,
                        hipStream_t s) {
  const uint64_t units = bytes / 8;
  HD_HIP(hipMemsetAsync(d_mask, 0, sizeof(uint32_t), s));
  if (!units) return;
  const uint64_t blocks = (units + 256 * kMaskLoads - 1) / (256 * kMaskLoads);
  HD_REQUIRE(blocks < (1ull << 31), "too many bytes for the nibble mask");
  hipLaunchKernelGGL(k_nibble_mask, dim3((uint32_t)blocks), dim3(256), 0, s,
                     reinterpret_cast<const uint2 *>(d_packed), units, d_mask);
  HD_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// consensus letter counts (input of the sparse/dense heuristic)
// ---------------------------------------------------------------------------
// Thread t owns the 4-byte word w (8 sites) of rows [r0, r0 + 512): the
// loads of a wavefront are 256 contiguous bytes of one row.
static constexpr uint32_t kCountRows = 512;

__global__ __launch_bounds__(256) void k_letter_counts(const uint8_t *__restrict__ packed,
                                                       uint32_t n, uint64_t bpr,
                                                       uint64_t L,
                                                       uint32_t *__restrict__ counts,
                                                       uint32_t *__restrict__ mask) {
  const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t words = bpr / 4;
  const uint32_t r0 = blockIdx.y * kCountRows;
  const uint32_t r1 = min(n, r0 + kCountRows);
  uint32_t c[8][4];
#pragma unroll
  for (int s = 0; s < 8; ++s)
#pragma unroll
    for (int l = 0; l < 4; ++l) c[s][l] = 0;
  uint32_t m = 0;
  if (w < words) {
    for (uint32_t r = r0; r < r1; ++r) {
      const uint32_t v = reinterpret_cast<const uint32_t *>(packed + (uint64_t)r * bpr)[w];
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        const uint32_t nib = (v >> (4 * s)) & 15u;
        if (8 * w + s < L) m |= 1u << nib;
        c[s][0] += nib == 1u;
        c[s][1] += nib == 2u;
        c[s][2] += nib == 4u;
        c[s][3] += nib == 8u;
      }
    }
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const uint64_t site = 8 * w + s;
      if (site < L)
#pragma unroll
        for (int l = 0; l < 4; ++l)
          if (c[s][l]) atomicAdd(&counts[site * 4 + l], c[s][l]);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m |= __shfl_xor(m, off);
  if ((threadIdx.x & 63) == 0 && m) atomicOr(mask, m);
}

// ---------------------------------------------------------------------------
// reference distances: one query against many packed rows
// ---------------------------------------------------------------------------
// A persistent grid of wavefronts: each holds the query in VGPRs (RegQuery)
// and walks rows k = wave, wave + waves, ...; lane l reads 8-byte units
// l, l + 64, ... of a row, all of them issued before the first is used.
// Rows longer than the register form covers read the query from memory.
template <int SEM, int T>
__global__ __launch_bounds__(256) void k_refdist(const uint2 *__restrict__ query,
                                                 const uint8_t *__restrict__ db,
                                                 uint64_t bpr,
                                                 const uint64_t *__restrict__ rows,
                                                 uint64_t m,
                                                 uint32_t *__restrict__ out) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t units = bpr / 8;
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / 64);
  const uint64_t k0 = (uint64_t)blockIdx.x * (blockDim.x / 64) +
                      __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
  if constexpr (T > 0) {
    RegQuery<SEM, T> q;
    q.load(query, (uint32_t)units, lane);
    for (uint64_t k = k0; k < m; k += waves) {
      const uint64_t r = rows ? rows[k] : k;
      const uint32_t d =
          q.distance(reinterpret_cast<const uint2 *>(db + r * bpr), (uint32_t)units, lane);
      if (lane == 0) out[k] = d;
    }
  } else {
    for (uint64_t k = k0; k < m; k += waves) {
      const uint64_t r = rows ? rows[k] : k;
      const uint32_t d = row_distance_mem<SEM>(
          query, reinterpret_cast<const uint2 *>(db + r * bpr), units, lane);
      if (lane == 0) out[k] = d;
    }
  }
}

// Workgroups of `block` threads of kernel `fn` resident on the whole current
// device at once; looked up once per (device, kernel).
static uint64_t resident_blocks(const void *fn, int block) {
  static std::mutex mu;
  static std::map<std::pair<int, const void *>, uint64_t> cache;
  int dev = 0;
  HD_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  const auto key = std::make_pair(dev, fn);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int cus = 0, per_cu = 0;
  HD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  HD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, 0));
  const uint64_t r = (uint64_t)std::max(cus, 1) * std::max(per_cu, 1);
  cache.emplace(key, r);
  return r;
}

template <int SEM>
static const void *refdist_kernel(int t) {
  switch (t) {
  case 4: return reinterpret_cast<const void *>(&k_refdist<SEM, 4>);
  case 8: return reinterpret_cast<const void *>(&k_refdist<SEM, 8>);
  case 16: return reinterpret_cast<const void *>(&k_refdist<SEM, 16>);
  case 32: return reinterpret_cast<const void *>(&k_refdist<SEM, 32>);
  default: return reinterpret_cast<const void *>(&k_refdist<SEM, 0>);
  }
}

static void launch_refdist(const uint8_t *d_query, const uint8_t *d_db,
                           uint64_t bpr, const uint64_t *d_rows, uint64_t m,
                           int semantics, uint32_t *d_out, hipStream_t s) {
  if (m == 0) return;
  if (semantics != HD_SEM_DENSE && semantics != HD_SEM_CHAR_X)
    throw Error(HD_ERR_INVALID,
                "reference distances support HD_SEM_DENSE and HD_SEM_CHAR_X");
  const int t = reg_query_t(bpr / 8);
  const void *fn = semantics == HD_SEM_DENSE ? refdist_kernel<HD_SEM_DENSE>(t)
                                             : refdist_kernel<HD_SEM_CHAR_X>(t);
  // Workgroups of 4 wavefronts; as many as are resident at once (each
  // wavefront loads the query once), but no more than the rows need.
  const uint64_t resident = resident_blocks(fn, 256);
  const uint32_t blocks = (uint32_t)std::max<uint64_t>(1, std::min(resident, (m + 3) / 4));
  const uint2 *q = reinterpret_cast<const uint2 *>(d_query);
  void *args[] = {&q, &d_db, &bpr, &d_rows, &m, &d_out};
  HD_HIP(hipLaunchKernel(fn, dim3(blocks), dim3(256), args, 0, s));
}

// ---------------------------------------------------------------------------
// sparse emit: cells <= threshold of a device triangle, row-major order
// ---------------------------------------------------------------------------
// Each thread owns 16 consecutive cells (one to four 16-byte loads), each
// 256-thread block kEmitCells = 4096 cells.  Pass 1 counts the qualifying
// cells per block; a two-level scan on the device turns the counts into
// block offsets (per 4096-block chunk, then across chunks); pass 2 re-reads
// the cells and writes each block's hits in cell order (k_emit_write).  Only
// the total goes to the host.
static constexpr uint32_t kEmitPerThread = 16;
static constexpr uint32_t kEmitCells = 256 * kEmitPerThread;
static constexpr uint32_t kScanChunk = 256 * 16; // block counts per scan block

// The 16 cells x0 .. x0+15 as uint32 values; returns how many of them
// exist (cells past n_cells read as 0 and must be ignored).  vec: x0 is
// 16-cell aligned in a 16-byte aligned buffer, so whole 16-byte loads where
// all 16 cells exist.
template <typename T>
__device__ __forceinline__ uint32_t load16(const T *__restrict__ cells, uint64_t n_cells,
                                           uint64_t x0, bool vec, uint32_t (&v)[16]) {
  constexpr uint32_t kMask = sizeof(T) >= 4 ? 0xFFFFFFFFu : (1u << (8 * sizeof(T))) - 1u;
  if (vec && x0 + 16 <= n_cells) {
    constexpr int kLoads = 16 * sizeof(T) / 16;
    using u32x4 = uint32_t __attribute__((ext_vector_type(4)));
    const u32x4 *p = reinterpret_cast<const u32x4 *>(cells + x0);
#pragma unroll
    for (int l = 0; l < kLoads; ++l) {
      // streaming loads: each pass reads a cell once
      const u32x4 q = __builtin_nontemporal_load(p + l);
      const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        constexpr int per = 4 / sizeof(T); // cells per 32-bit word
#pragma unroll
        for (int e = 0; e < per; ++e)
          v[(l * 4 + k) * per + e] = (w[k] >> (8 * sizeof(T) * e)) & kMask;
      }
    }
    return 16;
  }
#pragma unroll
  for (int e = 0; e < 16; ++e) v[e] = x0 + e < n_cells ? (uint32_t)cells[x0 + e] : 0u;
  return x0 >= n_cells ? 0u : (uint32_t)min<uint64_t>(16, n_cells - x0);
}

// Exclusive prefix of v over the 256 threads of the block; *total = the sum.
__device__ __forceinline__ uint64_t block_exclusive_scan(uint64_t v, uint64_t *total) {
  __shared__ uint64_t s_wave[4];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint64_t up = __shfl_up(incl, off);
    if (lane >= (uint32_t)off) incl += up;
  }
  if (lane == 63) s_wave[wave] = incl;
  __syncthreads();
  uint64_t before = 0;
  for (uint32_t w = 0; w < wave; ++w) before += s_wave[w];
  *total = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
  __syncthreads();
  return before + incl - v;
}

// Cell blocks counted per workgroup: one block of 4,096 uint8 cells per
// workgroup left the 1.25e9-cell count at 0.27 ms (4.6 TB/s), bound by the
// 305,000 workgroups' launches rather than by the bytes.
static constexpr uint32_t kCountBlocksPerGroup = 4;

template <typename T>
__global__ __launch_bounds__(256) void k_emit_count(const T *__restrict__ cells,
                                                    uint64_t n_cells, uint32_t thr, int vec,
                                                    uint64_t blocks,
                                                    uint32_t *__restrict__ counts) {
  constexpr uint32_t B = kCountBlocksPerGroup;
  const uint64_t b0 = (uint64_t)blockIdx.x * B;
  uint32_t c[B];
  const uint64_t xl = (b0 + B) * kEmitCells;
  if (vec != 0 && xl <= n_cells) {
    // every block whole: all the loads first, then the counts
    using u32x4 = uint32_t __attribute__((ext_vector_type(4)));
    constexpr int kLoads = 16 * sizeof(T) / 16;
    u32x4 q[B][kLoads];
#pragma unroll
    for (uint32_t k = 0; k < B; ++k)
#pragma unroll
      for (int l = 0; l < kLoads; ++l)
        q[k][l] = __builtin_nontemporal_load(
            reinterpret_cast<const u32x4 *>(cells + (b0 + k) * kEmitCells +
                                            threadIdx.x * kEmitPerThread) + l);
    constexpr uint32_t kMask = sizeof(T) >= 4 ? 0xFFFFFFFFu : (1u << (8 * sizeof(T))) - 1u;
#pragma unroll
    for (uint32_t k = 0; k < B; ++k) {
      c[k] = 0;
#pragma unroll
      for (int l = 0; l < kLoads; ++l)
#pragma unroll
        for (int wd = 0; wd < 4; ++wd)
#pragma unroll
          for (int e = 0; e < (int)(4 / sizeof(T)); ++e)
            c[k] += ((q[k][l][wd] >> (8 * sizeof(T) * e)) & kMask) <= thr ? 1u : 0u;
    }
  } else {
#pragma unroll
    for (uint32_t k = 0; k < B; ++k) {
      uint32_t v[16];
      const uint64_t x0 = (b0 + k) * kEmitCells + threadIdx.x * kEmitPerThread;
      const uint32_t valid = load16(cells, n_cells, x0, vec != 0, v);
      c[k] = 0;
#pragma unroll
      for (int e = 0; e < 16; ++e) c[k] += ((uint32_t)e < valid && v[e] <= thr) ? 1u : 0u;
    }
  }
  __shared__ uint32_t s_sum[B][4];
#pragma unroll
  for (uint32_t k = 0; k < B; ++k) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c[k] += __shfl_xor(c[k], off);
    if ((threadIdx.x & 63) == 0) s_sum[k][threadIdx.x >> 6] = c[k];
  }
  __syncthreads();
  if (threadIdx.x < B && b0 + threadIdx.x < blocks) {
    const uint32_t k = threadIdx.x;
    counts[b0 + k] = s_sum[k][0] + s_sum[k][1] + s_sum[k][2] + s_sum[k][3];
  }
}

// Level 1: exclusive prefix of the block counts inside chunks of kScanChunk
// (written over the counts) and each chunk's sum.
__global__ __launch_bounds__(256) void k_emit_scan_local(uint32_t *__restrict__ counts,
                                                         uint64_t blocks,
                                                         uint64_t *__restrict__ chunk_sums) {
  const uint64_t b0 = (uint64_t)blockIdx.x * kScanChunk + threadIdx.x * 16;
  uint32_t c[16];
  uint64_t sum = 0;
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    c[e] = b0 + e < blocks ? counts[b0 + e] : 0u;
    sum += c[e];
  }
  uint64_t total = 0;
  uint64_t run = block_exclusive_scan(sum, &total);
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    if (b0 + e < blocks) counts[b0 + e] = (uint32_t)run;
    run += c[e];
  }
  if (threadIdx.x == 0) chunk_sums[blockIdx.x] = total;
}

// Level 2 (one block): exclusive prefix of the chunk sums, in place, and
// the grand total in chunk_sums[chunks].
__global__ __launch_bounds__(256) void k_emit_scan_chunks(uint64_t *__restrict__ chunk_sums,
                                                          uint64_t chunks) {
  uint64_t carry = 0;
  for (uint64_t c0 = 0; c0 < chunks; c0 += 256) {
    const uint64_t c = c0 + threadIdx.x;
    const uint64_t v = c < chunks ? chunk_sums[c] : 0;
    uint64_t total = 0;
    const uint64_t ex = block_exclusive_scan(v, &total);
    if (c < chunks) chunk_sums[c] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) chunk_sums[chunks] = carry;
}

__device__ __forceinline__ uint64_t row_of_cell(uint64_t x) {
  // Largest i with i (i - 1) / 2 <= x.
  uint64_t i = (uint64_t)floor(sqrt(2.0 * (double)x + 0.25) + 0.5);
  while (i * (i - 1) / 2 > x) --i;
  while ((i + 1) * i / 2 <= x) ++i;
  return i;
}

// Pass 2.  Wave w of a block owns the block's cells [1024 w, 1024 (w + 1)):
// the 16-byte loads put them into LDS, then in step e (0..15) lane l takes
// cell 64 e + l, so the hits of one step are consecutive cells and their
// triplets land at consecutive positions (a ballot and mbcnt give each
// lane's position; triplet words 3 apart spread over all LDS banks).  Each
// wave stages the triplets of 8 steps at a time in its own LDS (shifted so
// that 16-byte quads of the output are 16-byte aligned in LDS) and stores
// them with consecutive lanes on consecutive quads; only the wave totals
// need a block barrier.
static constexpr uint32_t kEmitWaveCells = 1024;
static constexpr uint32_t kEmitHalf = 8;                        // steps per staging turn
static constexpr uint32_t kEmitStageWords = 3 * 64 * kEmitHalf + 4; // per wave

template <typename T>
__global__ __launch_bounds__(256) void k_emit_write(const T *__restrict__ cells,
                                                    uint64_t n_cells, uint32_t thr, int vec,
                                                    const uint32_t *__restrict__ local_prefix,
                                                    const uint64_t *__restrict__ chunk_prefix,
                                                    uint64_t capacity, int vec_out,
                                                    uint32_t *__restrict__ trip) {
  __shared__ T s_cells[kEmitCells];
  __shared__ __attribute__((aligned(16))) uint32_t s_stage[4][kEmitStageWords];
  __shared__ uint32_t s_wave_hits[4];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t xb = (uint64_t)blockIdx.x * kEmitCells;
  {
    uint32_t v[16];
    load16(cells, n_cells, xb + threadIdx.x * kEmitPerThread, vec != 0, v);
#pragma unroll
    for (int e = 0; e < 16; ++e) s_cells[threadIdx.x * kEmitPerThread + e] = (T)v[e];
  }
  wave_lds_sync(); // each wave reads back only the cells its own lanes wrote
  const uint64_t xw = xb + (uint64_t)wave * kEmitWaveCells;
  const T *wc = s_cells + wave * kEmitWaveCells;
  uint64_t bal[16];
  uint32_t val[16];
  uint32_t wave_hits = 0;
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    val[e] = wc[64 * e + lane];
    const bool hit = xw + 64 * e + lane < n_cells && val[e] <= thr;
    bal[e] = __ballot(hit);
    wave_hits += (uint32_t)__builtin_popcountll(bal[e]);
  }
  if (lane == 0) s_wave_hits[wave] = wave_hits;
  __syncthreads();
  uint64_t base = local_prefix[blockIdx.x] + chunk_prefix[blockIdx.x / kScanChunk];
  for (uint32_t w = 0; w < wave; ++w) base += s_wave_hits[w];
  if (wave_hits == 0) return;
  // row of this lane's first cell; each step moves 64 cells on
  uint64_t x = xw + lane;
  uint64_t i = row_of_cell(x), row_start = i * (i - 1) / 2;
  const uint64_t limit = 3 * capacity;
  uint32_t *const st = s_stage[wave];
#pragma unroll
  for (int h = 0; h < 16; h += kEmitHalf) {
    uint32_t turn = 0;
#pragma unroll
    for (int e = h; e < h + (int)kEmitHalf; ++e)
      turn += (uint32_t)__builtin_popcountll(bal[e]);
    if (turn == 0) {
      x += 64 * kEmitHalf;
      continue;
    }
    const uint64_t first = 3 * base;
    const uint32_t shift = (uint32_t)(first & 3u);
    uint32_t a = 0;
#pragma unroll
    for (int e = h; e < h + (int)kEmitHalf; ++e) {
      while (x >= row_start + i) { // next row (rows of length i)
        row_start += i;
        ++i;
      }
      if ((bal[e] >> lane) & 1u) {
        const uint32_t k =
            shift + 3 * (a + __builtin_amdgcn_mbcnt_hi((uint32_t)(bal[e] >> 32),
                                                       __builtin_amdgcn_mbcnt_lo(
                                                           (uint32_t)bal[e], 0u)));
        st[k] = (uint32_t)i;
        st[k + 1] = (uint32_t)(x - row_start);
        st[k + 2] = val[e];
      }
      a += (uint32_t)__builtin_popcountll(bal[e]);
      x += 64;
    }
    wave_lds_sync();
    const uint64_t words = first >= limit ? 0 : min<uint64_t>(3ull * turn, limit - first);
    if (!vec_out) {
      for (uint64_t w = lane; w < words; w += 64) trip[first + w] = st[shift + w];
    } else {
      // 16-byte stores from the first 16-byte boundary of the output on (the
      // output itself is 16-byte aligned); single words before and after.
      const uint32_t head = (uint32_t)min<uint64_t>((4u - shift) & 3u, words);
      const uint64_t quads = (words - head) / 4;
      if (lane < head) trip[first + lane] = st[shift + lane];
      using u32x4 = uint32_t __attribute__((ext_vector_type(4)));
      u32x4 *dst = reinterpret_cast<u32x4 *>(trip + first + head);
      const u32x4 *src = reinterpret_cast<const u32x4 *>(st + shift + head);
      // streaming stores: the triplets are not read back by this kernel
      for (uint64_t q = lane; q < quads; q += 64) __builtin_nontemporal_store(src[q], dst + q);
      for (uint64_t w = head + 4 * quads + lane; w < words; w += 64)
        trip[first + w] = st[shift + w];
    }
    wave_lds_sync(); // the stores have read the stage before the next turn writes it
    base += turn;
  }
}

// Scratch of one emit over n_cells cells: the scan's chunk sums
// (chunks + 1 uint64) followed by the block counts (blocks uint32).
static constexpr uint64_t emit_blocks(uint64_t n_cells) {
  return (n_cells + kEmitCells - 1) / kEmitCells;
}
static constexpr uint64_t emit_chunks(uint64_t blocks) {
  return (blocks + kScanChunk - 1) / kScanChunk;
}
static constexpr uint64_t emit_scratch_bytes(uint64_t n_cells) {
  return emit_blocks(n_cells) ? (emit_chunks(emit_blocks(n_cells)) + 1) * 8 +
                                    emit_blocks(n_cells) * 4
                              : 0;
}

// ---------------------------------------------------------------------------
// the emit's scratch without allocation (hd_sparse_emit_dev_noalloc)
// ---------------------------------------------------------------------------
// A file-scope device array sized for HD_EMIT_NOALLOC_MAX_CELLS cells (8 MiB),
// held by every GPU the library runs on until the process ends.  Host code
// never uses the array's name as an address: its device address is taken
// once per device with hipGetSymbolAddress, with that device current, and
// kept in g_scratch_ptr.  Every call on a GPU shares that GPU's array.
static constexpr uint64_t kNoAllocScratchWords =
    (emit_scratch_bytes(HD_EMIT_NOALLOC_MAX_CELLS) + 7) / 8;
static __device__ __align__(256) uint64_t scratch[kNoAllocScratchWords];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device.
static void *noalloc_scratch() {
  int dev = 0;
  HD_HIP(hipGetDevice(&dev));
  HD_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    HD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    HD_REQUIRE(p, "no device address for the emit scratch");
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

// Emits into device triplets, with the scratch at `work`
// (emit_scratch_bytes(n_cells) bytes, 8-byte aligned); returns the total
// number of qualifying cells.  Every scratch word is written before it is
// read, so the scratch needs no clearing.
template <typename T>
static uint64_t emit_t(const T *d_cells, uint64_t n_cells, uint32_t thr, uint32_t *d_trip,
                       uint64_t capacity, void *work, hipStream_t s) {
  const uint64_t blocks = emit_blocks(n_cells);
  if (!blocks) return 0;
  HD_REQUIRE(blocks < (1ull << 31), "triangle too large for the emit grid");
  const int vec = (reinterpret_cast<uintptr_t>(d_cells) % 16) == 0 ? 1 : 0;
  const uint64_t chunks = emit_chunks(blocks);
  uint64_t *d_chunks = static_cast<uint64_t *>(work);
  uint32_t *d_counts = reinterpret_cast<uint32_t *>(d_chunks + chunks + 1);
  hipLaunchKernelGGL((k_emit_count<T>),
                     dim3((uint32_t)((blocks + kCountBlocksPerGroup - 1) / kCountBlocksPerGroup)),
                     dim3(256), 0, s, d_cells, n_cells, thr, vec, blocks, d_counts);
  HD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_emit_scan_local, dim3((uint32_t)chunks), dim3(256), 0, s, d_counts,
                     blocks, d_chunks);
  HD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_emit_scan_chunks, dim3(1), dim3(256), 0, s, d_chunks, chunks);
  HD_HIP(hipGetLastError());
  uint64_t total = 0;
  HD_HIP(hipMemcpyAsync(&total, d_chunks + chunks, sizeof total, hipMemcpyDeviceToHost, s));
  HD_HIP(hipStreamSynchronize(s));
  if (total && capacity) {
    const int vec_out = (reinterpret_cast<uintptr_t>(d_trip) % 16) == 0 ? 1 : 0;
    hipLaunchKernelGGL((k_emit_write<T>), dim3((uint32_t)blocks), dim3(256), 0, s, d_cells,
                       n_cells, thr, vec, d_counts, d_chunks, capacity, vec_out, d_trip);
    HD_HIP(hipGetLastError());
    HD_HIP(hipStreamSynchronize(s));
  }
  return total;
}

// ---------------------------------------------------------------------------
// library stream
// ---------------------------------------------------------------------------
hipStream_t lib_stream() {
  static std::once_flag once;
  static hipStream_t s = nullptr;
  std::call_once(once, [] {
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
  });
  if (!s) throw Error(HD_ERR_HIP, "could not create the library stream");
  return s;
}

// A second library stream for downloads that overlap work on lib_stream().
hipStream_t lib_copy_stream() {
  static std::once_flag once;
  static hipStream_t s = nullptr;
  std::call_once(once, [] {
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
  });
  if (!s) throw Error(HD_ERR_HIP, "could not create the library copy stream");
  return s;
}

// Host-side copy spread over up to 16 threads (a fresh destination takes
// its page faults in parallel).
static void parallel_copy(uint8_t *dst, const uint8_t *src, uint64_t len) {
  const unsigned threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < threads; ++t)
    pool.emplace_back([=] {
      const uint64_t b = len * t / threads, e = len * (t + 1) / threads;
      std::memcpy(dst + b, src + b, e - b);
    });
  std::memcpy(dst, src, len / threads);
  for (auto &th : pool) th.join();
}

// Two pinned 64 MiB staging buffers and their events, one set per direction
// (each guarded by its own mutex: one large copy at a time shares it).  Held
// until the process ends (never freed, so no HIP call runs during static
// destruction).
constexpr uint64_t kStage = 64ull << 20;
struct Staging {
  std::mutex mu;
  PinnedBuffer *buf = nullptr;
  Event *done = nullptr;
  void ensure() {
    if (buf) return;
    auto *b = new PinnedBuffer[2];
    b[0].alloc(kStage);
    b[1].alloc(kStage);
    done = new Event[2];
    buf = b;
  }
};

// The first `end` bytes of a device buffer are final once `ev` has fired.
struct ReadyMark {
  uint64_t end;
  hipEvent_t ev;
};

// Copies `bytes` of device memory (ordered after the work on `s`, and, with
// `marks`, each chunk after the first mark that covers it) into pageable
// host memory.  Large copies go through the pinned staging buffers: the DMA
// of chunk c + 1 overlaps the copy of chunk c out of its staging buffer,
// which several host threads share (they also take the page faults of a
// freshly allocated destination in parallel).
static void download(void *host, const void *dev, uint64_t bytes, hipStream_t s,
                     const std::vector<ReadyMark> *marks = nullptr) {
  const TraceRange trace("download");
  if (bytes <= kStage) {
    if (marks && !marks->empty()) HD_HIP(hipStreamWaitEvent(s, marks->back().ev, 0));
    HD_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, s));
    HD_HIP(hipStreamSynchronize(s));
    return;
  }
  static Staging st;
  std::lock_guard<std::mutex> lock(st.mu);
  st.ensure();
  const uint64_t chunks = (bytes + kStage - 1) / kStage;
  size_t mark = 0;
  auto enqueue = [&](uint64_t c) {
    const uint64_t off = c * kStage, len = std::min(kStage, bytes - off);
    if (marks) {
      while (mark < marks->size() && (*marks)[mark].end < off + len) ++mark;
      if (mark < marks->size()) HD_HIP(hipStreamWaitEvent(s, (*marks)[mark].ev, 0));
    }
    HD_HIP(hipMemcpyAsync(st.buf[c & 1].get(), static_cast<const uint8_t *>(dev) + off, len,
                          hipMemcpyDeviceToHost, s));
    HD_HIP(hipEventRecord(st.done[c & 1].get(), s));
  };
  enqueue(0);
  for (uint64_t c = 0; c < chunks; ++c) {
    if (c + 1 < chunks) enqueue(c + 1); // its buffer was emptied in step c - 1
    HD_HIP(hipEventSynchronize(st.done[c & 1].get()));
    const uint64_t off = c * kStage, len = std::min(kStage, bytes - off);
    parallel_copy(static_cast<uint8_t *>(host) + off, st.buf[c & 1].get<uint8_t>(), len);
  }
}

// Copies `bytes` of pageable host memory to the device, ordered on `s`.
// Large copies go through the pinned staging buffers: several host threads
// fill chunk c + 1 while the DMA of chunk c runs.  Returns once every chunk
// has reached the device.
static void upload(void *dev, const void *host, uint64_t bytes, hipStream_t s) {
  const TraceRange trace("upload");
  if (bytes <= kStage) {
    HD_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, s));
    return;
  }
  static Staging st;
  std::lock_guard<std::mutex> lock(st.mu);
  st.ensure();
  const uint64_t chunks = (bytes + kStage - 1) / kStage;
  for (uint64_t c = 0; c < chunks; ++c) {
    const int slot = (int)(c & 1);
    if (c >= 2) HD_HIP(hipEventSynchronize(st.done[slot].get())); // chunk c - 2 has left
    const uint64_t off = c * kStage, len = std::min(kStage, bytes - off);
    parallel_copy(st.buf[slot].get<uint8_t>(), static_cast<const uint8_t *>(host) + off, len);
    HD_HIP(hipMemcpyAsync(static_cast<uint8_t *>(dev) + off, st.buf[slot].get(), len,
                          hipMemcpyHostToDevice, s));
    HD_HIP(hipEventRecord(st.done[slot].get(), s));
  }
  HD_HIP(hipEventSynchronize(st.done[0].get()));
  HD_HIP(hipEventSynchronize(st.done[1].get()));
}

// ---------------------------------------------------------------------------
// small datasets: one wavefront per pair, straight from the packed rows
// ---------------------------------------------------------------------------
// For a few hundred rows the tile kernels are bound by latency, not work:
// one wavefront walks a row's words in turn (n = 3, L = 29,903: about
// 0.3 ms of a 0.45-0.59 ms call, with the mask and encode before it).  Here
// each wavefront takes one pair and its lanes split the row's 32-bit words
// (8 nibble sites each), every load of a lane issued before the first is
// used, then a wavefront sum.  The per-site rules are the semantics'
// definitions on the nibble codes (include/hammingdist_amd.h), evaluated on
// the 8 nibbles of a word at once (bit 4k of a mask stands for nibble k).
constexpr uint32_t kNib0 = 0x11111111u;

// Nibbles of v that are not zero.
__device__ __forceinline__ uint32_t nib_any(uint32_t v) {
  return (v | (v >> 1) | (v >> 2) | (v >> 3)) & kNib0;
}

// Nibbles of v that are 15 ('-').
__device__ __forceinline__ uint32_t nib_gap(uint32_t v) {
  return v & (v >> 1) & (v >> 2) & (v >> 3) & kNib0;
}

// The sparse class of each nibble: the letters (1, 2, 4, 8), '-' (15) and,
// with x_class, X (3) keep their code; every other code becomes the one
// invalid class 0.
__device__ __forceinline__ uint32_t sparse_class(uint32_t v, bool x_class) {
  const uint32_t b0 = v & kNib0, b1 = (v >> 1) & kNib0, b2 = (v >> 2) & kNib0,
                 b3 = (v >> 3) & kNib0;
  const uint32_t any = b0 | b1 | b2 | b3;
  const uint32_t two = (b0 & b1) | (b0 & b2) | (b0 & b3) | (b1 & b2) | (b1 & b3) | (b2 & b3);
  uint32_t keep = (any & ~two) | (b0 & b1 & b2 & b3);
  if (x_class) keep |= b0 & b1 & ~b2 & ~b3;
  return v & (keep * 15u);
}

template <int SEM>
__device__ __forceinline__ uint32_t pair_word_count(uint32_t a, uint32_t b) {
  if constexpr (SEM == HD_SEM_DENSE) return __builtin_popcount(~nib_any(a & b) & kNib0);
  const uint32_t no_gap = ~(nib_gap(a) | nib_gap(b)) & kNib0;
  if constexpr (SEM == HD_SEM_CHAR_X)
    return __builtin_popcount((~nib_any(a & b) & kNib0) | (nib_any(a ^ b) & no_gap));
  const bool x = SEM == HD_SEM_SPARSE_X;
  return __builtin_popcount(nib_any(sparse_class(a, x) ^ sparse_class(b, x)) & no_gap);
}

template <int SEM, typename OutT>
__global__ __launch_bounds__(256) void k_pairs_small(const uint32_t *__restrict__ packed,
                                                     uint64_t words, uint64_t pairs,
                                                     uint32_t max_dist, OutT *__restrict__ out) {
  const uint64_t t = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= pairs) return; // the whole wavefront
  const uint32_t lane = threadIdx.x & 63;
  uint64_t i = (uint64_t)((1.0 + sqrt(1.0 + 8.0 * (double)t)) * 0.5);
  while (i * (i - 1) / 2 > t) --i;
  while ((i + 1) * i / 2 <= t) ++i;
  const uint64_t j = t - i * (i - 1) / 2;
  const uint32_t *ra = packed + i * words, *rb = packed + j * words;
  constexpr int kIn = 8; // words per lane in flight
  uint32_t acc = 0;
  for (uint64_t w0 = lane; w0 < words; w0 += 64 * kIn) {
    uint32_t a[kIn], b[kIn];
#pragma unroll
    for (int k = 0; k < kIn; ++k) {
      const uint64_t w = min(w0 + 64 * k, words - 1);
      a[k] = ra[w];
      b[k] = rb[w];
    }
#pragma unroll
    for (int k = 0; k < kIn; ++k)
      if (w0 + 64 * k < words) acc += pair_word_count<SEM>(a[k], b[k]);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if (lane == 0) out[t] = (OutT)min(acc, max_dist);
}

template <typename OutT>
static void launch_pairs_small_t(int semantics, const uint8_t *d_packed, uint64_t n,
                                 uint64_t bpr, uint32_t max_dist, OutT *d_out, hipStream_t s) {
  const uint64_t pairs = tri_cells(n);
  const dim3 grid((uint32_t)((pairs + 3) / 4)), block(256);
  const uint32_t *p = reinterpret_cast<const uint32_t *>(d_packed);
  const uint64_t words = bpr / 4;
  switch (semantics) {
  case HD_SEM_DENSE:
    hipLaunchKernelGGL((k_pairs_small<HD_SEM_DENSE, OutT>), grid, block, 0, s, p, words, pairs,
                       max_dist, d_out);
    break;
  case HD_SEM_SPARSE:
    hipLaunchKernelGGL((k_pairs_small<HD_SEM_SPARSE, OutT>), grid, block, 0, s, p, words, pairs,
                       max_dist, d_out);
    break;
  case HD_SEM_SPARSE_X:
    hipLaunchKernelGGL((k_pairs_small<HD_SEM_SPARSE_X, OutT>), grid, block, 0, s, p, words,
                       pairs, max_dist, d_out);
    break;
  case HD_SEM_CHAR_X:
    hipLaunchKernelGGL((k_pairs_small<HD_SEM_CHAR_X, OutT>), grid, block, 0, s, p, words, pairs,
                       max_dist, d_out);
    break;
  default:
    throw Error(HD_ERR_INVALID, "unknown semantics");
  }
  HD_HIP(hipGetLastError());
}

// Datasets the one-pair-per-wavefront kernel takes: at most 512 rows
// (HD_SMALL_PAIRS_MAX_ROWS overrides; 0 sends every dataset to the tile
// kernels) and 1 GiB of row bytes over all pairs (each pair reads both rows,
// from L2).
static bool small_pairs(uint64_t n, uint64_t bpr) {
  uint64_t max_rows = 512;
  if (const char *e = std::getenv("HD_SMALL_PAIRS_MAX_ROWS")) max_rows = std::strtoull(e, nullptr, 10);
  return n <= max_rows && tri_cells(n) * bpr <= (1ull << 30);
}

static void launch_pairs_small(int semantics, const uint8_t *d_packed, uint64_t n,
                               uint64_t bpr, int out_bytes, uint32_t max_dist, void *d_out,
                               hipStream_t s) {
  switch (out_bytes) {
  case 1:
    return launch_pairs_small_t(semantics, d_packed, n, bpr, max_dist,
                                static_cast<uint8_t *>(d_out), s);
  case 2:
    return launch_pairs_small_t(semantics, d_packed, n, bpr, max_dist,
                                static_cast<uint16_t *>(d_out), s);
  default:
    return launch_pairs_small_t(semantics, d_packed, n, bpr, max_dist,
                                static_cast<uint32_t *>(d_out), s);
  }
}

// Uploads n packed rows and builds their planes (plus the nibble mask).
struct Encoded {
  Geometry g;
  DeviceBuffer packed, planes, mask;
};

static void upload_encode(Encoded &e, const uint8_t *packed, uint64_t n,
                          uint64_t bpr, int semantics, hipStream_t s) {
  const TraceRange trace("upload + encode");
  geometry(n, bpr, HD_SEM_DENSE, 0); // argument checks
  e.packed.alloc(n * bpr);
  upload(e.packed.get(), packed, n * bpr, s);
  // The nibble mask picks the plane code (also for the dense semantics,
  // whose 3-op code needs the data to hold no X or other codes).
  uint32_t mask = 0xFFFFu;
  {
    e.mask.alloc(sizeof(uint32_t));
    launch_nibble_mask(e.packed.get<uint8_t>(), n * bpr, e.mask.get<uint32_t>(), s);
    HD_HIP(hipMemcpyAsync(&mask, e.mask.get(), 4, hipMemcpyDeviceToHost, s));
    HD_HIP(hipStreamSynchronize(s));
  }
  e.g = geometry(n, bpr, semantics, mask);
  e.planes.alloc(e.g.total_words() * 4);
  launch_encode(e.g, e.packed.get<uint8_t>(), e.planes.get<uint32_t>(), s);
}

} // namespace hd

using namespace hd;

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char *hd_last_error(void) { return g_last_error.c_str(); }

int hd_device_count(void) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess) return 0;
  return count;
}

int hd_device_arch(char *buf, size_t buf_len) {
  return guarded([&] {
    ensure_device();
    HD_REQUIRE(buf && buf_len > 0, "buffer required");
    hipDeviceProp_t prop;
    HD_HIP(hipGetDeviceProperties(&prop, 0));
    std::snprintf(buf, buf_len, "%s", prop.gcnArchName);
  });
}

int hd_variant_count(void) { return kNumVariants; }

int hd_variant_info(int variant, int *rows_per_lane, int *cols_per_wave,
                    int *waves, int *planes) {
  return guarded([&] {
    HD_REQUIRE(variant >= 0 && variant < kNumVariants, "unknown variant");
    if (rows_per_lane) *rows_per_lane = kVariants[variant].R;
    if (cols_per_wave) *cols_per_wave = kVariants[variant].C;
    if (waves) *waves = kVariants[variant].NW;
    if (planes) *planes = 0;
  });
}

int hd_planes_layout(uint64_t n, uint64_t bytes_per_seq, int semantics,
                     uint32_t nibble_mask, uint64_t *n_pad, uint32_t *n_words,
                     uint32_t *n_planes, uint64_t *total_words) {
  return guarded([&] {
    const Geometry g = geometry(n, bytes_per_seq, semantics, nibble_mask);
    if (n_pad) *n_pad = g.n_pad;
    if (n_words) *n_words = g.W;
    if (n_planes) *n_planes = (uint32_t)g.code.planes;
    if (total_words) *total_words = g.total_words();
  });
}

int hd_default_variant(uint64_t n, uint64_t bytes_per_seq, int semantics,
                       uint32_t nibble_mask, int early_exit, uint32_t max_dist,
                       int *variant) {
  return guarded([&] {
    HD_REQUIRE(variant, "variant pointer required");
    const Geometry g = geometry(n, bytes_per_seq, semantics, nibble_mask);
    *variant = default_variant(g, early_exit != 0, max_dist);
  });
}

int hd_nibble_mask_dev(const uint8_t *d_packed, uint64_t n, uint64_t bytes_per_seq,
                       uint32_t *mask, void *stream) {
  return guarded([&] {
    ensure_device();
    HD_REQUIRE(mask, "mask pointer required");
    hipStream_t s = (hipStream_t)stream;
    DeviceBuffer d_mask(4);
    launch_nibble_mask(d_packed, n * bytes_per_seq, d_mask.get<uint32_t>(), s);
    HD_HIP(hipMemcpyAsync(mask, d_mask.get(), 4, hipMemcpyDeviceToHost, s));
    HD_HIP(hipStreamSynchronize(s));
  });
}

int hd_encode_dev(const uint8_t *d_packed, uint64_t n, uint64_t bytes_per_seq,
                  int semantics, uint32_t nibble_mask, uint32_t *d_planes,
                  void *stream) {
  return guarded([&] {
    const TraceRange trace("hd_encode_dev");
    ensure_device();
    const Geometry g = geometry(n, bytes_per_seq, semantics, nibble_mask);
    launch_encode(g, d_packed, d_planes, (hipStream_t)stream);
  });
}

int hd_all_pairs_dev(const uint32_t *d_planes, uint64_t n, uint64_t bytes_per_seq,
                     int semantics, uint32_t nibble_mask, int out_bytes,
                     uint32_t max_dist, int early_exit, int variant, void *d_out,
                     void *stream) {
  return guarded([&] {
    const TraceRange trace("hd_all_pairs_dev");
    ensure_device();
    check_max_dist(out_bytes, max_dist);
    const Geometry g = geometry(n, bytes_per_seq, semantics, nibble_mask);
    if (n < 2) return;
    const int v = variant < 0 ? default_variant(g, early_exit != 0, max_dist) : variant;
    HD_REQUIRE(v < kNumVariants, "unknown variant");
    const uint64_t TI = tile_rows(v);
    const uint64_t nb = (n + TI - 1) / TI;
    run_allpairs(g, d_planes, out_bytes, max_dist, early_exit != 0, v, 0, nb,
                 d_out, 0, (hipStream_t)stream);
  });
}

int hd_all_pairs_dev_rows(const uint32_t *d_planes, uint64_t n, uint64_t bytes_per_seq,
                          int semantics, uint32_t nibble_mask, int out_bytes,
                          uint32_t max_dist, int early_exit, int variant, uint64_t row_begin,
                          uint64_t row_end, void *d_out, void *stream) {
  return guarded([&] {
    const TraceRange trace("hd_all_pairs_dev_rows");
    ensure_device();
    check_max_dist(out_bytes, max_dist);
    const Geometry g = geometry(n, bytes_per_seq, semantics, nibble_mask);
    HD_REQUIRE(row_begin <= row_end && row_end <= n, "row band outside the dataset");
    if (n < 2 || row_end <= std::max<uint64_t>(row_begin, 1)) return;
    const int v = variant < 0 ? default_variant(g, early_exit != 0, max_dist) : variant;
    HD_REQUIRE(v < kNumVariants, "unknown variant");
    const uint64_t TI = tile_rows(v);
    HD_REQUIRE(row_begin % TI == 0 && (row_end % TI == 0 || row_end == n),
               "row band must start and end on a multiple of the tile's " +
                   std::to_string(TI) + " rows (or end at n)");
    const uint64_t first = row_begin ? row_begin * (row_begin - 1) / 2 : 0;
    run_allpairs(g, d_planes, out_bytes, max_dist, early_exit != 0, v, row_begin / TI,
                 (row_end + TI - 1) / TI, d_out, first, (hipStream_t)stream);
  });
}

int hd_all_pairs(const uint8_t *packed, uint64_t n, uint64_t bytes_per_seq,
                 int semantics, int out_bytes, uint32_t max_dist, void *host_out) {
  return guarded([&] {
    const TraceRange trace("hd_all_pairs");
    check_max_dist(out_bytes, max_dist);
    if (n < 2) return;
    HD_REQUIRE(packed && host_out, "null buffer");
    ensure_device();
    hipStream_t s = lib_stream();
    if (small_pairs(n, bytes_per_seq)) {
      geometry(n, bytes_per_seq, semantics, 0); // argument checks
      DeviceBuffer d_packed(n * bytes_per_seq), d_cells(tri_cells(n) * out_bytes);
      upload(d_packed.get(), packed, n * bytes_per_seq, s);
      launch_pairs_small(semantics, d_packed.get<uint8_t>(), n, bytes_per_seq, out_bytes,
                         max_dist, d_cells.get(), s);
      download(host_out, d_cells.get(), tri_cells(n) * out_bytes, s);
      return;
    }
    Encoded e;
    upload_encode(e, packed, n, bytes_per_seq, semantics, s);
    const uint64_t cells = tri_cells(n);
    DeviceBuffer d_out(cells * out_bytes);
    const int v = default_variant(e.g, true, max_dist);
    const uint64_t TI = tile_rows(v);
    const uint64_t nb = (n + TI - 1) / TI;
    // The triangle is computed in row bands of about 1/8 of it (at least
    // one staging chunk, at most 512 MiB, so the transfer left after the
    // last band stays short), each launched on its own, so the download of a
    // band's cells overlaps the computation of the next ones.  With early-
    // exit votes and a small cap (below 64) the kernel is short and one
    // launch followed by the download is faster (50 vs 57 ms for 50,000 x
    // 29,903 at max_distance 2; banded 143 vs 159 ms at 255).
    const bool short_kernel = check_interval(true, max_dist, e.g.W) < e.g.W && max_dist < 64;
    const uint64_t band_bytes =
        short_kernel ? cells * out_bytes
                     : std::max(kStage, std::min<uint64_t>(cells * out_bytes / 8, 8 * kStage));
    std::vector<std::unique_ptr<Event>> evs;
    std::vector<ReadyMark> marks;
    for (uint64_t rb = 0; rb < nb;) {
      const uint64_t c0 = tri_cells(std::min(n, rb * TI));
      uint64_t rb1 = rb + 1;
      while (rb1 < nb && (tri_cells(std::min(n, (rb1 + 1) * TI)) - c0) * out_bytes <= band_bytes)
        ++rb1;
      run_allpairs(e.g, e.planes.get<uint32_t>(), out_bytes, max_dist, true, v, rb, rb1,
                   d_out.get(), 0, s);
      evs.push_back(std::make_unique<Event>());
      HD_HIP(hipEventRecord(evs.back()->get(), s));
      marks.push_back({tri_cells(std::min(n, rb1 * TI)) * out_bytes, evs.back()->get()});
      rb = rb1;
    }
    if (short_kernel) download(host_out, d_out.get(), cells * out_bytes, s);
    else download(host_out, d_out.get(), cells * out_bytes, lib_copy_stream(), &marks);
  });
}

int hd_all_pairs_stream(const uint8_t *packed, uint64_t n, uint64_t bytes_per_seq,
                        int semantics, int out_bytes, uint32_t max_dist,
                        uint64_t band_cells, hd_band_cb cb, void *user) {
  return guarded([&] {
    const TraceRange trace("hd_all_pairs_stream");
    check_max_dist(out_bytes, max_dist);
    HD_REQUIRE(cb, "callback required");
    if (n < 2) return;
    HD_REQUIRE(packed, "null buffer");
    ensure_device();
    hipStream_t s = lib_stream();
    Stream copy;
    Encoded e;
    upload_encode(e, packed, n, bytes_per_seq, semantics, s);
    const int v = default_variant(e.g, true, max_dist);
    const uint64_t TI = tile_rows(v);
    const uint64_t nb = (n + TI - 1) / TI;
    if (band_cells == 0) band_cells = 64ull << 20;

    // Bands of whole row blocks, each at least one block.
    std::vector<uint64_t> band_rb{0};
    while (band_rb.back() < nb) {
      uint64_t rb = band_rb.back() + 1;
      const uint64_t r0 = band_rb.back() * TI;
      while (rb < nb && tri_cells(std::min(n, (rb + 1) * TI)) - tri_cells(r0) <= band_cells)
        ++rb;
      band_rb.push_back(rb);
    }
    const size_t n_bands = band_rb.size() - 1;
    uint64_t max_band = 0;
    for (size_t b = 0; b < n_bands; ++b)
      max_band = std::max(max_band, tri_cells(std::min(n, band_rb[b + 1] * TI)) -
                                        tri_cells(band_rb[b] * TI));
    const uint64_t bytes = max_band * out_bytes;
    DeviceBuffer dbuf[2];
    PinnedBuffer hbuf[2];
    Event ev_k[2], ev_c[2];
    for (int i = 0; i < 2; ++i) {
      dbuf[i].alloc(bytes);
      hbuf[i].alloc(bytes);
    }
    auto band_range = [&](size_t b, uint64_t &r0, uint64_t &r1) {
      r0 = band_rb[b] * TI;
      r1 = std::min(n, band_rb[b + 1] * TI);
    };
    auto enqueue = [&](size_t b) {
      uint64_t r0, r1;
      band_range(b, r0, r1);
      const int slot = (int)(b & 1);
      run_allpairs(e.g, e.planes.get<uint32_t>(), out_bytes, max_dist, true, v,
                   band_rb[b], band_rb[b + 1], dbuf[slot].get(), tri_cells(r0), s);
      HD_HIP(hipEventRecord(ev_k[slot].get(), s));
      HD_HIP(hipStreamWaitEvent(copy.get(), ev_k[slot].get(), 0));
      HD_HIP(hipMemcpyAsync(hbuf[slot].get(), dbuf[slot].get(),
                            (tri_cells(r1) - tri_cells(r0)) * out_bytes,
                            hipMemcpyDeviceToHost, copy.get()));
      HD_HIP(hipEventRecord(ev_c[slot].get(), copy.get()));
    };
    for (size_t b = 0; b < std::min<size_t>(2, n_bands); ++b) enqueue(b);
    for (size_t b = 0; b < n_bands; ++b) {
      const int slot = (int)(b & 1);
      {
        const TraceRange wait("wait for band");
        HD_HIP(hipEventSynchronize(ev_c[slot].get()));
      }
      uint64_t r0, r1;
      band_range(b, r0, r1);
      const uint64_t c0 = tri_cells(r0), nc = tri_cells(r1) - c0;
      const TraceRange band("band callback");
      if (nc && cb(user, hbuf[slot].get(), c0, nc, r0 == 0 ? 1 : r0, r1) != 0) {
        HD_HIP(hipStreamSynchronize(s));
        HD_HIP(hipStreamSynchronize(copy.get()));
        throw Error(HD_ERR_CALLBACK, "band callback aborted the stream");
      }
      if (b + 2 < n_bands) enqueue(b + 2);
    }
    HD_HIP(hipStreamSynchronize(s));
    HD_HIP(hipStreamSynchronize(copy.get()));
  });
}

// One pair on the reference-distance kernel (one query row, one database
// row): the two rows and the result go through a device buffer and a pinned
// host buffer the library keeps (grown as needed; one call at a time) on a
// stream of their own, so a call costs two copies and one launch.  Through
// the all-pairs path (allocations, mask, encode, one wavefront walking the
// row's words) a pair of 29,903-site rows took 0.5 ms.  Never destroyed: no
// HIP call runs during static destruction.
struct PairScratch {
  std::mutex mu;
  hipStream_t s = nullptr;
  DeviceBuffer dev;
  PinnedBuffer host;
};

static uint32_t pair_distance_refdist(const uint8_t *a, const uint8_t *b, uint64_t bytes,
                                      int semantics) {
  static PairScratch *ps = new PairScratch;
  std::lock_guard<std::mutex> lock(ps->mu);
  if (!ps->s) HD_HIP(hipStreamCreateWithFlags(&ps->s, hipStreamNonBlocking));
  const uint64_t need = 2 * bytes + 8;
  if (ps->dev.size() < need) {
    ps->dev.alloc(need);
    ps->host.alloc(need);
  }
  uint8_t *h = ps->host.get<uint8_t>();
  uint8_t *d = ps->dev.get<uint8_t>();
  std::memcpy(h, a, bytes);
  std::memcpy(h + bytes, b, bytes);
  HD_HIP(hipMemcpyAsync(d, h, 2 * bytes, hipMemcpyHostToDevice, ps->s));
  launch_refdist(d, d + bytes, bytes, nullptr, 1, semantics,
                 reinterpret_cast<uint32_t *>(d + 2 * bytes), ps->s);
  HD_HIP(hipMemcpyAsync(h + 2 * bytes, d + 2 * bytes, 4, hipMemcpyDeviceToHost, ps->s));
  HD_HIP(hipStreamSynchronize(ps->s));
  uint32_t out;
  std::memcpy(&out, h + 2 * bytes, 4);
  return out;
}

int hd_distance_pair(const uint8_t *a, const uint8_t *b, uint64_t bytes,
                     int semantics, int32_t max_dist, int32_t *out) {
  return guarded([&] {
    const TraceRange trace("hd_distance_pair");
    HD_REQUIRE(a && b && out, "null buffer");
    HD_REQUIRE(max_dist >= 0, "max_dist must be >= 0");
    if (semantics == HD_SEM_DENSE || semantics == HD_SEM_CHAR_X) {
      HD_REQUIRE(bytes > 0 && bytes % 8 == 0, "bytes_per_seq must be a positive multiple of 8");
      ensure_device();
      *out = (int32_t)std::min<uint32_t>(pair_distance_refdist(a, b, bytes, semantics),
                                         (uint32_t)max_dist);
      return;
    }
    std::vector<uint8_t> both(2 * bytes);
    std::memcpy(both.data(), a, bytes);
    std::memcpy(both.data() + bytes, b, bytes);
    uint32_t d = 0;
    const int st = hd_all_pairs(both.data(), 2, bytes, semantics, 4,
                                (uint32_t)max_dist, &d);
    if (st != HD_OK) throw Error(st, g_last_error);
    *out = (int32_t)d;
  });
}

int hd_reference_distances_dev(const uint8_t *d_query, const uint8_t *d_db,
                               uint64_t n_db, uint64_t bytes_per_seq,
                               const uint64_t *d_rows, uint64_t m, int semantics,
                               uint32_t *d_out, void *stream) {
  return guarded([&] {
    ensure_device();
    HD_REQUIRE(bytes_per_seq > 0 && bytes_per_seq % 8 == 0,
               "bytes_per_seq must be a positive multiple of 8");
    HD_REQUIRE(d_rows || m <= n_db, "m exceeds the database rows");
    launch_refdist(d_query, d_db, bytes_per_seq, d_rows, m, semantics, d_out,
                   (hipStream_t)stream);
  });
}

int hd_reference_distances(const uint8_t *query, const uint8_t *db, uint64_t n_db,
                           uint64_t bytes_per_seq, const uint64_t *rows, uint64_t m,
                           int semantics, uint32_t *out) {
  return guarded([&] {
    const TraceRange trace("hd_reference_distances");
    HD_REQUIRE(bytes_per_seq > 0 && bytes_per_seq % 8 == 0,
               "bytes_per_seq must be a positive multiple of 8");
    HD_REQUIRE(semantics == HD_SEM_DENSE || semantics == HD_SEM_CHAR_X,
               "reference distances support HD_SEM_DENSE and HD_SEM_CHAR_X");
    if (m == 0) return;
    HD_REQUIRE(query && db && out, "null buffer");
    if (rows) {
      for (uint64_t k = 0; k < m; ++k)
        HD_REQUIRE(rows[k] < n_db, "row index out of range");
    } else {
      HD_REQUIRE(m <= n_db, "m exceeds the database rows");
    }
    ensure_device();
    hipStream_t s = lib_stream();
    DeviceBuffer d_q(bytes_per_seq), d_db(n_db * bytes_per_seq), d_out(m * 4), d_rows;
    HD_HIP(hipMemcpyAsync(d_q.get(), query, bytes_per_seq, hipMemcpyHostToDevice, s));
    upload(d_db.get(), db, n_db * bytes_per_seq, s);
    if (rows) {
      d_rows.alloc(m * 8);
      HD_HIP(hipMemcpyAsync(d_rows.get(), rows, m * 8, hipMemcpyHostToDevice, s));
    }
    launch_refdist(d_q.get<uint8_t>(), d_db.get<uint8_t>(), bytes_per_seq,
                   rows ? d_rows.get<uint64_t>() : nullptr, m, semantics,
                   d_out.get<uint32_t>(), s);
    HD_HIP(hipMemcpyAsync(out, d_out.get(), m * 4, hipMemcpyDeviceToHost, s));
    HD_HIP(hipStreamSynchronize(s));
  });
}

int hd_consensus_diff(const uint8_t *packed, uint64_t n, uint64_t bytes_per_seq,
                      uint64_t L, uint64_t *n_diff, uint32_t *nibble_mask) {
  return guarded([&] {
    const TraceRange trace("hd_consensus_diff");
    HD_REQUIRE(bytes_per_seq > 0 && bytes_per_seq % 8 == 0,
               "bytes_per_seq must be a positive multiple of 8");
    HD_REQUIRE(L <= 2 * bytes_per_seq, "L exceeds the packed row length");
    HD_REQUIRE(n_diff && nibble_mask, "null output");
    *n_diff = 0;
    *nibble_mask = 0;
    if (n == 0 || L == 0) return;
    HD_REQUIRE(packed, "null buffer");
    HD_REQUIRE(n < (1ull << 32), "too many sequences");
    ensure_device();
    hipStream_t s = lib_stream();
    DeviceBuffer d_packed(n * bytes_per_seq), d_counts(L * 16), d_mask(4);
    HD_HIP(hipMemcpyAsync(d_packed.get(), packed, n * bytes_per_seq,
                          hipMemcpyHostToDevice, s));
    HD_HIP(hipMemsetAsync(d_counts.get(), 0, L * 16, s));
    HD_HIP(hipMemsetAsync(d_mask.get(), 0, 4, s));
    const uint64_t words = bytes_per_seq / 4;
    dim3 grid((uint32_t)((words + 255) / 256), (uint32_t)((n + kCountRows - 1) / kCountRows));
    HD_REQUIRE(grid.y <= 65535, "too many sequences for the count grid");
    hipLaunchKernelGGL(k_letter_counts, grid, dim3(256), 0, s,
                       d_packed.get<uint8_t>(), (uint32_t)n, bytes_per_seq, L,
                       d_counts.get<uint32_t>(), d_mask.get<uint32_t>());
    HD_HIP(hipGetLastError());
    std::vector<uint32_t> counts(L * 4);
    HD_HIP(hipMemcpyAsync(counts.data(), d_counts.get(), L * 16, hipMemcpyDeviceToHost, s));
    HD_HIP(hipMemcpyAsync(nibble_mask, d_mask.get(), 4, hipMemcpyDeviceToHost, s));
    HD_HIP(hipStreamSynchronize(s));
    // Consensus letter = first maximum over A, C, G, T; every other entry of
    // the column differs from it (src/hamming_impl.cc:157-166, 177-185).
    uint64_t diff = 0;
    for (uint64_t site = 0; site < L; ++site) {
      const uint32_t *c = &counts[site * 4];
      uint32_t mx = c[0];
      for (int l = 1; l < 4; ++l) mx = std::max(mx, c[l]);
      diff += n - mx;
    }
    *n_diff = diff;
  });
}

// Sparse emit of a device triangle into device triplets with the scratch at
// `work` (emit_scratch_bytes(nc) bytes); returns the total.
static uint64_t emit_dev(const void *d_cells, uint64_t nc, int out_bytes, uint32_t thr,
                         uint32_t *d_trip, uint64_t capacity, void *work, hipStream_t s) {
  if (out_bytes == 1)
    return emit_t(static_cast<const uint8_t *>(d_cells), nc, thr, d_trip, capacity, work, s);
  if (out_bytes == 2)
    return emit_t(static_cast<const uint16_t *>(d_cells), nc, thr, d_trip, capacity, work, s);
  return emit_t(static_cast<const uint32_t *>(d_cells), nc, thr, d_trip, capacity, work, s);
}

// The same with the scratch in a device buffer of this call.
static uint64_t emit_dev_alloc(const void *d_cells, uint64_t nc, int out_bytes, uint32_t thr,
                               uint32_t *d_trip, uint64_t capacity, hipStream_t s) {
  DeviceBuffer work(emit_scratch_bytes(nc));
  return emit_dev(d_cells, nc, out_bytes, thr, d_trip, capacity, work.get(), s);
}

int hd_sparse_emit(const void *cells, uint64_t n, int out_bytes, int64_t threshold,
                   uint32_t *triplets, uint64_t capacity, uint64_t *count) {
  return guarded([&] {
    const TraceRange trace("hd_sparse_emit");
    HD_REQUIRE(count, "null count");
    HD_REQUIRE(out_bytes == 1 || out_bytes == 2 || out_bytes == 4,
               "out_bytes must be 1, 2 or 4");
    *count = 0;
    const uint64_t nc = tri_cells(n);
    if (nc == 0 || threshold < 0) return;
    HD_REQUIRE(cells, "null buffer");
    HD_REQUIRE(triplets || capacity == 0, "null triplets");
    ensure_device();
    const uint32_t thr = threshold > 0xFFFFFFFFll ? 0xFFFFFFFFu : (uint32_t)threshold;
    hipStream_t s = lib_stream();
    DeviceBuffer d_cells(nc * out_bytes);
    upload(d_cells.get(), cells, nc * out_bytes, s);
    DeviceBuffer d_trip(std::max<uint64_t>(capacity, 1) * 12);
    const uint64_t total =
        emit_dev_alloc(d_cells.get(), nc, out_bytes, thr, d_trip.get<uint32_t>(), capacity, s);
    const uint64_t got = std::min(total, capacity);
    if (got) {
      HD_HIP(hipMemcpyAsync(triplets, d_trip.get(), got * 12, hipMemcpyDeviceToHost, s));
      HD_HIP(hipStreamSynchronize(s));
    }
    *count = total;
    if (total > capacity)
      throw Error(HD_ERR_CAPACITY, "more cells below the threshold than capacity");
  });
}

int hd_sparse_emit_dev_noalloc(const void *d_cells, uint64_t n, int out_bytes,
                               int64_t threshold, uint32_t *d_triplets, uint64_t capacity,
                               uint64_t *count, void *stream) {
  return guarded([&] {
    const TraceRange trace("hd_sparse_emit_dev_noalloc");
    HD_REQUIRE(count, "null count");
    HD_REQUIRE(out_bytes == 1 || out_bytes == 2 || out_bytes == 4,
               "out_bytes must be 1, 2 or 4");
    *count = 0;
    const uint64_t nc = tri_cells(n);
    if (nc == 0 || threshold < 0) return;
    HD_REQUIRE(nc <= HD_EMIT_NOALLOC_MAX_CELLS,
               "triangle of " + std::to_string(nc) +
                   " cells exceeds the no-allocation emit's maximum of " +
                   std::to_string((unsigned long long)HD_EMIT_NOALLOC_MAX_CELLS) +
                   " cells (use hd_sparse_emit_dev)");
    HD_REQUIRE(d_cells, "null buffer");
    HD_REQUIRE(d_triplets || capacity == 0, "null triplets");
    ensure_device();
    const uint32_t thr = threshold > 0xFFFFFFFFll ? 0xFFFFFFFFu : (uint32_t)threshold;
    const uint64_t total = emit_dev(d_cells, nc, out_bytes, thr, d_triplets, capacity,
                                    noalloc_scratch(), (hipStream_t)stream);
    *count = total;
    if (total > capacity)
      throw Error(HD_ERR_CAPACITY, "more cells below the threshold than capacity");
  });
}

int hd_sparse_emit_dev(const void *d_cells, uint64_t n, int out_bytes, int64_t threshold,
                       uint32_t *d_triplets, uint64_t capacity, uint64_t *count,
                       void *stream) {
  return guarded([&] {
    const TraceRange trace("hd_sparse_emit_dev");
    HD_REQUIRE(count, "null count");
    HD_REQUIRE(out_bytes == 1 || out_bytes == 2 || out_bytes == 4,
               "out_bytes must be 1, 2 or 4");
    *count = 0;
    const uint64_t nc = tri_cells(n);
    if (nc == 0 || threshold < 0) return;
    HD_REQUIRE(d_cells, "null buffer");
    HD_REQUIRE(d_triplets || capacity == 0, "null triplets");
    ensure_device();
    const uint32_t thr = threshold > 0xFFFFFFFFll ? 0xFFFFFFFFu : (uint32_t)threshold;
    const uint64_t total = emit_dev_alloc(d_cells, nc, out_bytes, thr, d_triplets, capacity,
                                    (hipStream_t)stream);
    *count = total;
    if (total > capacity)
      throw Error(HD_ERR_CAPACITY, "more cells below the threshold than capacity");
  });
}

} // extern "C"
