// Batched all-pairs: many independent datasets of uneven n and L in one
// persistent triangular-tile launch, captured with the batched encode in a
// hipGraph.  C ABI: hd_batch_* in include/hammingdist_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "engine_common.h"
#include "kernels.h"

namespace hd {

// Tile shapes of the batched kernel (rows per lane R, columns per wave C; 4
// waves per workgroup, so a tile is 64 R rows by 4 C columns).  One launch
// uses one shape for every job; the autotuner picks it per batch.
constexpr int kBatchWaves = 4;
struct Shape {
  int R, C;
};
constexpr Shape kShapes[] = {{4, 16}, {4, 8}, {2, 8}, {2, 16}, {1, 8}};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

// One dataset of a batch, as the kernels see it.
struct BatchJobDev {
  const uint32_t *planes;
  const uint32_t *corr; // correction counts (codes with has_corr), else null
  void *out;
  uint64_t tile_begin; // first index of this job in the batch's tile list
  uint32_t n, n_pad, W, P, out_bytes, max_dist, check_every;
};

struct EncodeJobDev {
  const uint8_t *packed;
  uint32_t *planes;
  uint32_t *corr;       // correction counts (codes with has_corr), else null
  uint64_t bpr;
  uint64_t block_begin; // first encode block of this job
  uint64_t count_begin; // first totals block (256 rows) of this job
  uint32_t n, n_pad, W, P, runs, mode;
  uint32_t op; // plane code (kernels.h), for the correction counts
  Table16 tbl;
};

// Index of the job whose range [begin[k], begin[k+1]) holds x (begin sorted).
// x is wave-uniform; the result is made explicitly uniform so that the job's
// fields are scalar loads and everything derived from them stays in SGPRs.
template <typename Job>
__device__ __forceinline__ uint32_t find_job(const Job *__restrict__ jobs, uint32_t n_jobs,
                                             uint64_t x, uint64_t Job::*begin) {
  uint32_t lo = 0, hi = n_jobs;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (jobs[mid].*begin <= x) lo = mid;
    else hi = mid;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// Encode tiles of every job (kernels.h, encode_tile): workgroup b of job j
// takes its tile (row group, word block) from b - block_begin.
__global__ __launch_bounds__(256) void k_batch_encode(const EncodeJobDev *__restrict__ jobs,
                                                      uint32_t n_jobs) {
  __shared__ uint8_t s_tbl[16];
  __shared__ __attribute__((aligned(16))) uint8_t s_rows[kEncodeLdsBytes];
  const uint32_t j = find_job(jobs, n_jobs, blockIdx.x, &EncodeJobDev::block_begin);
  const EncodeJobDev &J = jobs[j];
  if (threadIdx.x < 16) s_tbl[threadIdx.x] = J.tbl.v[threadIdx.x];
  const uint64_t local = blockIdx.x - J.block_begin;
  const uint32_t groups = J.n_pad / kWave;
  encode_tile(J.packed, J.n, J.bpr, J.W, (int)J.P, (int)J.op, s_tbl, (int)J.mode, J.planes,
              J.corr, J.n_pad, (uint32_t)(local % groups),
              (uint32_t)(local / groups) * kEncodeWords, s_rows);
}

// Whole-row correction counts of every job whose code has them, from the
// chunk counts its encode tiles wrote: one thread per row, 256 rows per
// workgroup.
__global__ __launch_bounds__(256) void k_batch_totals(const EncodeJobDev *__restrict__ jobs,
                                                      uint32_t n_jobs) {
  const uint32_t j = find_job(jobs, n_jobs, blockIdx.x, &EncodeJobDev::count_begin);
  const EncodeJobDev &J = jobs[j];
  const uint32_t row = (uint32_t)(blockIdx.x - J.count_begin) * 256 + threadIdx.x;
  corr_totals_row(J.corr, J.W, J.n_pad, (int)J.op, row);
}

__global__ void k_reset_counter(unsigned long long *counter) { *counter = 0; }

// Persistent workgroups take tiles from the batch's list (largest per-tile
// work first) through one atomic counter until the list is exhausted; each
// tile is computed with its job's plane code and cell type.  ONLY (kDense3
// or 4): every job has that plane code, so only its tile loop is compiled in
// and the kernel keeps the register budget of the plain kernel; ONLY = 0
// handles any mix.
template <int R, int C, int ONLY>
__global__ __launch_bounds__(kWave *kBatchWaves) void k_batch_allpairs(
    const BatchJobDev *__restrict__ jobs, uint32_t n_jobs, uint64_t total_tiles,
    unsigned long long *__restrict__ counter) {
  constexpr uint32_t TI = kWave * R;
  constexpr uint32_t TJ = kBatchWaves * C;
  static_assert(TI % TJ == 0, "tile rows must be a multiple of tile columns");
  __shared__ unsigned long long s_t;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  for (;;) {
    if (threadIdx.x == 0) s_t = atomicAdd(counter, 1ull);
    __syncthreads();
    const uint64_t t = uniform64(s_t);
    __syncthreads();
    if (t >= total_tiles) break;
    const uint32_t j = find_job(jobs, n_jobs, t, &BatchJobDev::tile_begin);
    const BatchJobDev &J = jobs[j];
    uint32_t b, cb;
    tile_coords(t - J.tile_begin, TI / TJ, b, cb);
    const uint32_t i0 = b * TI, j0 = cb * TJ + wave * C;
    // 5 planes with 16 columns per wave would need more column words than
    // the SGPR file holds: such tiles run as two passes of 8 columns.
    constexpr int CW = C > 8 ? 8 : C;
#define HD_BATCH_TILE(PP, CC, JJ)                                                        \
  tile_wave<PP, R, CC>(J.planes, J.n, J.W, i0, JJ, J.out, J.out_bytes, 0, J.max_dist,     \
                       J.check_every, J.corr, J.n_pad);
    switch (ONLY ? (uint32_t)ONLY : J.P) {
    case kDense3:
      HD_BATCH_TILE(kDense3, C, j0)
      break;
    case 4:
      HD_BATCH_TILE(4, C, j0)
      break;
    case kGap2:
      HD_BATCH_TILE(kGap2, C, j0)
      break;
    case kGap3:
      HD_BATCH_TILE(kGap3, C, j0)
      break;
    case kClass3:
      HD_BATCH_TILE(kClass3, C, j0)
      break;
    default:
      for (uint32_t h = 0; h < (uint32_t)(C / CW); ++h) {
        HD_BATCH_TILE(5, CW, j0 + h * CW)
      }
      break;
    }
#undef HD_BATCH_TILE
  }
}

// The same tiles as one grid: workgroup g computes tile g of the batch's list
// (the hardware dispatcher balances the load; the list is in LPT order).
template <int R, int C, int ONLY>
__global__ __launch_bounds__(kWave *kBatchWaves) void k_batch_grid(
    const BatchJobDev *__restrict__ jobs, uint32_t n_jobs, uint64_t tile_begin) {
  constexpr uint32_t TI = kWave * R;
  constexpr uint32_t TJ = kBatchWaves * C;
  const uint64_t t = tile_begin + blockIdx.x;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const uint32_t j = find_job(jobs, n_jobs, t, &BatchJobDev::tile_begin);
  const BatchJobDev &J = jobs[j];
  uint32_t b, cb;
  tile_coords(t - J.tile_begin, TI / TJ, b, cb);
  const uint32_t i0 = b * TI, j0 = cb * TJ + wave * C;
  constexpr int CW = C > 8 ? 8 : C;
  const uint32_t P = ONLY ? (uint32_t)ONLY : J.P;
  if (P == kDense3) {
    tile_wave<kDense3, R, C>(J.planes, J.n, J.W, i0, j0, J.out, J.out_bytes, 0, J.max_dist,
                             J.check_every, J.corr, J.n_pad);
  } else if (P == 4) {
    tile_wave<4, R, C>(J.planes, J.n, J.W, i0, j0, J.out, J.out_bytes, 0, J.max_dist,
                       J.check_every, nullptr, J.n_pad);
  } else if (P == kGap2) {
    tile_wave<kGap2, R, C>(J.planes, J.n, J.W, i0, j0, J.out, J.out_bytes, 0, J.max_dist,
                           J.check_every, J.corr, J.n_pad);
  } else if (P == kGap3) {
    tile_wave<kGap3, R, C>(J.planes, J.n, J.W, i0, j0, J.out, J.out_bytes, 0, J.max_dist,
                           J.check_every, J.corr, J.n_pad);
  } else if (P == kClass3) {
    tile_wave<kClass3, R, C>(J.planes, J.n, J.W, i0, j0, J.out, J.out_bytes, 0, J.max_dist,
                             J.check_every, nullptr, J.n_pad);
  } else {
    for (uint32_t h = 0; h < (uint32_t)(C / CW); ++h)
      tile_wave<5, R, CW>(J.planes, J.n, J.W, i0, j0 + h * CW, J.out, J.out_bytes, 0,
                          J.max_dist, J.check_every, nullptr, J.n_pad);
  }
}

template <int V> static const void *grid_kernel_v(int only) {
  constexpr int R = kShapes[V].R, C = kShapes[V].C;
  switch (only) {
  case kDense3: return reinterpret_cast<const void *>(&k_batch_grid<R, C, kDense3>);
  case 4: return reinterpret_cast<const void *>(&k_batch_grid<R, C, 4>);
  default: return reinterpret_cast<const void *>(&k_batch_grid<R, C, 0>);
  }
}

static const void *grid_kernel_of(int v, int only) {
  switch (v) {
  case 0: return grid_kernel_v<0>(only);
  case 1: return grid_kernel_v<1>(only);
  case 2: return grid_kernel_v<2>(only);
  case 3: return grid_kernel_v<3>(only);
  default: return grid_kernel_v<4>(only);
  }
}

template <int V> static const void *batch_kernel_v(int only) {
  constexpr int R = kShapes[V].R, C = kShapes[V].C;
  switch (only) {
  case kDense3: return reinterpret_cast<const void *>(&k_batch_allpairs<R, C, kDense3>);
  case 4: return reinterpret_cast<const void *>(&k_batch_allpairs<R, C, 4>);
  default: return reinterpret_cast<const void *>(&k_batch_allpairs<R, C, 0>);
  }
}

static const void *batch_kernel_of(int v, int only) {
  switch (v) {
  case 0: return batch_kernel_v<0>(only);
  case 1: return batch_kernel_v<1>(only);
  case 2: return batch_kernel_v<2>(only);
  case 3: return batch_kernel_v<3>(only);
  default: return batch_kernel_v<4>(only);
  }
}

// Batch variant v: shape v % kNumShapes; persistent workgroups with an
// atomic tile queue for v < kNumShapes, one workgroup per tile above.
static constexpr int kNumBatchVariants = 2 * kNumShapes;

} // namespace hd

using namespace hd;

// One batch: per-job device buffers, the job tables and the captured graph.
struct hd_batch {
  struct Job {
    uint64_t n = 0, bpr = 0;
    int semantics = HD_SEM_DENSE, out_bytes = 1;
    uint32_t max_dist = 0, mask = 0xFFFF;
    Geometry g;
    DeviceBuffer packed, planes, out;
  };
  std::vector<std::unique_ptr<Job>> jobs;
  DeviceBuffer d_jobs, d_enc, d_counter;
  std::vector<BatchJobDev> host_jobs;
  uint32_t n_jobs_active = 0;
  uint64_t total_tiles = 0, enc_blocks = 0, count_blocks = 0;
  int variant = -1, early_exit = 1;
  int only = 4; // the plane code every job shares (kDense3 or 4), else 0
  uint32_t persistent_blocks = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool prepared = false;

  ~hd_batch() { drop_graph(); }
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    prepared = false;
  }
};

namespace {

// Enqueues the batched encode, the counter reset and the persistent kernel.
void enqueue_batch(hd_batch *b, hipStream_t s) {
  if (b->enc_blocks) {
    HD_REQUIRE(b->enc_blocks < (1ull << 31), "batch too large for one encode launch");
    hipLaunchKernelGGL(k_batch_encode, dim3((uint32_t)b->enc_blocks), dim3(256), 0, s,
                       b->d_enc.get<EncodeJobDev>(), (uint32_t)b->jobs.size());
    HD_HIP(hipGetLastError());
  }
  if (b->count_blocks) {
    HD_REQUIRE(b->count_blocks < (1ull << 31), "batch too large for one count launch");
    hipLaunchKernelGGL(k_batch_totals, dim3((uint32_t)b->count_blocks), dim3(256), 0, s,
                       b->d_enc.get<EncodeJobDev>(), (uint32_t)b->jobs.size());
    HD_HIP(hipGetLastError());
  }
  if (!b->total_tiles) return;
  const int shape = b->variant % kNumShapes;
  if (b->variant >= kNumShapes) {
    const BatchJobDev *jp = b->d_jobs.get<BatchJobDev>();
    uint32_t nj = b->n_jobs_active;
    for (uint64_t t0 = 0; t0 < b->total_tiles; t0 += 0x7fffffffull) {
      const uint64_t cnt = std::min<uint64_t>(0x7fffffffull, b->total_tiles - t0);
      uint64_t tb = t0;
      void *gargs[] = {(void *)&jp, (void *)&nj, (void *)&tb};
      HD_HIP(hipLaunchKernel(grid_kernel_of(shape, b->only), dim3((uint32_t)cnt),
                             dim3(kWave * kBatchWaves), gargs, 0, s));
    }
    return;
  }
  hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, s,
                     b->d_counter.get<unsigned long long>());
  HD_HIP(hipGetLastError());
  const BatchJobDev *jp = b->d_jobs.get<BatchJobDev>();
  uint32_t nj = b->n_jobs_active;
  uint64_t tt = b->total_tiles;
  unsigned long long *ctr = b->d_counter.get<unsigned long long>();
  void *args[] = {(void *)&jp, (void *)&nj, (void *)&tt, (void *)&ctr};
  HD_HIP(hipLaunchKernel(batch_kernel_of(shape, b->only), dim3(b->persistent_blocks),
                         dim3(kWave * kBatchWaves), args, 0, s));
}

void prepare(hd_batch *b, int variant, int early_exit) {
  HD_REQUIRE(!b->jobs.empty(), "empty batch");
  HD_REQUIRE(variant >= -1 && variant < kNumBatchVariants, "unknown batch variant");
  b->drop_graph();
  b->variant = variant < 0 ? 0 : variant;
  b->early_exit = early_exit;
  hipStream_t s = lib_stream();
  const Shape sh = kShapes[b->variant % kNumShapes];
  const uint64_t TI = 64ull * sh.R, TJ = (uint64_t)kBatchWaves * sh.C;
  auto tile_work = [&](const hd_batch::Job &J) { return (uint64_t)J.g.W * J.g.code.planes; };

  // Tiles of the longest rows first (LPT order); every job keeps its own
  // output buffer.
  std::vector<size_t> order(b->jobs.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
    return tile_work(*b->jobs[x]) > tile_work(*b->jobs[y]);
  });
  b->host_jobs.clear();
  std::vector<EncodeJobDev> enc;
  uint64_t tiles = 0, blocks = 0, count_blocks = 0;
  for (size_t k : order) {
    auto &J = *b->jobs[k];
    EncodeJobDev e{};
    e.packed = J.packed.get<uint8_t>();
    e.planes = J.planes.get<uint32_t>();
    e.corr = has_corr(J.g.code.op) ? e.planes + J.g.plane_words() : nullptr;
    e.bpr = J.bpr;
    e.block_begin = blocks;
    e.count_begin = count_blocks;
    e.n = (uint32_t)J.n;
    e.n_pad = (uint32_t)J.g.n_pad;
    e.W = J.g.W;
    e.P = (uint32_t)J.g.code.planes;
    e.op = (uint32_t)J.g.code.op;
    e.runs = (J.g.W + kEncodeWords - 1) / kEncodeWords; // word blocks
    std::memcpy(e.tbl.v, J.g.code.table, 16);
    e.mode = (uint32_t)encode_mode(J.g.code);
    blocks += (uint64_t)(e.n_pad / kWave) * e.runs;
    if (has_corr(J.g.code.op)) count_blocks += e.n_pad / 256;
    enc.push_back(e);
    if (J.n < 2) continue;
    BatchJobDev d{};
    d.planes = J.planes.get<uint32_t>();
    d.corr = e.corr;
    d.out = J.out.get();
    d.tile_begin = tiles;
    d.n = (uint32_t)J.n;
    d.n_pad = (uint32_t)J.g.n_pad;
    d.W = J.g.W;
    d.P = (uint32_t)J.g.code.op;
    d.out_bytes = (uint32_t)J.out_bytes;
    d.max_dist = J.max_dist;
    d.check_every = check_interval(early_exit != 0, J.max_dist, J.g.W);
    tiles += tiles_before((J.n + TI - 1) / TI, TI / TJ);
    b->host_jobs.push_back(d);
  }
  b->only = b->host_jobs.empty() ? 4 : (int)b->host_jobs[0].P;
  for (const auto &d : b->host_jobs)
    if ((int)d.P != b->only) b->only = 0;
  if (b->only != kDense3 && b->only != 4) b->only = 0;
  b->total_tiles = tiles;
  b->enc_blocks = blocks;
  b->count_blocks = count_blocks;
  b->n_jobs_active = (uint32_t)b->host_jobs.size();
  b->d_jobs.alloc(std::max<size_t>(1, b->host_jobs.size()) * sizeof(BatchJobDev));
  b->d_enc.alloc(enc.size() * sizeof(EncodeJobDev));
  b->d_counter.alloc(sizeof(unsigned long long));
  if (!b->host_jobs.empty())
    HD_HIP(hipMemcpyAsync(b->d_jobs.get(), b->host_jobs.data(),
                          b->host_jobs.size() * sizeof(BatchJobDev), hipMemcpyHostToDevice, s));
  HD_HIP(hipMemcpyAsync(b->d_enc.get(), enc.data(), enc.size() * sizeof(EncodeJobDev),
                        hipMemcpyHostToDevice, s));
  HD_HIP(hipStreamSynchronize(s));

  // Enough workgroups to fill every CU at the kernel's occupancy.
  int dev = 0, cus = 0, per_cu = 0;
  HD_HIP(hipGetDevice(&dev));
  HD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  HD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu,
                                                      batch_kernel_of(b->variant % kNumShapes,
                                                                      b->only),
                                                      kWave * kBatchWaves, 0));
  b->persistent_blocks = (uint32_t)std::max<uint64_t>(
      1, std::min<uint64_t>((uint64_t)cus * std::max(per_cu, 1), tiles));

  // Capture: batched encode -> dense 3-op counts -> counter reset (a kernel,
  // not a memset) -> persistent all-pairs.
  Stream cap;
  HD_HIP(hipStreamBeginCapture(cap.get(), hipStreamCaptureModeThreadLocal));
  try {
    enqueue_batch(b, cap.get());
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(cap.get(), &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  HD_HIP(hipStreamEndCapture(cap.get(), &b->graph));
  HD_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  b->prepared = true;
}

} // namespace

extern "C" {

int hd_batch_create(hd_batch **out) {
  return guarded([&] {
    HD_REQUIRE(out, "null output");
    *out = new hd_batch();
  });
}

int hd_batch_destroy(hd_batch *b) {
  return guarded([&] { delete b; });
}

int hd_batch_add(hd_batch *b, const uint8_t *packed, uint64_t n, uint64_t bytes_per_seq,
                 int semantics, int out_bytes, uint32_t max_dist, uint64_t *job) {
  return guarded([&] {
    HD_REQUIRE(b && job, "null batch or job output");
    HD_REQUIRE(out_bytes == 1 || out_bytes == 2, "batch jobs store uint8 or uint16 cells");
    HD_REQUIRE(max_dist <= (out_bytes == 1 ? 255u : 65535u),
               "max_dist does not fit the output cell type");
    HD_REQUIRE(b->jobs.size() < 65536, "too many jobs in one batch");
    HD_REQUIRE(n == 0 || packed, "null buffer");
    ensure_device();
    geometry(n, bytes_per_seq, HD_SEM_DENSE, 0); // argument checks
    hipStream_t s = lib_stream();
    auto J = std::make_unique<hd_batch::Job>();
    J->n = n;
    J->bpr = bytes_per_seq;
    J->semantics = semantics;
    J->out_bytes = out_bytes;
    J->max_dist = max_dist;
    J->packed.alloc(std::max<uint64_t>(n * bytes_per_seq, 8));
    if (n)
      HD_HIP(hipMemcpyAsync(J->packed.get(), packed, n * bytes_per_seq, hipMemcpyHostToDevice, s));
    if (n) { // the nibble mask picks the plane code
      DeviceBuffer d_mask(4);
      launch_nibble_mask(J->packed.get<uint8_t>(), n * bytes_per_seq, d_mask.get<uint32_t>(), s);
      HD_HIP(hipMemcpyAsync(&J->mask, d_mask.get(), 4, hipMemcpyDeviceToHost, s));
      HD_HIP(hipStreamSynchronize(s));
    }
    J->g = geometry(n, bytes_per_seq, semantics, J->mask);
    J->planes.alloc(J->g.total_words() * 4);
    J->out.alloc(std::max<uint64_t>(tri_cells(n) * out_bytes, 1));
    HD_HIP(hipStreamSynchronize(s));
    *job = b->jobs.size();
    b->jobs.push_back(std::move(J));
    b->prepared = false;
  });
}

int hd_batch_prepare(hd_batch *b, int variant, int early_exit) {
  return guarded([&] {
    HD_REQUIRE(b, "null batch");
    prepare(b, variant, early_exit);
  });
}

int hd_batch_run(hd_batch *b, void *stream) {
  return guarded([&] {
    HD_REQUIRE(b && b->prepared, "batch not prepared");
    HD_HIP(hipGraphLaunch(b->exec, (hipStream_t)stream));
  });
}

int hd_batch_run_sequential(hd_batch *b, void *stream) {
  return guarded([&] {
    HD_REQUIRE(b, "null batch");
    hipStream_t s = (hipStream_t)stream;
    for (auto &Jp : b->jobs) {
      auto &J = *Jp;
      if (J.n == 0) continue;
      launch_encode(J.g, J.packed.get<uint8_t>(), J.planes.get<uint32_t>(), s);
      if (J.n < 2) continue;
      const int v = default_variant(J.g, b->early_exit != 0, J.max_dist);
      const uint64_t TI = 64ull * variant_of(v).R;
      run_allpairs(J.g, J.planes.get<uint32_t>(), J.out_bytes, J.max_dist, b->early_exit != 0,
                   v, 0, (J.n + TI - 1) / TI, J.out.get(), 0, s);
    }
  });
}

int hd_batch_output(hd_batch *b, uint64_t job, void *host_out) {
  return guarded([&] {
    HD_REQUIRE(b && job < b->jobs.size(), "unknown job");
    auto &J = *b->jobs[job];
    const uint64_t bytes = tri_cells(J.n) * J.out_bytes;
    if (!bytes) return;
    HD_REQUIRE(host_out, "null buffer");
    hipStream_t s = lib_stream();
    HD_HIP(hipMemcpyAsync(host_out, J.out.get(), bytes, hipMemcpyDeviceToHost, s));
    HD_HIP(hipStreamSynchronize(s));
  });
}

int hd_batch_output_dev(hd_batch *b, uint64_t job, void **d_cells, uint64_t *cells) {
  return guarded([&] {
    HD_REQUIRE(b && job < b->jobs.size() && d_cells && cells, "unknown job or null output");
    *d_cells = b->jobs[job]->out.get();
    *cells = tri_cells(b->jobs[job]->n);
  });
}

int hd_batch_info(hd_batch *b, uint64_t *jobs, uint64_t *tiles, uint32_t *blocks,
                  int *variant) {
  return guarded([&] {
    HD_REQUIRE(b, "null batch");
    if (jobs) *jobs = b->jobs.size();
    if (tiles) *tiles = b->total_tiles;
    if (blocks) *blocks = b->persistent_blocks;
    if (variant) *variant = b->variant;
  });
}

int hd_batch_variant_count(void) { return kNumBatchVariants; }

int hd_batch_variant_info(int variant, int *rows_per_lane, int *cols_per_wave, int *waves) {
  return guarded([&] {
    HD_REQUIRE(variant >= 0 && variant < kNumBatchVariants, "unknown batch variant");
    if (rows_per_lane) *rows_per_lane = kShapes[variant % kNumShapes].R;
    if (cols_per_wave) *cols_per_wave = kShapes[variant % kNumShapes].C;
    if (waves) *waves = kBatchWaves;
  });
}

} // extern "C"
