// Serving loop: a packed sequence database resident in HBM, requests
// (reference-distance queries against row subsets, small all-pairs jobs on
// database rows) admitted by device-memory footprint and launched in batches
// as they arrive, with several batches in flight on separate streams.
// C ABI: hd_server_* in include/hammingdist_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine_common.h"
#include "kernels.h"

namespace hd {

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
struct QueryDev {
  const uint2 *query;
  const uint64_t *rows;
  uint32_t *out;
  uint64_t m;
  uint64_t block_begin; // first block of this query in the launch
  int semantics;
  int dense3; // the query holds only codes of the dense 3-op code (host-side flag)
};

// 4 wavefronts x 64 rows: each wavefront loads the query into VGPRs once per
// 64 rows (with 16 rows the query loads cost 26 % more kernel time).
constexpr uint32_t kRowsPerBlock = 256;

template <int SEM, int T>
__device__ __forceinline__ void query_rows(const QueryDev &Q, uint64_t k_begin, uint64_t k_end,
                                           const uint8_t *__restrict__ db, uint64_t bpr,
                                           uint32_t lane) {
  const uint64_t units = bpr / 8;
  if constexpr (T > 0) {
    RegQuery<SEM, T> q;
    q.load(Q.query, (uint32_t)units, lane);
    for (uint64_t k = k_begin; k < k_end; k += 4) {
      const uint32_t d = q.distance(reinterpret_cast<const uint2 *>(db + Q.rows[k] * bpr),
                                    (uint32_t)units, lane);
      if (lane == 0) Q.out[k] = d;
    }
  } else {
    for (uint64_t k = k_begin; k < k_end; k += 4) {
      const uint32_t d = row_distance_mem<SEM>(
          Q.query, reinterpret_cast<const uint2 *>(db + Q.rows[k] * bpr), units, lane);
      if (lane == 0) Q.out[k] = d;
    }
  }
}

// Many queries in one launch: block g belongs to the query whose block range
// holds g and scans kRowsPerBlock of its rows; each wavefront loads the query
// once (into VGPRs when T > 0) and scans every fourth of those rows.
template <int T>
__global__ __launch_bounds__(256) void k_query_batch(const QueryDev *__restrict__ qs,
                                                     uint32_t n_q,
                                                     const uint8_t *__restrict__ db,
                                                     uint64_t bpr) {
  uint32_t lo = 0, hi = n_q;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (qs[mid].block_begin <= blockIdx.x) lo = mid;
    else hi = mid;
  }
  const QueryDev &Q = qs[__builtin_amdgcn_readfirstlane(lo)];
  const uint64_t r0 = (blockIdx.x - Q.block_begin) * kRowsPerBlock;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t k_end = std::min<uint64_t>(r0 + kRowsPerBlock, Q.m);
  if (Q.semantics == HD_SEM_DENSE)
    query_rows<HD_SEM_DENSE, T>(Q, r0 + wave, k_end, db, bpr, lane);
  else
    query_rows<HD_SEM_CHAR_X, T>(Q, r0 + wave, k_end, db, bpr, lane);
}

static const void *query_batch_kernel(uint64_t bpr) {
  switch (reg_query_t(bpr / 8)) {
  case 4: return reinterpret_cast<const void *>(&k_query_batch<4>);
  case 8: return reinterpret_cast<const void *>(&k_query_batch<8>);
  case 16: return reinterpret_cast<const void *>(&k_query_batch<16>);
  case 32: return reinterpret_cast<const void *>(&k_query_batch<32>);
  default: return reinterpret_cast<const void *>(&k_query_batch<0>);
  }
}

// ---------------------------------------------------------------------------
// shared scan: a batch's queries of one semantics whose row sets overlap read
// each database row once.  Every (query, row position) pair becomes an entry;
// entries are bucketed by database row (count, exclusive scan, scatter), and
// then one wavefront per row holds the row in VGPRs and computes it against
// every query whose set holds it, the queries streaming from L2.  The site
// functions of both semantics are symmetric in their two operands.
// ---------------------------------------------------------------------------
struct SharedEntry {
  uint32_t q, pos; // query index in the group's table, position in its rows
};

static_assert(kRowsPerBlock == 256, "one thread per row position in the entry kernels");

__device__ __forceinline__ uint32_t block_query(const QueryDev *__restrict__ qs, uint32_t n_q) {
  uint32_t lo = 0, hi = n_q;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (qs[mid].block_begin <= blockIdx.x) lo = mid;
    else hi = mid;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

__global__ __launch_bounds__(256) void k_sq_count(const QueryDev *__restrict__ qs, uint32_t n_q,
                                                  uint32_t *__restrict__ counts) {
  const QueryDev &Q = qs[block_query(qs, n_q)];
  const uint64_t k = (blockIdx.x - Q.block_begin) * kRowsPerBlock + threadIdx.x;
  if (k < Q.m) {
    atomicAdd(&counts[Q.rows[k]], 1u);
    Q.out[k] = 0u; // the shared scan adds its two half-row sums into it
  }
}

// starts[row] (exclusive prefix of the counts) advances past each entry
// placed, so afterwards it holds the end of the row's entries.
__global__ __launch_bounds__(256) void k_sq_scatter(const QueryDev *__restrict__ qs,
                                                    uint32_t n_q, uint32_t *__restrict__ starts,
                                                    SharedEntry *__restrict__ ent) {
  const uint32_t qi = block_query(qs, n_q);
  const QueryDev &Q = qs[qi];
  const uint64_t k = (blockIdx.x - Q.block_begin) * kRowsPerBlock + threadIdx.x;
  if (k < Q.m) {
    const uint32_t p = atomicAdd(&starts[Q.rows[k]], 1u);
    ent[p] = SharedEntry{qi, (uint32_t)k};
  }
}

// Exclusive scan of one value per thread across a 256-thread workgroup.
__device__ __forceinline__ uint32_t block_scan256(uint32_t v, uint32_t *total) {
  __shared__ uint32_t s_w[4];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off);
    if (lane >= (uint32_t)off) x += y;
  }
  if (lane == 63) s_w[wave] = x;
  __syncthreads();
  uint32_t base = 0;
  for (uint32_t w = 0; w < wave; ++w) base += s_w[w];
  *total = s_w[0] + s_w[1] + s_w[2] + s_w[3];
  __syncthreads();
  return base + x - v;
}

constexpr uint32_t kScanBlock = 1024; // elements per scan workgroup (4 per thread)

__global__ __launch_bounds__(256) void k_scan_local(uint32_t *__restrict__ v, uint64_t N,
                                                    uint32_t *__restrict__ sums) {
  const uint64_t i0 = (uint64_t)blockIdx.x * kScanBlock + threadIdx.x * 4;
  uint32_t a[4], s = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    a[j] = i0 + j < N ? v[i0 + j] : 0u;
    s += a[j];
  }
  uint32_t total;
  uint32_t run = block_scan256(s, &total);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (i0 + j < N) v[i0 + j] = run;
    run += a[j];
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// One workgroup: exclusive scan of the nb workgroup sums.
__global__ __launch_bounds__(256) void k_scan_sums(uint32_t *__restrict__ sums, uint32_t nb) {
  uint32_t carry = 0;
  for (uint32_t base = 0; base < nb; base += 256) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t x = i < nb ? sums[i] : 0u;
    uint32_t total;
    const uint32_t ex = block_scan256(x, &total);
    if (i < nb) sums[i] = carry + ex;
    carry += total;
  }
}

__global__ __launch_bounds__(256) void k_scan_add(uint32_t *__restrict__ v, uint64_t N,
                                                  const uint32_t *__restrict__ sums) {
  const uint64_t i0 = (uint64_t)blockIdx.x * kScanBlock + threadIdx.x * 4;
  const uint32_t add = sums[blockIdx.x];
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if (i0 + j < N) v[i0 + j] += add;
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t *__restrict__ v, uint64_t N) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (uint64_t)gridDim.x * blockDim.x)
    v[i] = 0u;
}

// Persistent teams of kSqParts wavefronts over the database rows; ends[r]
// is the end of row r's entries (its start is ends[r - 1]).  Each wavefront
// of a team holds one part of row r in VGPRs (non-temporal loads: the row is
// read once), streams that part of each entry's query against it (cached
// loads: the group's queries are read again for every row) and adds its
// count into the entry's output cell, which the count pass zeroed.  Per
// (query, row) pair, 24 queries x 300,000 rows of 14,952 B: whole rows 1.48
// ns (228 VGPRs, 2 wavefronts per SIMD), halves 1.33 ns (98 VGPRs, 4 per
// SIMD), quarters 1.44 ns; scanning each query's rows on its own takes 2.5
// ns.  Staging the row in LDS for a workgroup's four wavefronts was slower
// (the row's load and its entries serialise within the workgroup).
constexpr uint32_t kSqParts = 2; // wavefronts per database row in the shared scan

template <int SEM, int T>
__global__ __launch_bounds__(256) void k_sq_scan(const QueryDev *__restrict__ qs,
                                                 const uint8_t *__restrict__ db, uint64_t bpr,
                                                 uint64_t n_db, const uint32_t *__restrict__ ends,
                                                 const SharedEntry *__restrict__ ent) {
  const uint32_t lane = threadIdx.x & 63;
  constexpr uint32_t kTeam = 64 * kSqParts;
  const uint64_t teams = (uint64_t)gridDim.x * (blockDim.x / kTeam);
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
  const uint64_t t0 = (uint64_t)blockIdx.x * (blockDim.x / kTeam) + wave / kSqParts;
  const uint32_t units = (uint32_t)(bpr / 8);
  const uint32_t part_units = (units + kSqParts - 1) / kSqParts;
  const uint32_t part = wave % kSqParts;
  const uint32_t first = min(units, part * part_units);           // this part's first unit
  const uint32_t count = min(units, first + part_units) - first;
  for (uint64_t r = t0; r < n_db; r += teams) {
    const uint32_t e1 = ends[r], e0 = r ? ends[r - 1] : 0u;
    if (e0 == e1) continue;
    RegQuery<SEM, T> row;
    row.template load<true>(reinterpret_cast<const uint2 *>(db + r * bpr) + first, count, lane);
    for (uint32_t e = e0; e < e1; ++e) {
      const SharedEntry E = ent[e];
      const QueryDev &Q = qs[E.q];
      const uint32_t d = row.template distance<false>(Q.query + first, count, lane);
      if (lane == 0) atomicAdd(Q.out + E.pos, d);
    }
  }
}

// A shared group holds at least this many row positions (below it the fixed
// cost of bucketing outweighs the rows it saves) and at most the server's
// shared_cap (8 bytes each in the stream's scratch; kSharedMaxEntries at
// most).
constexpr uint64_t kSharedMinEntries = 100000;
// ... and reads each of its rows for this many queries on average: a shared
// scan walks every database row and is a single long launch, so groups that
// overlap little (light load: a few queries per batch) stay on the
// per-query scan, where they do not hold up the requests queued behind them.
constexpr double kSharedMinReuse = 3.0;
constexpr uint64_t kSharedMaxEntries = 16ull << 20;
constexpr uint64_t kSharedScanBlocks = 2048; // persistent grid of the shared scan
// Rates behind the choice between a query rectangle and the group's own
// scans: database bytes per second of the per-query scans (k_query_batch,
// 6.2-6.5 TB/s measured) and (row, query) pair-words per second of the
// rectangle's tile loop (1.17e13 on the all-pairs headline), plus its fixed
// cost (query encode and launches).
constexpr double kScanBytesPerSec = 6.0e12;
constexpr double kRectPairWordsPerSec = 1.1e13;
constexpr double kRectFixedSec = 30e-6;
// Bytes of one query-major distance matrix (rect_q_cap x n_db) per stream.
constexpr uint64_t kRectDistBudget = 1ull << 30;

static const void *shared_scan_kernel(int semantics, uint64_t bpr) {
#define HD_SQ(T)                                                                          \
  return semantics == HD_SEM_DENSE ? reinterpret_cast<const void *>(&k_sq_scan<HD_SEM_DENSE, T>) \
                                   : reinterpret_cast<const void *>(&k_sq_scan<HD_SEM_CHAR_X, T>);
  switch (reg_query_t((bpr / 8 + kSqParts - 1) / kSqParts)) { // row parts
  case 4: HD_SQ(4)
  case 8: HD_SQ(8)
  case 16: HD_SQ(16)
  case 32: HD_SQ(32)
  default: return nullptr;
  }
#undef HD_SQ
}

// ---------------------------------------------------------------------------
// query rectangle: a group of DENSE queries against every database row
// through the all-pairs tile loop
// ---------------------------------------------------------------------------
// A database of only 0, A, C, G, T and '-' codes is also kept as dense 3-op
// planes (the all-pairs layout: planes[g][k][p][64], then the per-row
// counts).  A group of up to kRectMaxQ such queries is encoded the same way
// and one launch computes the distance of every (database row, query) pair
// of the group: workgroup b holds database rows [64 R b, 64 R (b + 1)) in
// its lanes' VGPRs (R rows per lane), wavefront w queries [16 w, 16 w + 16)
// in SGPRs (two of each query's three words copied to VGPRs), 1 v_xor + 2
// v_bitop3 + 1 v_bcnt per pair and 32 sites (kernels.h, dense3_word).
// The distances land in a query-major matrix; a gather then reads each
// query's row positions out of it.  At 32 queries against 1,000,000 rows of
// 29,903 sites this is about 2.6 ms, where the queries' own scans of 215,000
// rows each take 17 ms.
constexpr uint32_t kRectMaxQ = 64;
constexpr uint32_t kRectCols = 16; // queries per wavefront
// Every group takes 2 database rows per lane (128-row workgroups, up to 8
// wavefronts per SIMD), the XOR reading its query word from the SGPR: over
// 200,000 rows a group of 64 queries took 1.21 ms against 1.66 ms with 4
// rows per lane (256-row workgroups, each copied query word serving 4 rows),
// groups of up to 32 queries 0.78-0.79 ms; with few queries the loop waits
// on its row loads, and the smaller rows' wavefronts keep more in flight.
// Nibble codes the dense 3-op code encodes: invalid (0), A C G T, '-'.
constexpr uint32_t kDense3Codes = (1u << 0) | (1u << 1) | (1u << 2) | (1u << 4) | (1u << 8) |
                                  (1u << 15);

template <int R>
__global__ __launch_bounds__(256) void k_rect_d3(const uint32_t *__restrict__ dbp,
                                                 uint32_t n_db, uint32_t n_pad_db, uint32_t W,
                                                 const uint32_t *__restrict__ qp, uint32_t n_q,
                                                 uint32_t n_pad_q,
                                                 uint32_t *__restrict__ dist, uint64_t ld) {
  constexpr int C = kRectCols, PS = 4;
  constexpr uint32_t OFF = 0x80000000u;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const uint32_t j0 = blockIdx.y * (4 * C) + wave * C;
  if (j0 >= n_q) return;
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t i0 = blockIdx.x * (kWave * R);
  const size_t group = (size_t)W * PS * kWave;
  const uint32_t *dcorr = dbp + (size_t)W * PS * n_pad_db; // Z totals of the rows
  const uint32_t *qcorr = qp + (size_t)W * PS * n_pad_q;   // G totals at n_pad_q
  // acc = OFF + Z_row - G_query + popcounts (the dense 3-op code's per-row
  // and per-column terms; every partial sum stays within OFF +- 2^24).
  uint32_t acc[R][C];
  {
    uint32_t gj[C];
#pragma unroll
    for (int c = 0; c < C; ++c) gj[c] = qcorr[n_pad_q + j0 + c];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const uint32_t z = dcorr[min(i0 + kWave * q + lane, n_pad_db - 1)];
#pragma unroll
      for (int c = 0; c < C; ++c) acc[q][c] = OFF + z - gj[c];
    }
  }
  using cptr = const __attribute__((address_space(4))) uint32_t *;
  cptr pb = (cptr)(qp + (size_t)(j0 / kWave) * group + (j0 % kWave));
  // Rows through one buffer resource at the block's first row group: offsets
  // below (R - 1) * group * 4 + W * 1024 < 2^31 (W < kBufWords).
  const __amdgpu_buffer_rsrc_t rows_rsrc = __builtin_amdgcn_make_buffer_rsrc(
      (void *)(dbp + (size_t)(i0 / kWave) * group), (short)0, -1, kBufferRsrcWord3);
  int row_off[R];
#pragma unroll
  for (int q = 0; q < R; ++q) row_off[q] = (int)((lane + (uint32_t)(q * group)) * 4);
  for (uint32_t k = 0; k < W; ++k) {
    // rows: planes u, v, w; columns (the queries): v, w, y
    uint32_t a[R][3];
    const int step_off = (int)(k * (uint32_t)(PS * kWave * 4));
#pragma unroll
    for (int q = 0; q < R; ++q)
#pragma unroll
      for (int p = 0; p < 3; ++p)
        a[q][p] = __builtin_amdgcn_raw_buffer_load_b32(rows_rsrc, row_off[q],
                                                       step_off + p * kWave * 4, 0);
    uint32_t bw[C][3];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int c = 0; c < C; ++c) bw[c][p] = pb[(1 + p) * kWave + c];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      // with 2 rows per lane the XOR reads its column word from the SGPR,
      // as in the all-pairs tiles
      const uint32_t b0 = R == 2 ? bw[c][0] : to_vgpr_ordered(bw[c][0]),
                     b1 = to_vgpr_ordered(bw[c][1]), b2 = to_vgpr_ordered(bw[c][2]);
#pragma unroll
      for (int q = 0; q < R; ++q)
        acc[q][c] += __builtin_popcount(dense3_word(a[q][0], a[q][1], a[q][2], b0, b1, b2));
      // blocks of 4 x 1 or 2 x 4 independent chains, as in the all-pairs tiles
      constexpr int G = R >= 4 ? 1 : 4;
      if ((c + 1) % G == 0) pin_chains<R, C, G>(acc, c + 1 - G);
    }
    pb += PS * kWave;
  }
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const uint32_t i = i0 + kWave * q + lane;
    if (i < n_db) {
#pragma unroll
      for (int c = 0; c < C; ++c)
        if (j0 + c < n_q) dist[(uint64_t)(j0 + c) * ld + i] = acc[q][c] - OFF;
    }
  }
}

// Block q copies query q of a group (bpr bytes) into row q of a packed
// buffer, the input of the group's encode.
__global__ __launch_bounds__(256) void k_rect_queries(const QueryDev *__restrict__ qs,
                                                      uint64_t bpr, uint8_t *__restrict__ packed) {
  const uint2 *src = qs[blockIdx.x].query;
  uint2 *dst = reinterpret_cast<uint2 *>(packed + (uint64_t)blockIdx.x * bpr);
  for (uint64_t u = threadIdx.x; u < bpr / 8; u += blockDim.x) dst[u] = src[u];
}

// out[k] of each query of the group = its distance to database row rows[k]
// (one thread per row position, the group's blocks as in k_query_batch).
__global__ __launch_bounds__(256) void k_rect_gather(const QueryDev *__restrict__ qs,
                                                     uint32_t n_q,
                                                     const uint32_t *__restrict__ dist,
                                                     uint64_t ld) {
  const uint32_t qi = block_query(qs, n_q);
  const QueryDev &Q = qs[qi];
  const uint64_t k = (blockIdx.x - Q.block_begin) * kRowsPerBlock + threadIdx.x;
  if (k < Q.m) Q.out[k] = dist[(uint64_t)qi * ld + Q.rows[k]];
}

// Copies database rows rows[0..n) into a contiguous packed buffer (one block
// per row, 8-byte moves).
__global__ __launch_bounds__(256) void k_gather_rows(const uint8_t *__restrict__ db, uint64_t bpr,
                                                     const uint64_t *__restrict__ rows,
                                                     uint8_t *__restrict__ out) {
  const uint64_t r = rows[blockIdx.x];
  const uint2 *src = reinterpret_cast<const uint2 *>(db + r * bpr);
  uint2 *dst = reinterpret_cast<uint2 *>(out + (uint64_t)blockIdx.x * bpr);
  for (uint64_t u = threadIdx.x; u < bpr / 8; u += blockDim.x) dst[u] = src[u];
}

// ---------------------------------------------------------------------------
// first-fit arena over one allocation (device or pinned host)
// ---------------------------------------------------------------------------
class Arena {
public:
  void init(uint8_t *base, uint64_t size) {
    base_ = base;
    size_ = size;
    free_.clear();
    sizes_.clear();
    if (size) free_[0] = size;
    used_ = 0;
  }
  static uint64_t round(uint64_t b) { return std::max<uint64_t>(256, (b + 255) / 256 * 256); }
  // Offset of a block of `bytes` or UINT64_MAX when no free run is that large.
  uint64_t alloc(uint64_t bytes) {
    bytes = round(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        const uint64_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > bytes) free_[off + bytes] = len - bytes;
        used_ += bytes;
        sizes_[off] = bytes;
        return off;
      }
    }
    return UINT64_MAX;
  }
  void release(uint64_t off) {
    auto s = sizes_.find(off);
    if (s == sizes_.end()) return;
    uint64_t len = s->second;
    sizes_.erase(s);
    used_ -= len;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + len == next->first) {
      len += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        prev->second += len;
        return;
      }
    }
    free_[off] = len;
  }
  uint8_t *at(uint64_t off) const { return base_ + off; }
  uint64_t used() const { return used_; }
  uint64_t size() const { return size_; }

private:
  uint8_t *base_ = nullptr;
  uint64_t size_ = 0, used_ = 0;
  std::map<uint64_t, uint64_t> free_;
  std::unordered_map<uint64_t, uint64_t> sizes_;
};

// Device and pinned-host blocks of one request, as offsets inside them.
//   query    device [query | rows | out]   host [query | rows | out]
//   all-pairs device [rows | packed | planes | out]   host [rows | out]
struct Layout {
  uint64_t dev = 0, host = 0;
  uint64_t d_query = 0, d_rows = 0, d_packed = 0, d_planes = 0, d_out = 0;
  uint64_t h_rows = 0, h_out = 0, in_bytes = 0, out_bytes = 0;
};

} // namespace hd

using namespace hd;

struct hd_server {
  enum Kind { QUERY = 0, ALLPAIRS = 1 };
  struct Request {
    uint64_t id = 0;
    Kind kind = QUERY;
    int state = 0; // HD_REQ_*
    int semantics = HD_SEM_DENSE, out_bytes = 4;
    uint32_t max_dist = 0;
    uint64_t m = 0; // rows of a query, sequences of an all-pairs job
    bool dense3 = false; // a DENSE query of only 0, A, C, G, T and '-' codes
    std::vector<uint64_t> rows;
    std::vector<uint8_t> query;
    Layout lay;
    Geometry g;
    uint64_t d_off = UINT64_MAX, h_off = UINT64_MAX;
    std::chrono::steady_clock::time_point t_submit, t_done;
  };
  struct Batch {
    int slot = 0;
    std::vector<uint64_t> ids;
    uint64_t table_d = UINT64_MAX, table_h = UINT64_MAX;
  };

  uint64_t n_db = 0, bpr = 0, max_batch = 64;
  uint32_t db_mask = 0xFFFFu; // nibble codes present in the database
  int n_slots = 2;
  DeviceBuffer db, dev_mem;
  PinnedBuffer host_mem;
  Arena dev_arena, host_arena;
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> events;
  std::vector<bool> slot_busy;

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<uint64_t> pending;
  std::deque<uint64_t> admitted;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
  std::vector<Batch> running;
  uint64_t next_id = 1;
  bool stopping = false;
  std::string fatal;
  std::thread worker;

  uint64_t n_done = 0, n_batches = 0, n_overlapped = 0, max_running_reqs = 0,
           batched_reqs = 0, n_shared_groups = 0, shared_queries = 0;

  // Per stream slot: the shared scan's row counts, scan sums and entries
  // (held outside the request arena; allocated when the rows fit the
  // register-resident form).
  struct Scratch {
    DeviceBuffer counts, sums, ent;
    // query rectangle: the group's packed queries, their planes and the
    // query-major distance matrix (rect_q_cap x n_db)
    DeviceBuffer q_packed, q_planes, dist;
  };
  std::vector<std::unique_ptr<Scratch>> scratch;
  uint64_t shared_cap = kSharedMaxEntries; // row positions per shared group
  uint64_t shared_min = kSharedMinEntries;
  double shared_reuse = kSharedMinReuse;
  // Query rectangle (k_rect_d3): the database's dense 3-op planes, when its
  // codes allow them; groups of at most rect_q_cap queries.
  DeviceBuffer db_planes;
  Geometry g_db, g_q;
  bool rect = false, rect_force = false;
  bool shared = false; // shared scans available (the scratch holds their buffers)
  uint32_t rect_q_cap = 0;
  uint64_t n_rect_groups = 0, rect_queries = 0;

  ~hd_server() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stopping = true;
    }
    cv_work.notify_all();
    if (worker.joinable()) worker.join();
    for (auto s : streams) (void)hipStreamSynchronize(s);
    for (auto e : events) (void)hipEventDestroy(e);
    for (auto s : streams) (void)hipStreamDestroy(s);
  }

  Layout layout(Request &r) const {
    Layout L;
    auto up = [](uint64_t b) { return Arena::round(b); };
    if (r.kind == QUERY) {
      L.d_query = 0;
      L.d_rows = up(bpr);
      L.d_out = L.d_rows + up(r.m * 8);
      L.dev = L.d_out + up(r.m * 4);
      L.h_rows = L.d_rows;
      L.h_out = L.d_out;
      L.host = L.dev;
      L.in_bytes = L.d_out; // query and rows, staged together
      L.out_bytes = r.m * 4;
    } else {
      L.d_rows = 0;
      L.d_packed = up(r.m * 8);
      L.d_planes = L.d_packed + up(r.m * bpr);
      L.d_out = L.d_planes + up(r.g.total_words() * 4);
      L.out_bytes = tri_cells(r.m) * r.out_bytes;
      L.dev = L.d_out + up(L.out_bytes);
      L.h_rows = 0;
      L.h_out = up(r.m * 8);
      L.host = L.h_out + up(L.out_bytes);
      L.in_bytes = r.m * 8;
    }
    return L;
  }

  // Group counts of one launch, added to the server's under its lock.
  struct LaunchStats {
    uint64_t rect_groups = 0, rect_queries = 0, shared_groups = 0, shared_queries = 0;
  };

  // Issues every request of batch b (requests rs, in b.ids order) on the
  // batch's stream, without the server lock: it reads only what the
  // requests and the batch already own (their arena blocks, the query
  // table's blocks b.table_d / b.table_h, the slot's scratch) and the
  // server's immutable state.
  void launch(const Batch &b, const std::vector<Request *> &rs, LaunchStats &ls) {
    const int slot = b.slot;
    hipStream_t s = streams[slot];
    std::vector<QueryDev> qd;
    for (Request *rp : rs) {
      Request &r = *rp;
      uint8_t *dbase = dev_arena.at(r.d_off), *hbase = host_arena.at(r.h_off);
      if (r.kind == QUERY) {
        HD_HIP(hipMemcpyAsync(dbase, hbase, r.lay.in_bytes, hipMemcpyHostToDevice, s));
        QueryDev d{};
        d.query = reinterpret_cast<const uint2 *>(dbase + r.lay.d_query);
        d.rows = reinterpret_cast<const uint64_t *>(dbase + r.lay.d_rows);
        d.out = reinterpret_cast<uint32_t *>(dbase + r.lay.d_out);
        d.m = r.m;
        d.semantics = r.semantics;
        d.dense3 = r.dense3 ? 1 : 0;
        qd.push_back(d);
      } else {
        HD_HIP(hipMemcpyAsync(dbase + r.lay.d_rows, hbase + r.lay.h_rows, r.lay.in_bytes,
                              hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_gather_rows, dim3((uint32_t)r.m), dim3(256), 0, s,
                           db.get<uint8_t>(), bpr,
                           reinterpret_cast<const uint64_t *>(dbase + r.lay.d_rows),
                           dbase + r.lay.d_packed);
        HD_HIP(hipGetLastError());
        launch_encode(r.g, dbase + r.lay.d_packed,
                      reinterpret_cast<uint32_t *>(dbase + r.lay.d_planes), s);
        if (r.m >= 2) {
          const int v = default_variant(r.g, true, r.max_dist);
          const uint64_t TI = 64ull * variant_of(v).R;
          run_allpairs(r.g, reinterpret_cast<uint32_t *>(dbase + r.lay.d_planes), r.out_bytes,
                       r.max_dist, true, v, 0, (r.m + TI - 1) / TI, dbase + r.lay.d_out, 0, s);
        }
      }
    }
    if (!qd.empty()) {
      // Groups of queries for the shared scan (same semantics, at least two,
      // shared_min row positions and at most shared_cap, enough reuse), then the
      // queries scanned one by one; block_begin counts within each launch.
      struct Group {
        size_t first, count;
        uint64_t blocks;
        int semantics;
      };
      std::vector<QueryDev> table, singles, rest;
      std::vector<Group> groups, rect_groups;
      auto blocks_of = [](const QueryDev &d) { return (d.m + kRowsPerBlock - 1) / kRowsPerBlock; };
      // Query rectangles first: the batch's dense-code DENSE queries in
      // arrival order, in groups of at most rect_q_cap, each group taken as
      // one rectangle over the whole database when that is estimated faster
      // than the queries' own scans.
      if (rect) {
        std::vector<QueryDev> cur;
        auto flush = [&] {
          if (cur.empty()) return;
          uint64_t sum_m = 0;
          for (const auto &d : cur) sum_m += d.m;
          const double t_scan = (double)sum_m * (double)bpr / kScanBytesPerSec;
          const double t_rect =
              std::max((double)n_db * (double)bpr / kScanBytesPerSec,
                       (double)cur.size() * (double)n_db * g_db.W / kRectPairWordsPerSec) +
              (double)sum_m * 16.0 / kScanBytesPerSec + kRectFixedSec;
          if (rect_force || t_rect < t_scan) {
            Group g{table.size(), cur.size(), 0, HD_SEM_DENSE};
            for (auto &d : cur) {
              d.block_begin = g.blocks;
              g.blocks += blocks_of(d);
              table.push_back(d);
            }
            rect_groups.push_back(g);
          } else {
            rest.insert(rest.end(), cur.begin(), cur.end());
          }
          cur.clear();
        };
        for (const QueryDev &d : qd) {
          if (d.semantics == HD_SEM_DENSE && d.dense3) {
            cur.push_back(d);
            if (cur.size() == rect_q_cap) flush();
          } else {
            rest.push_back(d);
          }
        }
        flush();
      } else {
        rest = qd;
      }
      if (shared) {
        for (int sem : {HD_SEM_DENSE, HD_SEM_CHAR_X}) {
          std::vector<QueryDev> cur;
          uint64_t cur_m = 0;
          auto flush = [&] {
            // Reads each row saves: the group's row positions over the rows
            // they cover, estimated as for independent random subsets.
            double miss = 1.0;
            for (const auto &d : cur)
              miss *= 1.0 - std::min(1.0, (double)d.m / (double)n_db);
            const double covered = std::max(1.0, (1.0 - miss) * (double)n_db);
            const bool reuse = (double)cur_m >= shared_reuse * covered;
            if (cur.size() >= 2 && cur_m >= shared_min && reuse) {
              Group g{table.size(), cur.size(), 0, sem};
              for (auto &d : cur) {
                d.block_begin = g.blocks;
                g.blocks += blocks_of(d);
                table.push_back(d);
              }
              groups.push_back(g);
            } else {
              singles.insert(singles.end(), cur.begin(), cur.end());
            }
            cur.clear();
            cur_m = 0;
          };
          for (const QueryDev &d : rest) {
            if (d.semantics != sem) continue;
            if (d.m > shared_cap) {
              singles.push_back(d);
              continue;
            }
            if (cur_m + d.m > shared_cap) flush();
            cur.push_back(d);
            cur_m += d.m;
          }
          flush();
        }
      } else {
        singles = rest;
      }
      const size_t single_first = table.size();
      uint64_t blocks = 0;
      for (QueryDev &d : singles) {
        d.block_begin = blocks;
        blocks += blocks_of(d);
        table.push_back(d);
      }
      const uint64_t bytes = table.size() * sizeof(QueryDev);
      HD_REQUIRE(b.table_d != UINT64_MAX && b.table_h != UINT64_MAX,
                 "query table not allocated");
      std::memcpy(host_arena.at(b.table_h), table.data(), bytes);
      HD_HIP(hipMemcpyAsync(dev_arena.at(b.table_d), host_arena.at(b.table_h), bytes,
                            hipMemcpyHostToDevice, s));
      const QueryDev *dtab = reinterpret_cast<const QueryDev *>(dev_arena.at(b.table_d));
      const uint8_t *db_ptr = db.get<uint8_t>();
      for (const Group &g : rect_groups) {
        Scratch &S = *scratch[slot];
        const QueryDev *gt = dtab + g.first;
        const uint32_t n_g = (uint32_t)g.count;
        HD_REQUIRE(g.blocks < (1ull << 31), "query group too large");
        hipLaunchKernelGGL(k_rect_queries, dim3(n_g), dim3(256), 0, s, gt, bpr,
                           S.q_packed.get<uint8_t>());
        // The same dense 3-op code as the database's planes (its table does
        // not depend on which of its codes are present).
        const Geometry gq = geometry(n_g, bpr, HD_SEM_DENSE, db_mask);
        launch_encode(gq, S.q_packed.get<uint8_t>(), S.q_planes.get<uint32_t>(), s);
        constexpr uint64_t rows_per_block = 128;
        const dim3 grid((uint32_t)((n_db + rows_per_block - 1) / rows_per_block),
                        (n_g + 4 * kRectCols - 1) / (4 * kRectCols));
        hipLaunchKernelGGL(k_rect_d3<2>, grid, dim3(256), 0, s,
                           db_planes.get<uint32_t>(), (uint32_t)n_db, (uint32_t)g_db.n_pad,
                           g_db.W, S.q_planes.get<uint32_t>(), n_g, (uint32_t)gq.n_pad,
                           S.dist.get<uint32_t>(), n_db);
        hipLaunchKernelGGL(k_rect_gather, dim3((uint32_t)g.blocks), dim3(256), 0, s, gt, n_g,
                           S.dist.get<uint32_t>(), n_db);
        HD_HIP(hipGetLastError());
        ++ls.rect_groups;
        ls.rect_queries += g.count;
      }
      for (const Group &g : groups) {
        Scratch &S = *scratch[slot];
        uint32_t *counts = S.counts.get<uint32_t>();
        const uint64_t nb = (n_db + kScanBlock - 1) / kScanBlock;
        HD_REQUIRE(g.blocks < (1ull << 31) && nb < (1ull << 31), "query group too large");
        const QueryDev *gt = dtab + g.first;
        const uint32_t n_g = (uint32_t)g.count;
        hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)std::min<uint64_t>(nb, 4096)), dim3(256),
                           0, s, counts, n_db);
        hipLaunchKernelGGL(k_sq_count, dim3((uint32_t)g.blocks), dim3(256), 0, s, gt, n_g,
                           counts);
        hipLaunchKernelGGL(k_scan_local, dim3((uint32_t)nb), dim3(256), 0, s, counts, n_db,
                           S.sums.get<uint32_t>());
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, s, S.sums.get<uint32_t>(),
                           (uint32_t)nb);
        hipLaunchKernelGGL(k_scan_add, dim3((uint32_t)nb), dim3(256), 0, s, counts, n_db,
                           S.sums.get<uint32_t>());
        hipLaunchKernelGGL(k_sq_scatter, dim3((uint32_t)g.blocks), dim3(256), 0, s, gt, n_g,
                           counts, S.ent.get<SharedEntry>());
        const uint32_t *ends = counts;
        const SharedEntry *ent = S.ent.get<SharedEntry>();
        uint64_t n_rows = n_db, row_bytes = bpr;
        void *args[] = {&gt, &db_ptr, &row_bytes, &n_rows, &ends, &ent};
        const uint32_t grid = (uint32_t)std::min<uint64_t>((n_db + 3) / 4, kSharedScanBlocks);
        HD_HIP(hipLaunchKernel(shared_scan_kernel(g.semantics, bpr), dim3(grid), dim3(256),
                               args, 0, s));
        ++ls.shared_groups;
        ls.shared_queries += g.count;
      }
      if (!singles.empty()) {
        HD_REQUIRE(blocks < (1ull << 31), "query batch too large");
        const QueryDev *stab = dtab + single_first;
        uint32_t n_q = (uint32_t)singles.size();
        void *args[] = {&stab, &n_q, &db_ptr, &bpr};
        HD_HIP(hipLaunchKernel(query_batch_kernel(bpr), dim3((uint32_t)blocks), dim3(256), args,
                               0, s));
      }
      HD_HIP(hipGetLastError());
    }
    for (Request *rp : rs) {
      const Request &r = *rp;
      if (r.lay.out_bytes)
        HD_HIP(hipMemcpyAsync(host_arena.at(r.h_off) + r.lay.h_out,
                              dev_arena.at(r.d_off) + r.lay.d_out, r.lay.out_bytes,
                              hipMemcpyDeviceToHost, s));
    }
    HD_HIP(hipEventRecord(events[slot], s));
  }

  void complete(Batch &b, int state) {
    const auto now = std::chrono::steady_clock::now();
    for (uint64_t id : b.ids) {
      Request &r = *reqs[id];
      r.state = state;
      r.t_done = now;
      dev_arena.release(r.d_off);
      r.d_off = UINT64_MAX;
      ++n_done;
    }
    if (b.table_d != UINT64_MAX) dev_arena.release(b.table_d);
    if (b.table_h != UINT64_MAX) host_arena.release(b.table_h);
    slot_busy[b.slot] = false;
  }

  void loop() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      // 1. completed batches free their device memory
      bool progressed = false;
      for (size_t k = 0; k < running.size();) {
        const hipError_t q = hipEventQuery(events[running[k].slot]);
        if (q == hipErrorNotReady) {
          ++k;
          continue;
        }
        complete(running[k], q == hipSuccess ? HD_REQ_DONE : HD_REQ_FAILED);
        running.erase(running.begin() + (long)k);
        progressed = true;
      }
      if (progressed) cv_done.notify_all();
      // 2. admission by device footprint, in arrival order.  A request
      // whose inputs found no host memory at submit and find none now does
      // not hold back the requests behind it whose inputs are staged: they
      // already own their host memory, and running them is what frees it.
      for (auto it = pending.begin(); it != pending.end();) {
        Request &r = *reqs[*it];
        const uint64_t d = dev_arena.alloc(r.lay.dev);
        if (d == UINT64_MAX) break;
        if (r.h_off == UINT64_MAX) { // inputs not staged at submit
          const uint64_t h = host_arena.alloc(r.lay.host);
          if (h == UINT64_MAX) {
            dev_arena.release(d);
            ++it;
            continue;
          }
          r.h_off = h;
          uint8_t *hb = host_arena.at(h);
          if (r.kind == QUERY) std::memcpy(hb, r.query.data(), bpr);
          std::memcpy(hb + r.lay.h_rows, r.rows.data(), r.m * 8);
          std::vector<uint64_t>().swap(r.rows);
          std::vector<uint8_t>().swap(r.query);
        }
        r.d_off = d;
        r.state = HD_REQ_ADMITTED;
        admitted.push_back(r.id);
        it = pending.erase(it);
      }
      // 3. everything admitted so far goes out as one batch on a free stream
      if (!admitted.empty()) {
        int slot = -1;
        for (int s = 0; s < n_slots; ++s)
          if (!slot_busy[s]) {
            slot = s;
            break;
          }
        if (slot >= 0) {
          // Taken under the lock: the batch's requests, their state and the
          // query table's arena blocks; the launch itself runs unlocked, so
          // submissions and result reads go on meanwhile.
          Batch b;
          b.slot = slot;
          std::vector<Request *> rs;
          uint64_t n_queries = 0;
          while (!admitted.empty() && b.ids.size() < max_batch) {
            const uint64_t id = admitted.front();
            admitted.pop_front();
            Request *r = reqs[id].get();
            r->state = HD_REQ_RUNNING;
            n_queries += r->kind == QUERY ? 1 : 0;
            b.ids.push_back(id);
            rs.push_back(r);
          }
          slot_busy[slot] = true;
          std::string err;
          if (n_queries) {
            const uint64_t bytes = n_queries * sizeof(QueryDev);
            b.table_d = dev_arena.alloc(bytes);
            b.table_h = host_arena.alloc(bytes);
            if (b.table_d == UINT64_MAX || b.table_h == UINT64_MAX)
              err = "no arena space for the query table";
          }
          LaunchStats ls;
          if (err.empty()) {
            lk.unlock();
            try {
              launch(b, rs, ls);
            } catch (const std::exception &e) {
              err = e.what();
            }
            lk.lock();
          }
          if (!err.empty()) {
            // The stream may hold part of the batch: wait for it, then fail
            // every request of the batch.
            fatal = err;
            (void)hipStreamSynchronize(streams[slot]);
            complete(b, HD_REQ_FAILED);
            cv_done.notify_all();
            continue;
          }
          if (!running.empty()) ++n_overlapped;
          ++n_batches;
          batched_reqs += b.ids.size();
          n_rect_groups += ls.rect_groups;
          rect_queries += ls.rect_queries;
          n_shared_groups += ls.shared_groups;
          shared_queries += ls.shared_queries;
          running.push_back(std::move(b));
          uint64_t live = 0;
          for (auto &rb : running) live += rb.ids.size();
          max_running_reqs = std::max(max_running_reqs, live);
          continue;
        }
      }
      if (stopping && running.empty()) return;
      if (!running.empty()) cv_work.wait_for(lk, std::chrono::microseconds(20));
      else cv_work.wait(lk);
    }
  }
};

extern "C" {

int hd_server_create(const uint8_t *db, uint64_t n_db, uint64_t bytes_per_seq,
                     uint64_t device_bytes, uint64_t host_bytes, int streams,
                     uint64_t max_batch, hd_server **out) {
  return guarded([&] {
    HD_REQUIRE(out, "null output");
    HD_REQUIRE(db && n_db > 0, "empty database");
    HD_REQUIRE(bytes_per_seq > 0 && bytes_per_seq % 8 == 0,
               "bytes_per_seq must be a positive multiple of 8");
    HD_REQUIRE(streams >= 1 && streams <= 8, "streams must be 1..8");
    HD_REQUIRE(device_bytes >= 4096 && host_bytes >= 4096, "budgets too small");
    ensure_device();
    auto s = std::make_unique<hd_server>();
    s->n_db = n_db;
    s->bpr = bytes_per_seq;
    s->n_slots = streams;
    s->max_batch = max_batch ? max_batch : 64;
    s->db.alloc(n_db * bytes_per_seq);
    HD_HIP(hipMemcpy(s->db.get(), db, n_db * bytes_per_seq, hipMemcpyHostToDevice));
    {
      // Codes present anywhere in the database: every all-pairs request on
      // its rows uses the plane code this mask selects.
      DeviceBuffer d_mask(4);
      launch_nibble_mask(s->db.get<uint8_t>(), n_db * bytes_per_seq, d_mask.get<uint32_t>(),
                         nullptr);
      HD_HIP(hipMemcpy(&s->db_mask, d_mask.get(), 4, hipMemcpyDeviceToHost));
    }
    s->dev_mem.alloc(device_bytes);
    s->host_mem.alloc(host_bytes);
    // Every page of the pinned arena is written once now, from several
    // threads: its first write costs a page fault, and left to the requests
    // the first pass over the arena ran at a few GB/s per submitting thread
    // (4,000 requests from 8 threads: 0.47-0.50 s of submission against 0.17
    // s once the arena had been written).
    {
      const unsigned threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
      uint8_t *h = s->host_mem.get<uint8_t>();
      std::vector<std::thread> pool;
      for (unsigned t = 0; t < threads; ++t)
        pool.emplace_back([=] {
          const uint64_t b = host_bytes * t / threads, e = host_bytes * (t + 1) / threads;
          std::memset(h + b, 0, e - b);
        });
      for (auto &th : pool) th.join();
    }
    s->dev_arena.init(s->dev_mem.get<uint8_t>(), device_bytes);
    s->host_arena.init(s->host_mem.get<uint8_t>(), host_bytes);
    // HD_SHARED_MIN_ENTRIES overrides the smallest shared group and drops the
    // reuse condition (it lets the shared scan be exercised on small
    // requests).
    if (const char *e = std::getenv("HD_SHARED_MIN_ENTRIES")) {
      s->shared_min = std::strtoull(e, nullptr, 10);
      s->shared_reuse = 0.0;
    }
    s->shared = shared_scan_kernel(HD_SEM_DENSE, bytes_per_seq) && n_db < (1ull << 32);
    // Entries for 64 row positions per database row, within [2^16, 2^24].
    if (s->shared)
      s->shared_cap = std::min<uint64_t>(kSharedMaxEntries,
                                         std::max<uint64_t>(1ull << 16, 64 * n_db));
    // Query rectangles (HD_SERVER_RECT=0 turns them off; HD_RECT_FORCE=1
    // takes every group of dense-code queries that way, whatever the cost
    // estimate says): the database's dense 3-op planes, when its codes allow
    // the code and the planes, each stream's distance matrix and a margin
    // fit in free device memory.
    {
      const char *off = std::getenv("HD_SERVER_RECT");
      const char *force = std::getenv("HD_RECT_FORCE");
      s->rect_force = force && *force == '1';
      const Geometry g = geometry(n_db, bytes_per_seq, HD_SEM_DENSE, s->db_mask);
      if (!(off && *off == '0') && (s->db_mask & ~kDense3Codes & 0xFFFFu) == 0 &&
          is_dense3(g.code) && n_db < (1ull << 31)) {
        const uint64_t cap = kRectMaxQ * n_db * 4 <= kRectDistBudget ? kRectMaxQ : kRectCols;
        const Geometry gq = geometry(kRectMaxQ, bytes_per_seq, HD_SEM_DENSE, s->db_mask);
        const uint64_t per_stream = cap * n_db * 4 + kRectMaxQ * bytes_per_seq +
                                    gq.total_words() * 4;
        const uint64_t need = g.total_words() * 4 + (uint64_t)streams * per_stream;
        size_t free_b = 0, total_b = 0;
        HD_HIP(hipMemGetInfo(&free_b, &total_b));
        if (need + (2ull << 30) <= free_b) {
          s->db_planes.alloc(g.total_words() * 4);
          launch_encode(g, s->db.get<uint8_t>(), s->db_planes.get<uint32_t>(), nullptr);
          HD_HIP(hipDeviceSynchronize());
          s->g_db = g;
          s->g_q = gq;
          s->rect_q_cap = (uint32_t)cap;
          s->rect = true;
        }
      }
    }
    if (s->shared || s->rect) {
      const uint64_t nb = (n_db + kScanBlock - 1) / kScanBlock;
      for (int k = 0; k < streams; ++k) {
        auto sc = std::make_unique<hd_server::Scratch>();
        if (s->shared) {
          sc->counts.alloc(n_db * 4);
          sc->sums.alloc(nb * 4);
          sc->ent.alloc(s->shared_cap * sizeof(SharedEntry));
        }
        if (s->rect) {
          sc->q_packed.alloc(kRectMaxQ * bytes_per_seq);
          sc->q_planes.alloc(s->g_q.total_words() * 4);
          sc->dist.alloc((uint64_t)s->rect_q_cap * n_db * 4);
        }
        s->scratch.push_back(std::move(sc));
      }
    }
    for (int k = 0; k < streams; ++k) {
      hipStream_t st;
      HD_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
      s->streams.push_back(st);
      hipEvent_t ev;
      HD_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      s->events.push_back(ev);
      s->slot_busy.push_back(false);
    }
    hd_server *raw = s.get();
    s->worker = std::thread([raw] { raw->loop(); });
    *out = s.release();
  });
}

int hd_server_destroy(hd_server *s) {
  return guarded([&] { delete s; });
}

// The caller's thread stages the request's inputs (query, row list) in the
// pinned host arena when there is room, outside the server lock, so the
// worker thread only admits by device footprint and launches; when the host
// arena is full the inputs are kept in the request and staged at admission.
static int submit(hd_server *s, std::unique_ptr<hd_server::Request> r, const uint8_t *query,
                  const uint64_t *rows, uint64_t *id) {
  return guarded([&] {
    HD_REQUIRE(s && id, "null server or id");
    if (r->kind == hd_server::ALLPAIRS) r->g = geometry(r->m, s->bpr, r->semantics, s->db_mask);
    if (r->kind == hd_server::QUERY && r->semantics == HD_SEM_DENSE && s->rect) {
      // The codes of the query: a rectangle takes it when they are all
      // codes of the dense 3-op code.
      uint32_t mask = 0;
      for (uint64_t b = 0; b < s->bpr; ++b)
        mask |= (1u << (query[b] & 15u)) | (1u << (query[b] >> 4));
      r->dense3 = (mask & ~kDense3Codes) == 0;
    }
    r->lay = s->layout(*r);
    if (r->lay.dev > s->dev_arena.size() || r->lay.host > s->host_arena.size())
      throw Error(HD_ERR_CAPACITY, "request larger than the server's memory budget");
    r->t_submit = std::chrono::steady_clock::now();
    uint8_t *staged = nullptr;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      if (!s->fatal.empty()) throw Error(HD_ERR_HIP, s->fatal);
      const uint64_t h = s->host_arena.alloc(r->lay.host);
      if (h != UINT64_MAX) {
        r->h_off = h;
        staged = s->host_arena.at(h);
      }
    }
    // One pass over the row list: range check and copy.
    uint64_t *dst = nullptr;
    if (staged) {
      if (r->kind == hd_server::QUERY) std::memcpy(staged, query, s->bpr);
      dst = reinterpret_cast<uint64_t *>(staged + r->lay.h_rows);
    } else {
      if (r->kind == hd_server::QUERY) r->query.assign(query, query + s->bpr);
      r->rows.resize(r->m);
      dst = r->rows.data();
    }
    uint64_t worst = 0;
    for (uint64_t k = 0; k < r->m; ++k) {
      const uint64_t v = rows[k];
      worst = std::max(worst, v);
      dst[k] = v;
    }
    if (worst >= s->n_db) {
      if (staged) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->host_arena.release(r->h_off);
      }
      throw Error(HD_ERR_INVALID, "row index out of range");
    }
    {
      std::lock_guard<std::mutex> lk(s->mu);
      r->id = s->next_id++;
      r->state = HD_REQ_PENDING;
      *id = r->id;
      s->pending.push_back(r->id);
      s->reqs[r->id] = std::move(r);
    }
    s->cv_work.notify_one();
  });
}

int hd_server_submit_query(hd_server *s, const uint8_t *query, const uint64_t *rows,
                           uint64_t m, int semantics, uint64_t *id) {
  auto r = std::make_unique<hd_server::Request>();
  if (!s || !query || !rows || m == 0 ||
      (semantics != HD_SEM_DENSE && semantics != HD_SEM_CHAR_X)) {
    set_last_error("query needs a server, a query, m >= 1 rows and DENSE or CHAR_X semantics");
    return HD_ERR_INVALID;
  }
  r->kind = hd_server::QUERY;
  r->semantics = semantics;
  r->m = m;
  return submit(s, std::move(r), query, rows, id);
}

int hd_server_submit_all_pairs(hd_server *s, const uint64_t *rows, uint64_t n, int semantics,
                               int out_bytes, uint32_t max_dist, uint64_t *id) {
  auto r = std::make_unique<hd_server::Request>();
  if (!s || !rows || n == 0 || (out_bytes != 1 && out_bytes != 2 && out_bytes != 4) ||
      (out_bytes < 4 && max_dist >= (1u << (8 * out_bytes)))) {
    set_last_error("all-pairs job needs n >= 1 rows, out_bytes 1/2/4 and a fitting max_dist");
    return HD_ERR_INVALID;
  }
  r->kind = hd_server::ALLPAIRS;
  r->semantics = semantics;
  r->m = n;
  r->out_bytes = out_bytes;
  r->max_dist = max_dist;
  return submit(s, std::move(r), nullptr, rows, id);
}

int hd_server_wait(hd_server *s, uint64_t id, double timeout_s, int *state) {
  return guarded([&] {
    HD_REQUIRE(s && state, "null server or state");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    HD_REQUIRE(it != s->reqs.end(), "unknown request");
    auto done = [&] {
      const int st = s->reqs[id]->state;
      return st == HD_REQ_DONE || st == HD_REQ_FAILED;
    };
    if (timeout_s < 0) s->cv_done.wait(lk, done);
    else s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done);
    *state = s->reqs[id]->state;
  });
}

int hd_server_result(hd_server *s, uint64_t id, void *out, uint64_t bytes, double *latency_s) {
  return guarded([&] {
    HD_REQUIRE(s, "null server");
    const uint8_t *src = nullptr;
    uint64_t n = 0;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      auto it = s->reqs.find(id);
      HD_REQUIRE(it != s->reqs.end(), "unknown request");
      const hd_server::Request &r = *it->second;
      HD_REQUIRE(r.state == HD_REQ_DONE, "request not done");
      HD_REQUIRE(bytes >= r.lay.out_bytes, "result buffer too small");
      HD_REQUIRE(out || r.lay.out_bytes == 0, "null buffer");
      src = s->host_arena.at(r.h_off) + r.lay.h_out;
      n = r.lay.out_bytes;
      if (latency_s) *latency_s = std::chrono::duration<double>(r.t_done - r.t_submit).count();
    }
    // A done request's result stays in its host block until the request is
    // released, which the caller does not do while it reads the result.
    if (n) std::memcpy(out, src, n);
  });
}

int hd_server_release(hd_server *s, uint64_t id) {
  return guarded([&] {
    HD_REQUIRE(s, "null server");
    {
      std::lock_guard<std::mutex> lk(s->mu);
      auto it = s->reqs.find(id);
      HD_REQUIRE(it != s->reqs.end(), "unknown request");
      const int st = it->second->state;
      HD_REQUIRE(st == HD_REQ_DONE || st == HD_REQ_FAILED, "request still in flight");
      if (it->second->h_off != UINT64_MAX) s->host_arena.release(it->second->h_off);
      s->reqs.erase(it);
    }
    s->cv_work.notify_one();
  });
}

int hd_server_stats(hd_server *s, uint64_t *done, uint64_t *batches, uint64_t *overlapped,
                    uint64_t *max_running, uint64_t *device_used) {
  return guarded([&] {
    HD_REQUIRE(s, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (done) *done = s->n_done;
    if (batches) *batches = s->n_batches;
    if (overlapped) *overlapped = s->n_overlapped;
    if (max_running) *max_running = s->max_running_reqs;
    if (device_used) *device_used = s->dev_arena.used();
  });
}

int hd_server_shared_stats(hd_server *s, uint64_t *groups, uint64_t *queries) {
  return guarded([&] {
    HD_REQUIRE(s, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (groups) *groups = s->n_shared_groups;
    if (queries) *queries = s->shared_queries;
  });
}

int hd_server_rect_stats(hd_server *s, int *enabled, uint64_t *groups, uint64_t *queries) {
  return guarded([&] {
    HD_REQUIRE(s, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (enabled) *enabled = s->rect ? 1 : 0;
    if (groups) *groups = s->n_rect_groups;
    if (queries) *queries = s->rect_queries;
  });
}

} // extern "C"
