// ref_harness.cc — drives the reference's own SIMD pair kernels
// (src/distance_sse2.cc, src/distance_avx2.cc, src/distance_avx512.cc, built
// unchanged from the reference tree by oracle/Makefile into oracle/_ref/).
//
// TEST INFRASTRUCTURE ONLY: used by tests/ to check the oracle's dense kernel
// against the reference's code and by bench.py's cpu_baseline leg.  The
// product never loads it.
//
// The rest of the reference CPU path (src/hamming_impl.cc, src/hamming.cc)
// needs the cpu_features library, which is absent from this image, so it is
// not built; the all-pairs loop below restates distances<T>'s dense CPU route
// (include/hamming/hamming_impl.hh:141-173: OpenMP schedule(static,1) over
// rows, cell i*(i-1)/2 + j).
#include <chrono>
#include <cstdint>
#include <vector>

#include <omp.h>

#include "hamming/distance_avx2.hh"
#include "hamming/distance_avx512.hh"
#include "hamming/distance_sse2.hh"

namespace {
using Row = std::vector<hamming::GeneBlock>;
using Kernel = int (*)(const Row &, const Row &, int);

Kernel pick(int which) {
  switch (which) {
  case 1:
    return hamming::distance_sse2;
  case 2:
    return hamming::distance_avx2;
  default:
    return hamming::distance_avx512;
  }
}

std::vector<Row> rows_of(const uint8_t *packed, uint64_t n, uint64_t bps) {
  std::vector<Row> rows(n);
  for (uint64_t i = 0; i < n; ++i)
    rows[i].assign(packed + i * bps, packed + (i + 1) * bps);
  return rows;
}
} // namespace

extern "C" {

// One pair through the reference kernel `which` (1 sse2, 2 avx2, 3 avx512).
int ref_distance(int which, const uint8_t *a, const uint8_t *b, uint64_t bytes,
                 int max_dist) {
  Row ra(a, a + bytes), rb(b, b + bytes);
  return pick(which)(ra, rb, max_dist);
}

// Rows [row0, row1) of the dense all-pairs triangle with the reference kernel
// `which`; cells are written relative to row0's first cell.  The rows are
// copied into the reference's vector-of-vectors form first (as to_dense_data
// does before the timed distance loop); *seconds receives the loop's wall
// time.  Returns the OpenMP thread count used.
int ref_all_pairs(int which, const uint8_t *packed, uint64_t n, uint64_t bps,
                  uint64_t row0, uint64_t row1, int out_bytes, int max_dist,
                  void *out, int threads, double *seconds) {
  if (threads > 0) omp_set_num_threads(threads);
  const int used = omp_get_max_threads();
  const std::vector<Row> rows = rows_of(packed, row1 < n ? row1 : n, bps);
  const auto t0 = std::chrono::steady_clock::now();
  const Kernel f = pick(which);
  const uint64_t base = row0 ? row0 * (row0 - 1) / 2 : 0;
#pragma omp parallel for schedule(static, 1)
  for (int64_t i = (int64_t)row0; i < (int64_t)row1; ++i) {
    const uint64_t off = (uint64_t)i * (uint64_t)(i - 1) / 2 - base;
    for (int64_t j = 0; j < i; ++j) {
      const int d = f(rows[i], rows[j], max_dist);
      if (out_bytes == 1) static_cast<uint8_t *>(out)[off + j] = (uint8_t)d;
      else if (out_bytes == 2) static_cast<uint16_t *>(out)[off + j] = (uint16_t)d;
      else static_cast<uint32_t *>(out)[off + j] = (uint32_t)d;
    }
  }
  if (seconds)
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return used;
}
}
