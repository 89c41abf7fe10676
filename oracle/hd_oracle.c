/*
 * hd_oracle.c — CPU restatement of the reference (hammingdist) algorithms on
 * the all-pairs Hamming distance path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (hammingdist_amd) never links, loads or calls it.
 *
 * Every function names the reference file:line it restates.  How each is
 * pinned (golden vectors, known-answer tests, closed forms, brute force) is
 * listed in tests/test_oracle.py; no function here is left unpinned.
 *
 * Conventions: sequences are char arrays of length L, n of them back to back
 * (data + i*L).  Packed rows use the reference dense format (2 sites / byte,
 * low nibble first, padded with 0xFF to bytes_per_seq, a multiple of 8).
 * Triangle cell (i, j), i > j, lies at i*(i-1)/2 + j.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#if defined(__x86_64__)
#include <immintrin.h>
#endif

#define GAP 0xFFu

/* A uint32 cell bound as the int max_dist of the reference kernels. */
static int md_int(uint32_t m) { return m > 0x7FFFFFFFu ? 0x7FFFFFFF : (int)m; }

/* lookupTable, src/hamming_impl.cc:31-44 */
void oracle_lookup(int include_x, uint8_t out[256]) {
  memset(out, 0, 256);
  out['-'] = 0xFF;
  out['A'] = 0x11;
  out['C'] = 0x22;
  out['G'] = 0x44;
  out['T'] = 0x88;
  if (include_x) out['X'] = 0x33;
}

/* bytes of one packed row: from_string, src/hamming_impl.cc:247-264 */
uint64_t oracle_bytes_per_seq(uint64_t L) {
  uint64_t blocks = (L + 1) / 2;
  return (blocks + 7) / 8 * 8;
}

/* from_string, src/hamming_impl.cc:244-267 (table: lookupTable(include_x)) */
void oracle_pack(const char *s, uint64_t L, int include_x, uint8_t *out,
                 uint64_t bytes_per_seq) {
  uint8_t lut[256];
  oracle_lookup(include_x, lut);
  uint64_t full = L / 2, k = 0;
  for (uint64_t b = 0; b < full; ++b) {
    out[k++] = (uint8_t)((lut[(uint8_t)s[2 * b]] & 0x0F) |
                         (lut[(uint8_t)s[2 * b + 1]] & 0xF0));
  }
  if (L & 1) out[k++] = (uint8_t)((lut[(uint8_t)s[L - 1]] & 0x0F) | 0xF0);
  while (k < bytes_per_seq) out[k++] = 0xFF;
}

/* distance_cpp, src/hamming_impl.cc:131-140 */
int oracle_distance_dense(const uint8_t *a, const uint8_t *b, uint64_t bytes,
                          int max_dist) {
  int r = 0;
  for (uint64_t i = 0; i < bytes; ++i) {
    uint8_t c = (uint8_t)(a[i] & b[i]);
    r += ((c & 0x0F) == 0) + ((c & 0xF0) == 0);
  }
  return r < max_dist ? r : max_dist;
}

/* distance(), src/hamming.cc:54-74 (the caller checks equal lengths) */
uint32_t oracle_distance_char(const char *s0, const char *s1, uint64_t L,
                              int include_x) {
  uint8_t lut[256];
  oracle_lookup(include_x, lut);
  uint32_t d = 0;
  for (uint64_t i = 0; i < L; ++i) {
    uint8_t a = lut[(uint8_t)s0[i]], b = lut[(uint8_t)s1[i]];
    int invalid = (a & b) == 0;
    int differ = s0[i] != s1[i];
    if (invalid || differ) {
      int nodash = a != 0xFF && b != 0xFF;
      d += (uint32_t)(invalid || (differ && nodash));
    }
  }
  return d;
}

/* fasta_reference_distances inner loop, src/hamming.cc:123-137: the
 * sequence seq (length Ls) against reference ref (length >= Ls).            */
uint32_t oracle_reference_distance(const char *ref, const char *seq, uint64_t Ls,
                                   int include_x) {
  return oracle_distance_char(ref, seq, Ls, include_x);
}

/* get_reference_expression, src/hamming_impl.cc:142-168 */
void oracle_consensus(const char *data, uint64_t n, uint64_t L, int include_x,
                      char *out) {
  static const char itoc[6] = {'A', 'A', 'C', 'G', 'T', 'X'};
  int ctoi[256];
  memset(ctoi, 0, sizeof ctoi);
  ctoi['A'] = 1;
  ctoi['C'] = 2;
  ctoi['G'] = 3;
  ctoi['T'] = 4;
  if (include_x) ctoi['X'] = 5;
  uint64_t *counts = (uint64_t *)calloc(L * 6, sizeof(uint64_t));
  for (uint64_t s = 0; s < n; ++s)
    for (uint64_t i = 0; i < L; ++i) ++counts[i * 6 + ctoi[(uint8_t)data[s * L + i]]];
  for (uint64_t i = 0; i < L; ++i) {
    /* std::max_element over [1, 6): first maximum */
    int best = 1;
    for (int c = 2; c < 6; ++c)
      if (counts[i * 6 + c] > counts[i * 6 + best]) best = c;
    out[i] = itoc[best];
  }
  free(counts);
}

/* to_sparse_data, src/hamming_impl.cc:170-188: the (index, code) pairs of the
 * sites where seq differs from the consensus; returns the pair count.        */
static uint64_t sparse_list(const char *seq, const char *cons, uint64_t L,
                            const uint8_t *lut, uint64_t *idx, uint8_t *code) {
  uint64_t m = 0;
  for (uint64_t i = 0; i < L; ++i)
    if (cons[i] != seq[i]) {
      idx[m] = i;
      code[m] = lut[(uint8_t)seq[i]];
      ++m;
    }
  return m;
}

/* distance_sparse, src/hamming_impl.cc:93-129 */
static int sparse_merge(const uint64_t *ia_, const uint8_t *ca, uint64_t na,
                        const uint64_t *ib_, const uint8_t *cb, uint64_t nb,
                        int max_dist) {
  int r = 0;
  uint64_t ia = 0, ib = 0;
  while (ia < na && ib < nb) {
    if (ia_[ia] < ib_[ib]) {
      r += ca[ia] != GAP;
      ++ia;
    } else if (ia_[ia] > ib_[ib]) {
      r += cb[ib] != GAP;
      ++ib;
    } else {
      int nodash = ca[ia] != GAP && cb[ib] != GAP;
      int differ = ca[ia] != cb[ib];
      r += nodash && differ;
      ++ia;
      ++ib;
    }
    if (r >= max_dist) return max_dist;
  }
  while (ia < na) r += ca[ia++] != GAP;
  while (ib < nb) r += cb[ib++] != GAP;
  return r < max_dist ? r : max_dist;
}

/* One cell of the sparse route: to_sparse_data of both sequences against the
 * dataset's consensus `cons` (src/hamming_impl.cc:170-188), then
 * distance_sparse (src/hamming_impl.cc:93-129).  For sampled cells of
 * datasets too large for oracle_all_pairs_sparse; cons comes from
 * oracle_consensus over the whole dataset.                                   */
int oracle_distance_sparse_pair(const char *a, const char *b, const char *cons, uint64_t L,
                                int include_x, uint32_t max_dist) {
  uint8_t lut[256];
  oracle_lookup(include_x, lut);
  uint64_t *ia = (uint64_t *)malloc((L ? L : 1) * sizeof(uint64_t));
  uint64_t *ib = (uint64_t *)malloc((L ? L : 1) * sizeof(uint64_t));
  uint8_t *ca = (uint8_t *)malloc(L ? L : 1);
  uint8_t *cb = (uint8_t *)malloc(L ? L : 1);
  uint64_t na = sparse_list(a, cons, L, lut, ia, ca);
  uint64_t nb = sparse_list(b, cons, L, lut, ib, cb);
  int d = sparse_merge(ia, ca, na, ib, cb, nb, md_int(max_dist));
  free(ia);
  free(ib);
  free(ca);
  free(cb);
  return d;
}

/* safe_int_cast<T>(max_distance), include/hamming/hamming_impl.hh:37-45 */
uint32_t oracle_clamp_max_distance(int64_t max_distance, int out_bytes) {
  uint64_t tmax = out_bytes == 1 ? 0xFFu : out_bytes == 2 ? 0xFFFFu : 0xFFFFFFFFu;
  if (max_distance > (int64_t)tmax) return (uint32_t)tmax;
  return (uint32_t)((uint64_t)max_distance & tmax);
}

static void store_cell(void *out, int out_bytes, uint64_t x, uint32_t v) {
  if (out_bytes == 1) ((uint8_t *)out)[x] = (uint8_t)v;
  else if (out_bytes == 2) ((uint16_t *)out)[x] = (uint16_t)v;
  else ((uint32_t *)out)[x] = v;
}

/* The sparse/dense heuristic, include/hamming/hamming_impl.hh:95, 104-118:
 * returns 1 when the reference CPU route uses distance_sparse.              */
int oracle_use_sparse(const char *data, uint64_t n, uint64_t L, int include_x) {
  if (include_x) return 1;
  char *cons = (char *)malloc(L);
  oracle_consensus(data, n, L, 0, cons);
  uint64_t n_diff = 0;
  for (uint64_t s = 0; s < n; ++s)
    for (uint64_t i = 0; i < L; ++i) n_diff += cons[i] != data[s * L + i];
  free(cons);
  double frac = (double)n_diff / (double)(n * L);
  return frac < 0.005;
}

/* All-pairs of the sparse route: distances<T>, hamming_impl.hh:95, 119-137 */
void oracle_all_pairs_sparse(const char *data, uint64_t n, uint64_t L, int include_x,
                             int out_bytes, uint32_t max_dist, void *out) {
  uint8_t lut[256];
  oracle_lookup(include_x, lut);
  char *cons = (char *)malloc(L ? L : 1);
  oracle_consensus(data, n, L, include_x, cons);
  uint64_t *len = (uint64_t *)malloc(n * sizeof(uint64_t));
  uint64_t **idx = (uint64_t **)malloc(n * sizeof(uint64_t *));
  uint8_t **code = (uint8_t **)malloc(n * sizeof(uint8_t *));
  for (uint64_t s = 0; s < n; ++s) {
    idx[s] = (uint64_t *)malloc((L ? L : 1) * sizeof(uint64_t));
    code[s] = (uint8_t *)malloc(L ? L : 1);
    len[s] = sparse_list(data + s * L, cons, L, lut, idx[s], code[s]);
  }
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    uint64_t off = (uint64_t)i * (uint64_t)(i - 1) / 2;
    for (int64_t j = 0; j < i; ++j)
      store_cell(out, out_bytes, off + (uint64_t)j,
                 (uint32_t)sparse_merge(idx[i], code[i], len[i], idx[j], code[j],
                                        len[j], md_int(max_dist)));
  }
  for (uint64_t s = 0; s < n; ++s) {
    free(idx[s]);
    free(code[s]);
  }
  free(idx);
  free(code);
  free(len);
  free(cons);
}

/* All-pairs of the dense route (CPU SIMD path and the CUDA kernel alike):
 * distances<T>, hamming_impl.hh:142-173, pair kernel distance_cpp.          */
void oracle_all_pairs_dense(const uint8_t *packed, uint64_t n, uint64_t bytes_per_seq,
                            int out_bytes, uint32_t max_dist, void *out) {
#pragma omp parallel for schedule(static, 1)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    uint64_t off = (uint64_t)i * (uint64_t)(i - 1) / 2;
    for (int64_t j = 0; j < i; ++j)
      store_cell(out, out_bytes, off + (uint64_t)j,
                 (uint32_t)oracle_distance_dense(packed + (uint64_t)i * bytes_per_seq,
                                                 packed + (uint64_t)j * bytes_per_seq,
                                                 bytes_per_seq, md_int(max_dist)));
  }
}

/* distances<T>(data, include_x, ..., use_gpu, max_distance),
 * include/hamming/hamming_impl.hh:68-176, for equal-length sequences.
 * gpu_route != 0 restates the reference GPU route (always dense); otherwise
 * the CPU route with its sparse/dense heuristic.  Returns the route taken:
 * 0 dense, 1 sparse.                                                         */
int oracle_distances(const char *data, uint64_t n, uint64_t L, int include_x,
                     int gpu_route, int out_bytes, int64_t max_distance, void *out) {
  uint32_t md = oracle_clamp_max_distance(max_distance, out_bytes);
  int sparse = gpu_route ? 0 : oracle_use_sparse(data, n, L, include_x);
  if (sparse) {
    oracle_all_pairs_sparse(data, n, L, include_x, out_bytes, md, out);
    return 1;
  }
  uint64_t bps = oracle_bytes_per_seq(L);
  uint8_t *packed = (uint8_t *)malloc(n * bps + 1);
  for (uint64_t s = 0; s < n; ++s) oracle_pack(data + s * L, L, 0, packed + s * bps, bps);
  oracle_all_pairs_dense(packed, n, bps, out_bytes, md, out);
  free(packed);
  return 0;
}

/* ---- CPU baseline: the reference's SIMD kernel, restated ----------------- */
/* distance_avx512, src/distance_avx512.cc:6-76: u8 lane counters, n_inner
 * early-exit cadence, scalar tail.  Same result as distance_cpp.            */
#if defined(__x86_64__)
__attribute__((target("avx512f,avx512bw"))) static int
dist_avx512(const uint8_t *a, const uint8_t *b, uint64_t bytes, int max_dist) {
  const uint64_t blk = 64;
  int r = 0;
  const __m512i lsb = _mm512_set1_epi8(1);
  const __m512i m0 = _mm512_set1_epi8(0x0F);
  const __m512i m1 = _mm512_set1_epi8((char)0xF0);
  uint64_t n_iter = bytes / blk;
  uint64_t n_inner = max_dist >= 255 ? 127 : 16;
  uint64_t n_outer = 1 + n_iter / n_inner;
  for (uint64_t j = 0; j < n_outer; ++j) {
    uint64_t end = (j + 1) * n_inner < n_iter ? (j + 1) * n_inner : n_iter;
    __m512i acc = _mm512_setzero_si512();
    for (uint64_t i = j * n_inner; i < end; ++i) {
      __m512i va = _mm512_loadu_si512((const void *)(a + blk * i));
      __m512i vb = _mm512_loadu_si512((const void *)(b + blk * i));
      __m512i c = _mm512_and_si512(va, vb);
      __mmask64 z0 = _mm512_cmpeq_epi8_mask(_mm512_and_si512(c, m0), _mm512_setzero_si512());
      acc = _mm512_mask_add_epi8(acc, z0, lsb, acc);
      __mmask64 z1 = _mm512_cmpeq_epi8_mask(_mm512_and_si512(c, m1), _mm512_setzero_si512());
      acc = _mm512_mask_add_epi8(acc, z1, lsb, acc);
    }
    __m512i sums = _mm512_sad_epu8(acc, _mm512_setzero_si512());
    uint64_t part[8];
    _mm512_storeu_si512((void *)part, sums);
    for (int k = 0; k < 8; ++k) r += (int)part[k];
    if (r >= max_dist) return max_dist;
  }
  for (uint64_t i = blk * n_iter; i < bytes; ++i) {
    uint8_t c = (uint8_t)(a[i] & b[i]);
    r += ((c & 0x0F) == 0) + ((c & 0xF0) == 0);
  }
  return r < max_dist ? r : max_dist;
}
#endif

/* 1 when this CPU runs the AVX-512BW restatement.                            */
int oracle_have_avx512(void) {
#if defined(__x86_64__)
  __builtin_cpu_init();
  return __builtin_cpu_supports("avx512bw") != 0;
#else
  return 0;
#endif
}

/* The dense CPU route of distances<T> as the reference runs it:
 * OpenMP schedule(static,1) over rows (hamming_impl.hh:161-173) with the
 * fastest SIMD kernel (hamming_impl.cc:46-78).  Rows [row0, row1) only, so
 * that a bounded sample can be timed.  Returns the thread count used.       */
int oracle_all_pairs_dense_simd(const uint8_t *packed, uint64_t n, uint64_t bytes_per_seq,
                                uint64_t row0, uint64_t row1, int out_bytes,
                                uint32_t max_dist, void *out, int threads) {
  int used = 1;
  int simd = oracle_have_avx512();
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
  used = omp_get_max_threads();
#endif
  (void)n;
  uint64_t base = row0 * (row0 ? row0 - 1 : 0) / 2;
#pragma omp parallel for schedule(static, 1)
  for (int64_t i = (int64_t)row0; i < (int64_t)row1; ++i) {
    uint64_t off = (uint64_t)i * (uint64_t)(i - 1) / 2 - base;
    const uint8_t *a = packed + (uint64_t)i * bytes_per_seq;
    for (int64_t j = 0; j < i; ++j) {
      const uint8_t *b = packed + (uint64_t)j * bytes_per_seq;
      int d;
#if defined(__x86_64__)
      if (simd) d = dist_avx512(a, b, bytes_per_seq, md_int(max_dist));
      else d = oracle_distance_dense(a, b, bytes_per_seq, md_int(max_dist));
#else
      d = oracle_distance_dense(a, b, bytes_per_seq, md_int(max_dist));
#endif
      store_cell(out, out_bytes, off + (uint64_t)j, (uint32_t)d);
    }
  }
  return used;
}

/* ---- writers ------------------------------------------------------------- */
/* lower_triangular_lines / partial_write_lower_triangular for a whole
 * triangle, include/hamming/hamming_utils.hh:19-110: row i (1..n-1) holds
 * cells (i,0)..(i,i-1) as decimals joined by ',' and ends in '\n'.  n < 2
 * writes nothing (the reference returns before opening the file).  Returns
 * bytes written or -1.                                                       */
int64_t oracle_write_lower_triangular(const char *path, const void *cells, uint64_t n,
                                      int out_bytes) {
  if (n < 2) return 0;
  FILE *f = fopen(path, "wb");
  if (!f) return -1;
  int64_t total = 0;
  uint64_t x = 0;
  for (uint64_t i = 1; i < n; ++i) {
    for (uint64_t j = 0; j < i; ++j, ++x) {
      uint32_t v = out_bytes == 1 ? ((const uint8_t *)cells)[x]
                   : out_bytes == 2 ? ((const uint16_t *)cells)[x]
                                    : ((const uint32_t *)cells)[x];
      total += fprintf(f, "%u%c", v, j + 1 == i ? '\n' : ',');
    }
  }
  fclose(f);
  return total;
}

/* DataSet::dump_sparse serial order, include/hamming/hamming.hh:104-112:
 * "i j d\n" for every cell with d <= threshold, row-major.                  */
int64_t oracle_write_sparse(const char *path, const void *cells, uint64_t n,
                            int out_bytes, int64_t threshold) {
  FILE *f = fopen(path, "wb");
  if (!f) return -1;
  int64_t total = 0;
  uint64_t x = 0;
  for (uint64_t i = 1; i < n; ++i)
    for (uint64_t j = 0; j < i; ++j, ++x) {
      uint32_t v = out_bytes == 1 ? ((const uint8_t *)cells)[x]
                   : out_bytes == 2 ? ((const uint16_t *)cells)[x]
                                    : ((const uint32_t *)cells)[x];
      if ((int64_t)v <= threshold) total += fprintf(f, "%llu %llu %u\n",
                                                    (unsigned long long)i,
                                                    (unsigned long long)j, v);
    }
  fclose(f);
  return total;
}

/* DataSet::dump, include/hamming/hamming.hh:63-73: full symmetric matrix,
 * ", " separated, diagonal 0.                                                */
int64_t oracle_write_full(const char *path, const void *cells, uint64_t n, int out_bytes) {
  FILE *f = fopen(path, "wb");
  if (!f) return -1;
  int64_t total = 0;
  for (uint64_t i = 0; i < n; ++i) {
    for (uint64_t j = 0; j < n; ++j) {
      uint32_t v = 0;
      if (i != j) {
        uint64_t a = i > j ? i : j, b = i > j ? j : i;
        uint64_t x = a * (a - 1) / 2 + b;
        v = out_bytes == 1 ? ((const uint8_t *)cells)[x]
            : out_bytes == 2 ? ((const uint16_t *)cells)[x]
                             : ((const uint32_t *)cells)[x];
      }
      total += fprintf(f, "%u", v);
      if (j != n - 1) total += fprintf(f, ", ");
    }
    total += fprintf(f, "\n");
  }
  fclose(f);
  return total;
}
