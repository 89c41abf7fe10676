// Device kernels of the Hamming engine (gfx950, wave64, integer VALU).
//
// Data layout in HBM
//   packed  : the reference's dense nibble rows, row-major, bytes_per_seq each
//             (src/hamming_impl.cc:244-267).  Used by the reference-distance
//             scan and as the upload format.
//   planes  : bit-sliced, in groups of 64 rows:
//             planes[g][k][p][64] uint32, row i = 64 g + lane.  Word k of plane
//             p of row i holds bit p of the plane code of sites 32k .. 32k+31.
//             A wavefront reading word (k, p) of one group reads 256
//             contiguous bytes; the same word of 16 consecutive columns is one
//             64-byte scalar load; all P planes of one step of a group lie in
//             P*256 contiguous bytes, so every per-step offset is an
//             instruction immediate.
//   triangle: cell (i, j), i > j, at i*(i-1)/2 + j (64-bit index).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>
#include <utility>

#include "hammingdist_amd.h"

namespace hd {

constexpr int kWave = 64;

struct Table16 {
  uint8_t v[16];
};

// ---------------------------------------------------------------------------
// all-pairs tile kernel
//
// A workgroup of NW waves owns a TI x TJ tile of the lower triangle
// (TI = 64 R rows, TJ = NW C columns).  Lane l of every wave holds rows
// i0 + l + 64 q (q < R) in VGPRs; wave w holds columns j0 + w C + c (c < C)
// in SGPRs (wave-uniform scalar loads), copied into VGPRs one column at a
// time right before the column's R rows use them (the first op of every code
// reads its word from the SGPR when R <= 2), so each VALU op
// pairs a per-lane row word with a broadcast column word and no LDS traffic
// is needed.
//
// Plane codes (template parameter P):
//   P = 4, 5, 6  P one-hot-style planes; a site counts when no plane is set
//                on both sides.  Per pair and 32 sites: P-1 AND/AND-OR ops,
//                one NOR-with-AND and one popcount-accumulate (P + 1 ops).
//   P = kDense3  the dense nibble semantics, (a & b) == 0, for data whose
//                nibbles are only 0, A, C, G, T and '-' (the reference's
//                lookup without X).  Four planes per row, (u, v, w, y):
//                  u = {A, C, invalid}, v = {C, T, invalid},
//                  w = {'-', invalid}, y = {G, T};
//                a row reads u, v, w, a column v, w, y.  Per pair and 32
//                sites one XOR, two v_bitop3 (one column plane each) and
//                one popcount-accumulate, 4 ops:
//                  s1 = v_i ^ v_j
//                  s2 = ~s1 & (u_i ^ y_j)                    (table 0x14)
//                  s3 = w_i ? (w_j & ~s2) : (w_j | ~s2)      (table 0x4d)
//                Per site, s3 + [site of row i invalid] - [site of row j
//                '-'] is the site's distance term, so popcount(s3) + Z_i - G_j
//                is the distance, where Z_i counts the invalid sites of row i
//                and G_j the '-' sites of row j (padding included).  With the
//                early exit on, the counts of each kCheckWords-word chunk are
//                added after that chunk, so every vote sees the exact partial
//                distance; both are written by the encode tiles (encode_tile).
//   P = kGap2, kGap3  the sparse semantics: a site counts when the classes
//                differ and neither is '-'.  Every class present and '-'
//                get distinct 2- or 3-bit codes (planes 0 .. B-1), plane B
//                marks the '-' sites.  Per pair and 32 sites: one XOR, one
//                (x ^ y) | s per further code bit and one popcount-accumulate
//                (3 or 4 ops), f = codes differ.  Per site f - [row i '-'] -
//                [row j '-'] + 2 [both '-'] is the site's distance term, so
//                popcount(f) - G_i - G_j + 2 |g_i & g_j| is the distance,
//                where G counts the '-' sites of a row.  Sites past the
//                packed row are '-' in every row and add nothing to any
//                pair: they keep the '-' code but stay out of plane B and
//                of G.  The last term is summed over the words where a
//                column has a '-' only: a per-row bitmap of such words and
//                a summary of its chunks (gap_corr_words) drive a
//                wave-uniform loop after the main one.  With the
//                early exit on, chunk counts of G make the votes use
//                popcount - G_i - G_j, a lower bound of the partial distance.
//   P = kClass3  the sparse semantics for rows of kBufWords words or more: 3
//                code bits and a '-' plane, s & ~g_i & ~g_j per pair (5 ops).
// ---------------------------------------------------------------------------
constexpr int kDense3 = 3;
constexpr int kGap2 = 7;
constexpr int kClass3 = 8;
constexpr int kGap3 = 9;

// The offset codes (kDense3, kGap2, kGap3) serve rows of fewer than
// kBufWords 32-site words (8,388,608 sites): their tile loop addresses the
// rows of a tile (at most 8 row groups of 4 planes) with 31-bit buffer
// offsets.  Longer rows take the nibble and class codes.
constexpr uint32_t kBufWords = 1u << 18;
// Word 3 of a raw buffer resource on gfx9-family GPUs (no format conversion,
// 32-bit data, stride 0).
constexpr int kBufferRsrcWord3 = 0x00020000;

// Words between two early-exit votes (1024 sites).
constexpr uint32_t kCheckWords = 32;

// kDense3 correction counts after the planes, in words from the first:
// Z[n_pad], G[n_pad] (whole rows), then Zc[T][n_pad], Gc[T][n_pad] per chunk
// of kCheckWords words, T = ceil(W / kCheckWords).
__host__ __device__ constexpr uint64_t dense3_corr_words(uint32_t W, uint64_t n_pad) {
  return (2 + 2 * (uint64_t)((W + kCheckWords - 1) / kCheckWords)) * n_pad;
}

// kGap2 / kGap3 correction words after the planes (same size as kDense3's):
// G[n_pad] (whole rows), S[n_pad], then Gc[T][n_pad] per chunk of
// kCheckWords words and BM[T][n_pad], bit k of BM[t][row] set when word
// 32 t + k of the row holds a '-'.  Bit b of S[row] is set when one of the
// chunks [b s, (b + 1) s) has a bit of BM set, s = gap_summary_span(T).
__host__ __device__ constexpr uint64_t gap_corr_words(uint32_t W, uint64_t n_pad) {
  return dense3_corr_words(W, n_pad);
}

// Chunks of kCheckWords words per bit of the kGap codes' summary word S.
__host__ __device__ constexpr uint32_t gap_summary_span(uint32_t T) {
  return (T + 31) / 32;
}

// Planes stored per row and word for plane code P.
__host__ __device__ constexpr int stored_planes(int P) {
  return P == kDense3 ? 4 : P == kGap2 ? 3 : P == kClass3 ? 4 : P == kGap3 ? 4 : P;
}

// True for the codes whose rows carry correction counts after the planes.
__host__ __device__ constexpr bool has_corr(int P) {
  return P == kDense3 || P == kGap2 || P == kGap3;
}

// A wave-uniform word copied into a VGPR (one v_mov_b32), kept in program
// order with the other volatile asm of the loop (pin_chains): the copies of
// a column block are issued right before the block's ops instead of all at
// the start of the step.
__device__ __forceinline__ uint32_t to_vgpr_ordered(uint32_t x) {
  uint32_t v;
  __asm__ volatile("v_mov_b32 %0, %1" : "=v"(v) : "s"(x));
  return v;
}

// Marks the accumulators of columns [c0, c0 + G) as written here (no
// instruction): their popcount-accumulates complete before any later
// ordered copy, so the scheduler interleaves the R x G independent chains of
// one column block instead of running long dependent runs, and keeps only
// that block's column words in VGPRs.
template <int R, int C, int G>
__device__ __forceinline__ void pin_chains(uint32_t (&acc)[R][C], int c0) {
#pragma unroll
  for (int c = c0; c < c0 + G; ++c)
#pragma unroll
    for (int q = 0; q < R; ++q) __asm__ volatile("" : "+v"(acc[q][c]));
}

// Tile t -> (row block b, column block cb) of a triangle tiled with M column
// blocks per row block: row block b holds tiles [M b(b+1)/2, M (b+1)(b+2)/2).
__device__ __forceinline__ void tile_coords(uint64_t t, uint32_t M, uint32_t &b,
                                            uint32_t &cb) {
  uint64_t r = (uint64_t)((sqrt(8.0 * (double)t / (double)M + 1.0) - 1.0) * 0.5);
  while ((uint64_t)M * (r + 1) * (r + 2) / 2 <= t) ++r;
  while (r > 0 && (uint64_t)M * r * (r + 1) / 2 > t) --r;
  b = (uint32_t)r;
  cb = (uint32_t)(t - (uint64_t)M * r * (r + 1) / 2);
}

// One 32-site word of the dense 3-op code: row planes (u, v, w) against
// column planes (v, w, y); the popcount of the result plus the row's invalid
// count minus the column's '-' count is the word's distance (summed over a
// row).  The first op is a plain XOR (a 4-byte VOP2 instruction: the 8-byte
// v_bitop3 it replaced, g | (c1 ^ b1) on a fourth row plane, made the
// headline loop 1.7 % slower).  v_bitop3 truth tables are over src0 = 0xF0,
// src1 = 0xCC, src2 = 0xAA.
__device__ __forceinline__ uint32_t dense3_word(uint32_t u, uint32_t v, uint32_t w,
                                                uint32_t bv, uint32_t bw, uint32_t by) {
  uint32_t s = v ^ bv;
  s = __builtin_amdgcn_bitop3_b32(u, by, s, 0x14);
  return __builtin_amdgcn_bitop3_b32(w, bw, s, 0x4d);
}

// LDS bytes each wave of a staged-store kernel owns.
constexpr int kStageBytes = 4096;

// Orders this wave's LDS writes before its later LDS reads (the LDS serves
// one wave's accesses in order; the fences keep the compiler from moving
// them).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Stores one wave's R x C cells, v[q][c] for row i0 + 64 q + lane and column
// j0 + c (already clamped), as SB-byte cells of the triangle.  Stored
// straight from the registers, each store instruction writes 64 rows, one
// cell each: 64 cache lines.  Here the cells go through the wave's LDS
// (kStageBytes at `stage`, QP row groups of 64 at a time) and each store
// instruction writes 64 / C whole row pieces of C consecutive cells.
template <int R, int C, int SB>
__device__ __forceinline__ void tile_store_staged(const uint32_t (&v)[R][C], uint32_t n,
                                                  uint32_t i0, uint32_t j0, void *out,
                                                  uint64_t cell_offset, uint8_t *stage) {
  using T = typename std::conditional<
      SB == 1, uint8_t, typename std::conditional<SB == 2, uint16_t, uint32_t>::type>::type;
  constexpr int RB = C * SB;                 // bytes of one row piece
  constexpr int QB = kWave * RB;             // bytes of 64 row pieces
  constexpr int QP = QB * R <= kStageBytes ? R : kStageBytes / QB;
  constexpr int RS = kWave / C;              // row pieces per store instruction
  static_assert(QP >= 1 && R % QP == 0, "stage must hold whole row groups");
  static_assert(RB % 8 == 0 && kWave % C == 0, "row pieces of 8-byte multiples");
  const uint32_t lane = threadIdx.x & (kWave - 1);
  T *const st = reinterpret_cast<T *>(stage);
  T *const dst = static_cast<T *>(out);
  const uint32_t sub = lane / C, cl = lane % C;
#pragma unroll
  for (int q0 = 0; q0 < R; q0 += QP) {
    if (q0) wave_lds_sync(); // the previous pass's reads come first
#pragma unroll
    for (int dq = 0; dq < QP; ++dq) {
      uint32_t w[RB / 4];
#pragma unroll
      for (int k = 0; k < RB / 4; ++k) {
        uint32_t x = 0;
#pragma unroll
        for (int e = 0; e < 4 / SB; ++e)
          x |= (uint32_t)(T)v[q0 + dq][k * (4 / SB) + e] << (8 * SB * e);
        w[k] = x;
      }
      uint8_t *row = stage + (dq * kWave + lane) * RB;
      if constexpr (RB % 16 == 0) {
#pragma unroll
        for (int k = 0; k < RB / 16; ++k)
          reinterpret_cast<uint4 *>(row)[k] =
              make_uint4(w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]);
      } else {
#pragma unroll
        for (int k = 0; k < RB / 8; ++k)
          reinterpret_cast<uint2 *>(row)[k] = make_uint2(w[2 * k], w[2 * k + 1]);
      }
    }
    wave_lds_sync();
    // Store s: row piece RS s + sub of this pass, cell cl.  Row i's first
    // cell is i (i - 1) / 2; RS rows further on it is RS i + RS (RS - 1) / 2
    // cells further.
    uint32_t i = i0 + kWave * q0 + sub;
    uint64_t x = (uint64_t)i * (i - 1) / 2 + j0 + cl - cell_offset;
#pragma unroll
    for (int s = 0; s < QP * C; ++s) {
      const T val = st[s * kWave + lane];
      if (i < n && j0 + cl < i) dst[x] = val;
      x += (uint64_t)RS * i + RS * (RS - 1) / 2;
      i += RS;
    }
  }
}

// One wave's share of filling the cells of rows [i0, i1), one contiguous run
// of the triangle (cells i0 (i0 - 1) / 2 .. i1 (i1 - 1) / 2 - 1, less
// cell_offset), with max_dist as SB-byte cells: share `part` of `parts` of
// the run's 16-byte units, streaming stores, and with part 0 the cells before
// the first unit and after the last.
template <int SB>
__device__ __forceinline__ void fill_rows_share(void *out, uint32_t i0, uint32_t i1,
                                                uint64_t cell_offset, uint32_t max_dist,
                                                uint32_t part, uint32_t parts) {
  using T = typename std::conditional<
      SB == 1, uint8_t, typename std::conditional<SB == 2, uint16_t, uint32_t>::type>::type;
  using u32x4 = uint32_t __attribute__((ext_vector_type(4)));
  constexpr uint32_t EPU = 16 / SB;
  const uint32_t lane = threadIdx.x & (kWave - 1);
  T *const dst = static_cast<T *>(out);
  const uint64_t e0 = (uint64_t)i0 * (i0 ? i0 - 1 : 0) / 2 - cell_offset;
  const uint64_t e1 = (uint64_t)i1 * (i1 ? i1 - 1 : 0) / 2 - cell_offset;
  const uint64_t mis = ((uintptr_t)dst / SB + e0) % EPU;
  const uint64_t a = min(e1, e0 + (mis ? EPU - mis : 0));
  const uint64_t units = (e1 - a) / EPU;
  const uint64_t tail = a + units * EPU;
  const T v = (T)max_dist;
  uint32_t w = v;
  if (SB == 1) w *= 0x01010101u;
  else if (SB == 2) w |= w << 16;
  if (part == 0) {
    if (e0 + lane < a) dst[e0 + lane] = v;
    if (tail + lane < e1) dst[tail + lane] = v;
  }
  const u32x4 q = {w, w, w, w};
  u32x4 *const u = reinterpret_cast<u32x4 *>(dst + a);
  const uint64_t u0 = units * part / parts, u1 = units * (part + 1) / parts;
  // streaming stores: nothing reads the cells back in this launch
  for (uint64_t k = u0 + lane; k < u1; k += kWave) __builtin_nontemporal_store(q, u + k);
}

// One wave's share of a tile: rows i0 + 64 q + lane (q < R) against columns
// j0 .. j0 + C - 1, stored into the triangle at cell i(i-1)/2 + j - cell_offset
// as out_bytes-wide cells (1, 2 or 4; a constant where the caller's is).
// corr (codes with has_corr only): the correction counts of the planes' rows
// (dense3_corr_words / gap_corr_words); check_every is W or kCheckWords.
// EE = false compiles the loop without early-exit votes (check_every is then
// ignored).  SB = out_bytes (1, 2 or 4) stores the cells through `stage`,
// kStageBytes of LDS owned by this wave (tile_store_staged); SB = 0 stores
// them straight from the accumulators.  fill_parts > 0 (kDense3 with votes,
// SB != 0, every tile of the row block in the launch): when every row of the
// tile's row block holds at least max_dist invalid sites, the wave writes
// share fill_part of fill_parts of the block's cells instead of its own
// (fill_rows_share).
template <int P, int R, int C, bool EE = true, int SB = 0>
__device__ __forceinline__ void tile_wave(const uint32_t *__restrict__ planes, uint32_t n,
                                          uint32_t W, uint32_t i0, uint32_t j0,
                                          void *__restrict__ out, uint32_t out_bytes,
                                          uint64_t cell_offset, uint32_t max_dist,
                                          uint32_t check_every,
                                          const uint32_t *__restrict__ corr,
                                          uint32_t n_pad, uint8_t *stage = nullptr,
                                          uint32_t fill_part = 0, uint32_t fill_parts = 0) {
  constexpr bool D3 = P == kDense3;
  constexpr bool GAP = P == kGap2 || P == kGap3;
  constexpr bool CLS = P == kClass3;
  constexpr bool OFFSET = D3 || GAP;     // accumulators carry OFF + counts
  constexpr int PS = stored_planes(P);   // planes per row and word
  constexpr int GB = P == kGap2 ? 2 : 3; // kGap*: code bits; plane GB = '-'
  constexpr int PA = D3 ? 3 : GAP ? GB : PS; // row planes used by the main loop
  constexpr int PB = D3 ? 3 : PA;        // column planes used
  constexpr int B0 = D3 ? 1 : 0;         // first column plane used
  // Offset codes: accumulators carry OFF + popcounts + the row and column
  // counts added so far.  Rows have fewer than 2^23 sites (kBufWords; the
  // codes are only chosen then), so every partial sum of a pair stays
  // within [OFF - 2^24, OFF + 3 * 2^23], the cap can be clamped to 2^28 and
  // pairs outside the triangle start at kD3Idle, which stays above OFF + md.
  constexpr uint32_t OFF = OFFSET ? 0x80000000u : 0u;
  constexpr uint32_t kD3Idle = 0xD0000000u;
  const uint32_t md = OFFSET ? min(max_dist, 0x10000000u) : max_dist;
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t i_end = min(i0 + (uint32_t)(kWave * R), n);

  // kDense3 with votes: an invalid site counts against every site of every
  // other row ((0 & b) == 0), so d(i, j) >= max(Z_i, Z_j) with Z the row's
  // invalid sites.  A pair whose bound reaches the cap is settled before the
  // loop: it starts at kD3Idle like a pair outside the triangle and stores
  // max_dist (the bound applies only when max_dist <= Z < 2^23, so md is
  // max_dist).  A wave whose pairs are all settled skips the loop.  When
  // every row of the row block is settled, all the block's cells are
  // max_dist and the wave fills its share of them (every wave of the block's
  // tiles holds the same rows and takes the same branch).
  const bool bound = D3 && EE && check_every < W;
  uint32_t zi[R];
  if (bound) {
    bool rows_settled = true;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const uint32_t i = i0 + kWave * q + lane;
      zi[q] = corr[min(i, n_pad - 1)];
      rows_settled = rows_settled && (i >= n || zi[q] >= md);
    }
    if constexpr (SB != 0) {
      if (fill_parts && __all(rows_settled)) {
        fill_rows_share<SB>(out, i0, i_end, cell_offset, max_dist, fill_part, fill_parts);
        return;
      }
    }
  }
  if (j0 + 1 >= i_end) return; // no pair j < i inside this wave's block

  // Pairs outside the triangle start saturated so that the early-exit vote
  // only waits for real pairs; they are never stored.
  uint32_t acc[R][C];
  // Adds row counts zc[row] - gc[column] (kDense3) or -zc[row] - gc[column]
  // (kGap*) to every accumulator.  The column counts are read into VGPRs
  // (the offset `vzero` is an opaque per-lane 0), which keeps them out of
  // the SGPR file the column planes fill.
  auto add_counts = [&](const uint32_t *__restrict__ zc, const uint32_t *__restrict__ gc) {
    uint32_t vzero;
    __asm__ volatile("v_mov_b32 %0, 0" : "=v"(vzero));
    uint32_t gj[C];
#pragma unroll
    for (int c = 0; c < C; ++c) gj[c] = gc[j0 + c + vzero];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const uint32_t z = zc[min(i0 + kWave * q + lane, n_pad - 1)];
      const uint32_t zi = GAP ? 0u - z : z;
#pragma unroll
      for (int c = 0; c < C; ++c) acc[q][c] += zi - gj[c];
    }
  };
  bool settled = false;
  if constexpr (OFFSET && EE) {
    uint32_t zj[C];
    if (bound) {
      uint32_t vzero;
      __asm__ volatile("v_mov_b32 %0, 0" : "=v"(vzero));
#pragma unroll
      for (int c = 0; c < C; ++c) zj[c] = corr[j0 + c + vzero];
    }
    uint32_t lo = kD3Idle;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const uint32_t i = i0 + kWave * q + lane;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const bool open = j0 + c < i && i < n && !(bound && max(zi[q], zj[c]) >= md);
        acc[q][c] = open ? OFF : kD3Idle;
        lo = min(lo, acc[q][c]);
      }
    }
    if (bound) settled = __all(lo >= OFF + md);
  } else if constexpr (OFFSET) {
    // No votes: the whole-row counts go in at the start.
    uint32_t gj[C];
#pragma unroll
    for (int c = 0; c < C; ++c) gj[c] = corr[(GAP ? 0 : n_pad) + j0 + c];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const uint32_t z = corr[min(i0 + kWave * q + lane, n_pad - 1)];
      const uint32_t zi = GAP ? 0u - z : z;
#pragma unroll
      for (int c = 0; c < C; ++c) acc[q][c] = OFF + zi - gj[c];
    }
  } else {
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const uint32_t i = i0 + kWave * q + lane;
#pragma unroll
      for (int c = 0; c < C; ++c)
        acc[q][c] = (j0 + c < i && i < n) ? 0u : max_dist;
    }
  }

  // Row group g starts at word g * W * PS * 64; one step of it is PS * 64 words.
  const size_t group = (size_t)W * PS * kWave;
  // The planes are read-only for the whole kernel: reading them through the
  // constant address space lets the wave-uniform column words be scalar
  // loads even where the pointer itself came from memory.
  using cptr = const __attribute__((address_space(4))) uint32_t *;
  const cptr pb0 = (cptr)(planes + (size_t)(j0 / kWave) * group + (j0 % kWave));
  cptr pa = (cptr)(planes + (size_t)(i0 / kWave) * group + lane), pb = pb0;
  // The offset codes (rows of fewer than kBufWords words) read the rows
  // through one buffer resource at the tile's first row group: the per-lane
  // offsets stay fixed, the step's offset is a scalar, so no VALU op goes
  // into addresses.  Every offset is below R * group * 4 < 2^31 bytes.
  constexpr bool BUF = OFFSET;
  const __amdgpu_buffer_rsrc_t rows_rsrc = __builtin_amdgcn_make_buffer_rsrc(
      (void *)(planes + (size_t)(i0 / kWave) * group), (short)0, -1, kBufferRsrcWord3);
  int row_off[R];
#pragma unroll
  for (int q = 0; q < R; ++q) row_off[q] = (int)((lane + (uint32_t)(q * group)) * 4);
  uint32_t k = settled ? W : 0, chunk = 0;
  while (k < W) {
    const uint32_t k_end = EE ? min(W, k + check_every) : W;
    for (; k < k_end; ++k) {
      uint32_t a[R][PA];
      const int step_off = (int)(k * (uint32_t)(PS * kWave * 4));
#pragma unroll
      for (int q = 0; q < R; ++q)
#pragma unroll
        for (int p = 0; p < PA; ++p)
          a[q][p] = BUF ? __builtin_amdgcn_raw_buffer_load_b32(rows_rsrc, row_off[q],
                                                               step_off + p * kWave * 4, 0)
                        : pa[q * group + p * kWave];
      uint32_t bw[C][PB];
#pragma unroll
      for (int p = 0; p < PB; ++p)
#pragma unroll
        for (int c = 0; c < C; ++c) bw[c][p] = pb[(B0 + p) * kWave + c];
#pragma unroll
      for (int c = 0; c < C; ++c) {
        // The column's words are copied into VGPRs first (one v_mov_b32 per
        // column plane and step): v_bitop3 with three VGPR operands issues
        // up to twice per CU and clock on gfx950, with an SGPR operand once.
        // The dense 3-op code at n = 50,000, L = 29,903 went from 111.0 to
        // 99.5 ms; the gap-3 code at n = 100,000, L = 10,000 from 143.8 to
        // 124.3 ms and the nibble and one-hot codes at n = 30,000, L =
        // 10,000 from 17.0 to 15.1 and 20.4 to 17.9 ms, each with its
        // fastest tile shape before and after (the gap and pointer-addressed
        // codes gained only once the copies were ordered per column block,
        // below: issued all at the start of the step they held a VGPR each
        // for the whole step and cost a wave per SIMD).
        // With at most 2 rows per lane every code's first op (an XOR or AND,
        // a VOP2 instruction) takes its column word as an SGPR operand
        // instead: the dense code on 128 x 64 tiles at n = 50,000, L = 29,903
        // 92.6 -> 90.4 ms, on 128 x 32 tiles at n = 6,000 1.68 -> 1.61 ms;
        // the nibble and one-hot codes on 128 x 32 tiles at n = 1,000-3,000,
        // L = 10,000 2-3 % faster; with 4 rows per lane it was slower (92.5
        // -> 101.9 ms), and 2-row tiles without any copy slower still (117.1
        // ms).
        constexpr bool SXOR = R <= 2;
        uint32_t b[PB];
#pragma unroll
        for (int p = 0; p < PB; ++p)
          b[p] = SXOR && p == 0 ? bw[c][p] : to_vgpr_ordered(bw[c][p]);
#pragma unroll
        for (int q = 0; q < R; ++q) {
          // v_bitop3 truth tables are over src0 = 0xF0, src1 = 0xCC,
          // src2 = 0xAA.
          uint32_t s;
          if constexpr (D3) {
            s = dense3_word(a[q][0], a[q][1], a[q][2], b[0], b[1], b[2]);
          } else if constexpr (GAP) {
            // codes differ: XOR, then 0xBE = (x ^ y) | z per further bit
            s = a[q][0] ^ b[0];
#pragma unroll
            for (int p = 1; p < GB; ++p)
              s = __builtin_amdgcn_bitop3_b32(a[q][p], b[p], s, 0xBE);
          } else if constexpr (CLS) {
            // code bits differ (0xBE = (x^y)|z), then z & ~x & ~y (0x02)
            // with the two '-' planes
            s = a[q][0] ^ b[0];
#pragma unroll
            for (int p = 1; p < PS - 1; ++p)
              s = __builtin_amdgcn_bitop3_b32(a[q][p], b[p], s, 0xBE);
            s = __builtin_amdgcn_bitop3_b32(a[q][PS - 1], b[PS - 1], s, 0x02);
          } else {
            // s = a0&b0; s = (ap&bp)|s ...; s = ~((a_last&b_last)|s):
            // 0xEA = (x&y)|z, 0x15 = its NOT.
            s = a[q][0] & b[0];
#pragma unroll
            for (int p = 1; p < P - 1; ++p)
              s = __builtin_amdgcn_bitop3_b32(a[q][p], b[p], s, 0xEA);
            s = __builtin_amdgcn_bitop3_b32(a[q][P - 1], b[P - 1], s, 0x15);
          }
          acc[q][c] += __builtin_popcount(s);
        }
        // Column blocks of 4 rows x 1 column (R >= 4) or 2 rows x 4 columns:
        // the dense 3-op code at n = 50,000, L = 29,903 went from 101.0 to
        // 97.7 ms on 256 x 64 tiles, from 103.4 to 99.1 ms on 128 x 64 (the
        // latter 104.5 ms with blocks of 2 x 8, the former 101.4 with 4 x 2).
        constexpr int G = R >= 4 ? 1 : (4 < C ? 4 : C);
        if ((c + 1) % G == 0) pin_chains<R, C, G>(acc, c + 1 - G);
      }
      pa += PS * kWave;
      pb += PS * kWave;
    }
    if constexpr (OFFSET && EE) {
      if (check_every < W) {
        // The barrier keeps the count loads here, after the chunk, instead
        // of holding them in registers across it.
        __asm__ volatile("" ::: "memory");
        const uint64_t T = (W + kCheckWords - 1) / kCheckWords;
        const uint32_t *zc = corr + (2 + chunk) * (uint64_t)n_pad;
        add_counts(zc, D3 ? corr + (2 + T + chunk) * (uint64_t)n_pad : zc);
      }
    }
    ++chunk;
    if (EE && check_every < W) {
      uint32_t lo = acc[0][0];
#pragma unroll
      for (int q = 0; q < R; ++q)
#pragma unroll
        for (int c = 0; c < C; ++c) lo = min(lo, acc[q][c]);
      if (__all(lo >= OFF + md)) break;
    }
  }
  if constexpr (OFFSET && EE) {
    if (check_every >= W) { // one chunk: the whole-row counts
      __asm__ volatile("" ::: "memory");
      add_counts(corr, GAP ? corr : corr + n_pad);
    }
  }
  if constexpr (GAP) {
    // + 2 |g_i & g_j| over the words where column j has a '-': per column, a
    // wave-uniform loop over the chunks its summary word marks and the set
    // bits of their bitmap words.  After an early exit every real pair's
    // count is already a lower bound >= md, and adding to it keeps it there.
    __asm__ volatile("" ::: "memory");
    const uint32_t T = (W + kCheckWords - 1) / kCheckWords;
    const uint32_t span = gap_summary_span(T);
    const cptr bm = (cptr)(corr + (2 + (uint64_t)T) * n_pad);
    const cptr summary = (cptr)(corr + n_pad);
    uint32_t sm[C];
#pragma unroll
    for (int c = 0; c < C; ++c) sm[c] = summary[j0 + c];
    // The rows' base address is recomputed from a lane id the compiler
    // cannot tie to `lane`, so no copy of it stays live across the main loop.
    const uint32_t lane_b = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    const cptr pr = (cptr)(planes + (size_t)(i0 / kWave) * group + lane_b);
#pragma unroll
    for (int c = 0; c < C; ++c) {
      uint32_t chunks = j0 + c < n ? sm[c] : 0u;
      while (chunks) {
        const uint32_t t0 = (uint32_t)__builtin_ctz(chunks) * span;
        chunks &= chunks - 1;
        for (uint32_t t = t0; t < min(T, t0 + span); ++t) {
          uint32_t u = bm[(uint64_t)t * n_pad + j0 + c];
          while (u) {
            const uint32_t w = t * kCheckWords + (uint32_t)__builtin_ctz(u);
            u &= u - 1;
            const uint64_t step = ((uint64_t)w * PS + GB) * kWave;
            const uint32_t gb = pb0[step + c];
#pragma unroll
            for (int q = 0; q < R; ++q) {
              const uint32_t x = pr[q * group + step] & gb;
              acc[q][c] += __builtin_popcount(x);
              acc[q][c] += __builtin_popcount(x);
            }
          }
        }
      }
    }
  }

  if constexpr (SB != 0) {
#pragma unroll
    for (int q = 0; q < R; ++q)
#pragma unroll
      for (int c = 0; c < C; ++c) acc[q][c] = min(acc[q][c] - OFF, max_dist);
    tile_store_staged<R, C, SB>(acc, n, i0, j0, out, cell_offset, stage);
    return;
  }
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const uint32_t i = i0 + kWave * q + lane;
    if (i < n) {
      const uint64_t x = (uint64_t)i * (i - 1) / 2 + j0 - cell_offset;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        if (j0 + c < i) {
          const uint32_t v = min(acc[q][c] - OFF, max_dist);
          if (out_bytes == 1) static_cast<uint8_t *>(out)[x + c] = (uint8_t)v;
          else if (out_bytes == 2) static_cast<uint16_t *>(out)[x + c] = (uint16_t)v;
          else static_cast<uint32_t *>(out)[x + c] = v;
        }
      }
    }
  }
}

// Tile t of a launch over tiles [tile_begin, tile_begin + n_tiles).
template <int P, int R, int C, int NW, typename OutT, bool EE>
__device__ __forceinline__ void allpairs_tile(
    const uint32_t *__restrict__ planes, uint32_t n, uint32_t W, uint64_t tile_begin,
    uint32_t n_tiles, OutT *__restrict__ out, uint64_t cell_offset, uint32_t max_dist,
    uint32_t check_every, const uint32_t *__restrict__ corr, uint32_t n_pad) {
  constexpr uint32_t TI = kWave * R;
  constexpr uint32_t TJ = NW * C;
  static_assert(TI % TJ == 0, "tile rows must be a multiple of tile columns");

  // The dispatcher hands consecutive workgroups to the 8 XCDs in turn; map
  // them so that each XCD walks a contiguous run of tiles (same row block,
  // neighbouring column blocks) and shares the row data in its own L2.
  const uint32_t g = blockIdx.x;
  const uint32_t q8 = n_tiles >> 3, r8 = n_tiles & 7u, x = g & 7u;
  const uint64_t t = tile_begin + (uint64_t)x * q8 + (x < r8 ? x : r8) + (g >> 3);
  uint32_t b, cb;
  tile_coords(t, TI / TJ, b, cb);
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  if constexpr (EE) {
    // Early-exit tiles stop after a few chunks, and their stores weigh: they
    // go through LDS.  Tiles that run every word store straight from the
    // registers (measured as fast, with no LDS held).
    __shared__ __attribute__((aligned(16))) uint8_t stage[NW][kStageBytes];
    // Every tile of a row block is in the launch (launches cover whole row
    // blocks), so the block's M (b + 1) tiles x NW waves share its fill.
    tile_wave<P, R, C, EE, (int)sizeof(OutT)>(planes, n, W, b * TI, cb * TJ + wave * C, out,
                                              sizeof(OutT), cell_offset, max_dist,
                                              check_every, corr, n_pad, stage[wave],
                                              cb * NW + wave, (TI / TJ) * (b + 1) * NW);
  } else {
    tile_wave<P, R, C, EE>(planes, n, W, b * TI, cb * TJ + wave * C, out, sizeof(OutT),
                           cell_offset, max_dist, check_every, corr, n_pad);
  }
}

// Waves per SIMD a kernel is held to (1: no limit).  The vote-compiled
// 128 x 64 tiles at 6 (74 VGPRs; the dense 3-op code took 85 and ran 5):
// 3-7 % faster with the early exit at n = 50,000, L = 29,903, settled or
// not.  Holding the 128 x 32 ones to 8 (43 VGPRs instead of 66) made their
// unsettled launches 2-4 % slower.
template <int R, int C, bool EE>
constexpr int allpairs_min_waves() {
  return EE && R == 2 && C == 16 ? 6 : 1;
}

template <int P, int R, int C, int NW, typename OutT, bool EE>
__global__ __launch_bounds__(kWave *NW, (allpairs_min_waves<R, C, EE>())) void k_allpairs(
    const uint32_t *__restrict__ planes, uint32_t n, uint32_t W, uint64_t tile_begin,
    uint32_t n_tiles, OutT *__restrict__ out, uint64_t cell_offset, uint32_t max_dist,
    uint32_t check_every, const uint32_t *__restrict__ corr, uint32_t n_pad) {
  allpairs_tile<P, R, C, NW, OutT, EE>(planes, n, W, tile_begin, n_tiles, out, cell_offset,
                                       max_dist, check_every, corr, n_pad);
}

// The kGap codes' kernel, held to 4 waves per SIMD (at most 128 VGPRs):
// left free, the compiler spends registers on the tile loop's schedule and
// runs 3 waves per SIMD, which measured slower.
template <int P, int R, int C, int NW, typename OutT, bool EE>
__global__ __launch_bounds__(kWave *NW, 4) void k_allpairs_gap(
    const uint32_t *__restrict__ planes, uint32_t n, uint32_t W, uint64_t tile_begin,
    uint32_t n_tiles, OutT *__restrict__ out, uint64_t cell_offset, uint32_t max_dist,
    uint32_t check_every, const uint32_t *__restrict__ corr, uint32_t n_pad) {
  allpairs_tile<P, R, C, NW, OutT, EE>(planes, n, W, tile_begin, n_tiles, out, cell_offset,
                                       max_dist, check_every, corr, n_pad);
}

// ---------------------------------------------------------------------------
// encode: packed nibble rows -> bit-sliced planes
// ---------------------------------------------------------------------------
// A 256-thread workgroup encodes one tile: row group g (64 rows) by words
// k0 .. k0 + kEncodeWords - 1 (512 bytes of each row; one chunk of
// kCheckWords words).  The rows' bytes are read coalesced (32 threads per
// row segment) into LDS; then lane r of each wavefront takes row r,
// wavefront w its words w, w + 4, ..., and each store of one word of one
// plane for the 64 rows is 256 contiguous bytes.  For the codes with
// correction counts the tile also writes its chunk's counts (the layout of
// dense3_corr_words / gap_corr_words); the whole-row slots are summed
// afterwards (corr_totals_row).  Rows >= n and sites past the row end
// encode as '-' (nibble 15); the kGap codes keep the latter out of their
// '-' plane.
constexpr int kEncodeWords = 32;
// LDS row stride in bytes: 16 bytes of padding make lanes reading the same
// word of consecutive rows fall into different banks.
constexpr int kEncodeStride = kEncodeWords * 16 + 16;
constexpr int kEncodeLdsBytes = kWave * kEncodeStride;

// Site order inside a plane word.  The 16 bytes of a 32-site word hold site
// 8 k + j as nibble j of byte word k (k < 4, j < 8).  Exchanging the word
// index bits with the nibble's bit index (two rounds of masked swaps between
// word pairs) leaves bit p of that nibble in plane word p at bit 4 j + k:
// every plane word holds its 32 sites in that fixed order, the same for every
// row, and nothing outside the encoder depends on it (plane words meet only
// their counterparts of other rows in bitwise ops and popcounts).
__device__ __forceinline__ void nibble_planes(const uint32_t (&bytes)[4], uint32_t (&n)[4]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) n[k] = bytes[k];
  // a holds the pair's lower word index, b the higher one; bits of a whose
  // index has `sh` set trade places with bits of b whose index has not.
  auto swap_bits = [](uint32_t &a, uint32_t &b, int sh, uint32_t m) {
    const uint32_t t = ((a >> sh) ^ b) & m;
    b ^= t;
    a ^= t << sh;
  };
  swap_bits(n[0], n[1], 1, 0x55555555u);
  swap_bits(n[2], n[3], 1, 0x55555555u);
  swap_bits(n[0], n[2], 2, 0x33333333u);
  swap_bits(n[1], n[3], 2, 0x33333333u);
}

// Bits of the first `count` sites of a plane word (count <= 32), in the site
// order of nibble_planes.
__device__ __forceinline__ uint32_t first_sites_mask(uint32_t count) {
  uint32_t m = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t c = min(8u, count > 8u * k ? count - 8u * k : 0u);
    m |= (c >= 8 ? 0x11111111u : 0x11111111u & ((1u << (4 * c)) - 1u)) << k;
  }
  return m;
}

// Sites whose nibble is V, from the nibble planes.
template <int V>
__device__ __forceinline__ uint32_t nibble_is(const uint32_t (&n)[4]) {
  return ((V & 1) ? n[0] : ~n[0]) & ((V & 2) ? n[1] : ~n[1]) & ((V & 4) ? n[2] : ~n[2]) &
         ((V & 8) ? n[3] : ~n[3]);
}

// Table planes: plane p holds the sites whose nibble v has bit p of tbl[v]
// set.  Nibble values whose entry is 0 add nothing and are skipped (the
// table is the same for every lane).
template <int V>
__device__ __forceinline__ void table_term(const uint32_t (&n)[4], int P, const uint8_t *s_tbl,
                                           uint32_t (&w)[6]) {
  const uint32_t code = __builtin_amdgcn_readfirstlane(s_tbl[V]);
  if (!code) return;
  const uint32_t x = nibble_is<V>(n);
#pragma unroll
  for (int p = 0; p < 6; ++p)
    if (p < P && ((code >> p) & 1u)) w[p] |= x;
}

template <int... V>
__device__ __forceinline__ void table_planes(const uint32_t (&n)[4], int P, const uint8_t *s_tbl,
                                             uint32_t (&w)[6],
                                             std::integer_sequence<int, V...>) {
  (table_term<V>(n, P, s_tbl, w), ...);
}

// How the planes are built: through the 16-entry table, as the nibble itself
// (4 planes), or as the kDense3 planes derived from the nibble bits (valid
// for nibbles 0, 1, 2, 4, 8 and 15 only).
enum EncodeMode { kEncTable = 0, kEncIdentity = 1, kEncDense3 = 2 };

// Planes of one 32-site word (16 bytes, sites 2b and 2b + 1 in byte b) into
// w[0 .. P-1], sites in the order of nibble_planes.
__device__ __forceinline__ void encode_word(const uint32_t (&bytes)[4], int P,
                                            const uint8_t *s_tbl, int mode, uint32_t (&w)[6]) {
  uint32_t n[4];
  nibble_planes(bytes, n);
  if (mode == kEncDense3) {
    // Among 0, A=1, C=2, G=4, T=8 and '-'=15 only '-' has bits 0 and 1, and
    // only the invalid code has no bit.
    const uint32_t inv = ~(n[0] | n[1] | n[2] | n[3]);
    const uint32_t gt = (n[2] | n[3]) & ~n[0];    // G T
    w[0] = (n[0] ^ n[1]) | inv;                   // u: A C invalid
    w[1] = ((n[1] | n[3]) & ~n[0]) | inv;         // v: C T invalid
    w[2] = (n[0] & n[1]) | inv;                   // w: '-' invalid
    w[3] = gt;                                    // y: G T
    return;
  }
  if (mode == kEncIdentity) {
#pragma unroll
    for (int p = 0; p < 4; ++p) w[p] = n[p];
    return;
  }
#pragma unroll
  for (int p = 0; p < 6; ++p) w[p] = 0;
  table_planes(n, P, s_tbl, w, std::make_integer_sequence<int, 16>{});
}

// The tile (g, k0) of a dataset of n rows of bpr bytes, W words, P stored
// planes, plane code `op`; called by all 256 threads of a workgroup with
// `lds` holding kEncodeLdsBytes bytes.  corr: the correction counts after
// the planes when has_corr(op), else null.
__device__ __forceinline__ void encode_tile(const uint8_t *__restrict__ packed, uint32_t n,
                                            uint64_t bpr, uint32_t W, int P, int op,
                                            const uint8_t *s_tbl, int mode,
                                            uint32_t *__restrict__ planes,
                                            uint32_t *__restrict__ corr, uint32_t n_pad,
                                            uint32_t g, uint32_t k0, uint8_t *lds) {
  static_assert(kEncodeWords == (int)kCheckWords, "one encode tile per count chunk");
  __shared__ uint32_t s_cnt[2][256 / kWave][kWave];
  const uint32_t tid = threadIdx.x;
#pragma unroll
  for (int it = 0; it < kWave * kEncodeWords / 256; ++it) {
    const uint32_t c = tid + 256u * it, r = c / kEncodeWords, part = c % kEncodeWords;
    const uint32_t row = g * kWave + r, k = k0 + part;
    uint2 lo = make_uint2(~0u, ~0u), hi = lo;
    if (row < n && k < W) {
      const uint8_t *src = packed + (uint64_t)row * bpr + (uint64_t)k * 16;
      const uint64_t off = (uint64_t)k * 16;
      if (off + 8 <= bpr) lo = *reinterpret_cast<const uint2 *>(src);
      if (off + 16 <= bpr) hi = *reinterpret_cast<const uint2 *>(src + 8);
    }
    *reinterpret_cast<uint4 *>(lds + r * kEncodeStride + part * 16) =
        make_uint4(lo.x, lo.y, hi.x, hi.y);
  }
  __syncthreads();
  const uint32_t r = tid % kWave, w0 = tid / kWave;
  const bool d3 = op == kDense3;
  const int gp = op == kGap2 ? 2 : 3; // '-' plane of the kGap codes
  // kDense3: a = invalid sites, b = '-' sites; kGap*: a = '-' sites, b =
  // bitmap of the words holding a '-' (bit k - k0).
  uint32_t ca = 0, cb = 0;
  for (uint32_t part = w0; part < (uint32_t)kEncodeWords; part += 256 / kWave) {
    const uint32_t k = k0 + part;
    if (k >= W) break;
    const uint4 v = *reinterpret_cast<const uint4 *>(lds + r * kEncodeStride + part * 16);
    const uint32_t bytes[4] = {v.x, v.y, v.z, v.w};
    uint32_t w[6];
    encode_word(bytes, P, s_tbl, mode, w);
    if (!d3 && corr) {
      // kGap*: sites past the packed row leave the '-' plane (their '-'
      // code, the same in every row, stays).
      const uint64_t first = (uint64_t)k * 32, sites = 2 * bpr;
      if (first + 32 > sites) w[gp] &= first < sites ? first_sites_mask(sites - first) : 0u;
    }
    uint32_t *dst = planes + (((uint64_t)g * W + k) * P) * kWave + r;
#pragma unroll
    for (int p = 0; p < 6; ++p)
      if (p < P) dst[p * kWave] = w[p];
    if (corr) {
      if (d3) {
        ca += __builtin_popcount(w[0] & w[1] & w[2]); // invalid: u, v and w
        cb += __builtin_popcount(w[2] & ~w[0]);       // '-': w without u
      } else {
        const uint32_t gw = gp == 2 ? w[2] : w[3];
        ca += __builtin_popcount(gw);
        cb |= (gw != 0 ? 1u : 0u) << part;
      }
    }
  }
  if (!corr) return;
  s_cnt[0][w0][r] = ca;
  s_cnt[1][w0][r] = cb;
  __syncthreads();
  if (w0 == 0) {
    uint32_t a = s_cnt[0][0][r], b = s_cnt[1][0][r];
#pragma unroll
    for (int q = 1; q < 256 / kWave; ++q) {
      a += s_cnt[0][q][r];
      b = d3 ? b + s_cnt[1][q][r] : (b | s_cnt[1][q][r]);
    }
    const uint64_t T = (W + kCheckWords - 1) / kCheckWords, t = k0 / kCheckWords;
    const uint64_t row = (uint64_t)g * kWave + r;
    corr[(2 + t) * n_pad + row] = a;
    corr[(2 + T + t) * n_pad + row] = b;
  }
}

// Whole-row slots of the correction counts of `row` from its chunk counts:
// kDense3 Z and G, kGap* G and the chunk summary S of its bitmap.
__device__ __forceinline__ void corr_totals_row(uint32_t *__restrict__ corr, uint32_t W,
                                                uint32_t n_pad, int op, uint32_t row) {
  const uint32_t T = (W + kCheckWords - 1) / kCheckWords;
  const uint32_t span = gap_summary_span(T);
  uint32_t a = 0, b = 0, s = 0;
  for (uint32_t t = 0; t < T; ++t) {
    a += corr[(2 + (uint64_t)t) * n_pad + row];
    const uint32_t v = corr[(2 + (uint64_t)T + t) * n_pad + row];
    b += v;
    s |= (v != 0 ? 1u : 0u) << (t / span);
  }
  corr[row] = a;
  corr[n_pad + row] = op == kDense3 ? b : s;
}

// Number of tiles of the rows blocks [0, rb) for an M column blocks per row
// block tiling.
inline uint64_t tiles_before(uint64_t rb, uint64_t M) {
  return M * rb * (rb + 1) / 2;
}

// ---------------------------------------------------------------------------
// reference distances on the packed nibble layout (one query, many rows)
// ---------------------------------------------------------------------------
// Sites of one 32-bit word (8 nibbles) of query q and row r that count
// towards the distance:
//   HD_SEM_DENSE  (q & r) == 0                    (src/hamming_impl.cc:131-140)
//   HD_SEM_CHAR_X (q & r) == 0 || (q != r && neither is '-')
//                                                 (src/hamming.cc:62-71)
// Each nibble's bits are smeared into its bit 3 and the bit-3 lanes counted.
// Nibbles are the reference's lookup codes (src/hamming_impl.cc:22-41):
// 0 invalid, 1 2 4 8 for A C G T, 3 for X, 15 for '-'.  Among these only 15
// has both bit 3 and bit 0 set, so '-' is one shift and one AND per side.
template <int SEM>
__device__ __forceinline__ uint32_t site_count32(uint32_t q, uint32_t r) {
  if constexpr (SEM == HD_SEM_DENSE) {
    uint32_t c = q & r;
    c |= c << 1;
    c |= c << 2;
    return __builtin_popcount(~c & 0x88888888u);
  } else {
    uint32_t z = q & r;
    z |= z << 1;
    z |= z << 2; // bit 3: q & r != 0
    uint32_t d = q ^ r;
    d |= d << 1;
    d |= d << 2;                                         // bit 3: q != r
    const uint32_t g = (q & (q << 3)) | (r & (r << 3)); // bit 3: q or r is '-'
    return __builtin_popcount((~z | (d & ~g)) & 0x88888888u);
  }
}

// Largest row, in 8-byte units, a register-resident query covers per lane
// count T: lane l holds units l, l + 64, ..., l + 64 (T - 1).
constexpr int kRegQueryMaxT = 32; // 16 KiB rows (32,768 sites)

// A query held in VGPRs by one wavefront, scanned against rows of the same
// length.  Units past the row end read as all-ones on both sides, which
// count zero sites under either semantics.  distance() issues all T row
// loads (each 512 contiguous bytes per wavefront) before the first use, so a
// wavefront keeps T * 512 bytes in flight; the loads are non-temporal, as
// every row byte is read once (about 7% more bandwidth than cached loads
// for a 750 MB scan).
template <int SEM, int T>
struct RegQuery {
  typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

  uint2 q[T];

  // NT: non-temporal loads (an operand read once, which should not push
  // operands read again out of the caches).
  template <bool NT = false>
  __device__ __forceinline__ void load(const uint2 *__restrict__ query, uint32_t units,
                                       uint32_t lane) {
#pragma unroll
    for (int t = 0; t < T; ++t) {
      const uint32_t u = lane + kWave * t;
      if (u >= units) {
        q[t] = make_uint2(~0u, ~0u);
      } else if constexpr (NT) {
        const u32x2 v = __builtin_nontemporal_load(reinterpret_cast<const u32x2 *>(query) + u);
        q[t] = make_uint2(v.x, v.y);
      } else {
        q[t] = query[u];
      }
    }
  }

  // NT: non-temporal loads (rows read once); false for operands read again
  // soon (a query streamed against many rows held in registers).
  template <bool NT = true>
  __device__ __forceinline__ uint32_t distance(const uint2 *__restrict__ row, uint32_t units,
                                               uint32_t lane) const {
    uint2 a[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
      const uint32_t u = lane + kWave * t;
      if (u < units) {
        if constexpr (NT) {
          const u32x2 v = __builtin_nontemporal_load(reinterpret_cast<const u32x2 *>(row) + u);
          a[t] = make_uint2(v.x, v.y);
        } else {
          a[t] = row[u];
        }
      } else {
        a[t] = make_uint2(~0u, ~0u);
      }
    }
    uint32_t cnt = 0;
    // Steps whose units all lie past the row end count nothing on any lane
    // and are skipped (the test is the same for the whole wavefront).
#pragma unroll
    for (int t = 0; t < T; ++t)
      if ((uint32_t)(kWave * t) < units)
        cnt += site_count32<SEM>(q[t].x, a[t].x) + site_count32<SEM>(q[t].y, a[t].y);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
    return cnt;
  }
};

// Row distance with the query read from memory (rows longer than the
// register-resident form covers).
template <int SEM>
__device__ __forceinline__ uint32_t row_distance_mem(const uint2 *__restrict__ q,
                                                     const uint2 *__restrict__ row,
                                                     uint64_t units, uint32_t lane) {
  uint32_t cnt = 0;
  for (uint64_t u = lane; u < units; u += kWave) {
    const uint2 a = row[u];
    const uint2 b = q[u];
    cnt += site_count32<SEM>(b.x, a.x) + site_count32<SEM>(b.y, a.y);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
  return cnt;
}

// Lanes per unit count T for a row of `units` 8-byte units: 0 when the row
// is longer than the register-resident form covers.
inline int reg_query_t(uint64_t units) {
  for (int t = 4; t <= kRegQueryMaxT; t *= 2)
    if (units <= (uint64_t)kWave * t) return t;
  return 0;
}

} // namespace hd
