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
// in SGPRs (wave-uniform scalar loads), so each VALU op pairs a per-lane row
// word with a broadcast column word and no LDS traffic is needed.  Per pair
// and per 32 sites the work is P-1 AND/AND-OR ops, one NOR-with-AND and one
// popcount-accumulate: 5 ops for the 4-plane dense code.
// ---------------------------------------------------------------------------
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

// One wave's share of a tile: rows i0 + 64 q + lane (q < R) against columns
// j0 .. j0 + C - 1, stored into the triangle at cell i(i-1)/2 + j - cell_offset
// as out_bytes-wide cells (1, 2 or 4; a constant where the caller's is).
template <int P, int R, int C>
__device__ __forceinline__ void tile_wave(const uint32_t *__restrict__ planes, uint32_t n,
                                          uint32_t W, uint32_t i0, uint32_t j0,
                                          void *__restrict__ out, uint32_t out_bytes,
                                          uint64_t cell_offset, uint32_t max_dist,
                                          uint32_t check_every) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t i_end = min(i0 + (uint32_t)(kWave * R), n);
  if (j0 + 1 >= i_end) return; // no pair j < i inside this wave's block

  // Pairs outside the triangle start saturated so that the early-exit vote
  // only waits for real pairs; they are never stored.
  uint32_t acc[R][C];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const uint32_t i = i0 + kWave * q + lane;
#pragma unroll
    for (int c = 0; c < C; ++c)
      acc[q][c] = (j0 + c < i && i < n) ? 0u : max_dist;
  }

  // Row group g starts at word g * W * P * 64; one step of it is P * 64 words.
  const size_t group = (size_t)W * P * kWave;
  // The planes are read-only for the whole kernel: reading them through the
  // constant address space lets the wave-uniform column words be scalar
  // loads even where the pointer itself came from memory.
  using cptr = const __attribute__((address_space(4))) uint32_t *;
  cptr pa = (cptr)(planes + (size_t)(i0 / kWave) * group + lane);
  cptr pb = (cptr)(planes + (size_t)(j0 / kWave) * group + (j0 % kWave));
  uint32_t k = 0;
  while (k < W) {
    const uint32_t k_end = min(W, k + check_every);
    for (; k < k_end; ++k) {
      uint32_t a[R][P];
#pragma unroll
      for (int q = 0; q < R; ++q)
#pragma unroll
        for (int p = 0; p < P; ++p) a[q][p] = pa[q * group + p * kWave];
      uint32_t bw[C][P];
#pragma unroll
      for (int p = 0; p < P; ++p)
#pragma unroll
        for (int c = 0; c < C; ++c) bw[c][p] = pb[p * kWave + c];
#pragma unroll
      for (int q = 0; q < R; ++q)
#pragma unroll
        for (int c = 0; c < C; ++c) {
          // s = a0&b0; s = (ap&bp)|s ...; s = ~((a_last&b_last)|s); one
          // v_bitop3 per plane after the first (truth tables over
          // src0=0xF0, src1=0xCC, src2=0xAA: 0xEA = (x&y)|z, 0x15 = its NOT).
          uint32_t s = a[q][0] & bw[c][0];
#pragma unroll
          for (int p = 1; p < P - 1; ++p)
            s = __builtin_amdgcn_bitop3_b32(a[q][p], bw[c][p], s, 0xEA);
          s = __builtin_amdgcn_bitop3_b32(a[q][P - 1], bw[c][P - 1], s, 0x15);
          acc[q][c] += __builtin_popcount(s);
        }
      pa += P * kWave;
      pb += P * kWave;
    }
    if (check_every < W) {
      uint32_t lo = acc[0][0];
#pragma unroll
      for (int q = 0; q < R; ++q)
#pragma unroll
        for (int c = 0; c < C; ++c) lo = min(lo, acc[q][c]);
      if (__all(lo >= max_dist)) break;
    }
  }

#pragma unroll
  for (int q = 0; q < R; ++q) {
    const uint32_t i = i0 + kWave * q + lane;
    if (i < n) {
      const uint64_t x = (uint64_t)i * (i - 1) / 2 + j0 - cell_offset;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        if (j0 + c < i) {
          const uint32_t v = min(acc[q][c], max_dist);
          if (out_bytes == 1) static_cast<uint8_t *>(out)[x + c] = (uint8_t)v;
          else if (out_bytes == 2) static_cast<uint16_t *>(out)[x + c] = (uint16_t)v;
          else static_cast<uint32_t *>(out)[x + c] = v;
        }
      }
    }
  }
}

template <int P, int R, int C, int NW, typename OutT>
__global__ __launch_bounds__(kWave *NW) void k_allpairs(
    const uint32_t *__restrict__ planes, uint32_t n, uint32_t W, uint64_t tile_begin,
    uint32_t n_tiles, OutT *__restrict__ out, uint64_t cell_offset, uint32_t max_dist,
    uint32_t check_every) {
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
  tile_wave<P, R, C>(planes, n, W, b * TI, cb * TJ + wave * C, out, sizeof(OutT), cell_offset,
                     max_dist, check_every);
}

// ---------------------------------------------------------------------------
// encode: packed nibble rows -> bit-sliced planes
// ---------------------------------------------------------------------------
// Row i, words k0 .. k0+7: reads 128 contiguous bytes of the row and writes
// 8 x P plane words; the 64 rows of a wavefront form one row group, so each
// store is 256 contiguous bytes.  Rows >= n and sites past the row end
// encode as '-' (table entry 15, all planes set), which never counts.
constexpr int kEncodeRun = 8;

// Bit p of each of the 8 nibbles of v, gathered into bits 0..7 (nibble k ->
// bit k): mask the bits, then halve the spacing three times.
__device__ __forceinline__ uint32_t nibble_bits(uint32_t v, int p) {
  uint32_t x = (v >> p) & 0x11111111u;
  x = (x | (x >> 3)) & 0x03030303u;
  x = (x | (x >> 6)) & 0x000F000Fu;
  return (x | (x >> 12)) & 0xFFu;
}

__device__ __forceinline__ void encode_run(const uint8_t *__restrict__ packed, uint32_t n,
                                           uint64_t bpr, uint32_t W, int P,
                                           const uint8_t *s_tbl,
                                           uint32_t *__restrict__ planes, uint32_t i,
                                           uint32_t k0, bool identity = false) {
  const uint8_t *row = packed + (uint64_t)i * bpr;
  if (identity) {
    // The 4-plane dense code is the nibble itself: plane p is bit p of every
    // nibble, with no table lookup.
    for (int kk = 0; kk < kEncodeRun; ++kk) {
      const uint32_t k = k0 + kk;
      if (k >= W) break;
      const uint64_t off = (uint64_t)k * 16;
      uint32_t bytes[4];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const uint64_t o = off + 8 * h;
        if (i < n && o + 8 <= bpr) {
          const uint2 u = *reinterpret_cast<const uint2 *>(row + o);
          bytes[2 * h] = u.x;
          bytes[2 * h + 1] = u.y;
        } else {
          bytes[2 * h] = 0xFFFFFFFFu;
          bytes[2 * h + 1] = 0xFFFFFFFFu;
        }
      }
      uint32_t *dst = planes + (((uint64_t)(i / 64) * W + k) * 4) * 64 + (i % 64);
#pragma unroll
      for (int p = 0; p < 4; ++p)
        dst[p * 64] = nibble_bits(bytes[0], p) | (nibble_bits(bytes[1], p) << 8) |
                      (nibble_bits(bytes[2], p) << 16) | (nibble_bits(bytes[3], p) << 24);
    }
    return;
  }
  for (int kk = 0; kk < kEncodeRun; ++kk) {
    const uint32_t k = k0 + kk;
    if (k >= W) break;
    uint32_t w[6] = {0, 0, 0, 0, 0, 0};
    const uint64_t off = (uint64_t)k * 16;
    uint32_t bytes[4];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const uint64_t o = off + 8 * h;
      if (i < n && o + 8 <= bpr) {
        const uint2 u = *reinterpret_cast<const uint2 *>(row + o);
        bytes[2 * h] = u.x;
        bytes[2 * h + 1] = u.y;
      } else {
        bytes[2 * h] = 0xFFFFFFFFu;
        bytes[2 * h + 1] = 0xFFFFFFFFu;
      }
    }
#pragma unroll
    for (int s = 0; s < 32; ++s) {
      const uint32_t nib = (bytes[s >> 3] >> (4 * (s & 7))) & 15u;
      const uint32_t m = s_tbl[nib];
#pragma unroll
      for (int p = 0; p < 6; ++p)
        if (p < P) w[p] |= ((m >> p) & 1u) << s;
    }
    uint32_t *dst = planes + (((uint64_t)(i / 64) * W + k) * P) * 64 + (i % 64);
#pragma unroll
    for (int p = 0; p < 6; ++p)
      if (p < P) dst[p * 64] = w[p];
  }
}

// Number of tiles of the rows blocks [0, rb) for an M column blocks per row
// block tiling.
inline uint64_t tiles_before(uint64_t rb, uint64_t M) {
  return M * rb * (rb + 1) / 2;
}

} // namespace hd
