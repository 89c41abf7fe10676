/*
 * hammingdist_amd.h — C ABI of the MI355X (gfx950) Hamming-distance engine.
 *
 * This is the process -> device boundary of the all-pairs Hamming distance path.
 * It replaces the C++ CUDA backend of the reference (hammingdist):
 *   include/hamming/distance_cuda.hh:12-29 and src/distance_cuda.cu:9-238.
 * Every entry point takes plain pointers and sizes and returns an int status
 * (HD_OK on success); the message of the last failure on the calling thread is
 * returned by hd_last_error().  No C++ or torch types cross this boundary.
 *
 * Input format ("packed"): the reference's dense encoding (from_string,
 * src/hamming_impl.cc:244-267): one byte holds two sites, site 2k in the low
 * nibble and site 2k+1 in the high nibble, nibble codes A=1 C=2 G=4 T=8 '-'=15,
 * any other char 0 (X=3 only where the caller packed with the include_x table,
 * src/hamming_impl.cc:31-44).  Each sequence is padded with 0xFF to
 * bytes_per_seq bytes (a multiple of 8); all n sequences lie back to back.
 *
 * Output format: the linear lower triangle of the reference (distances<T>,
 * include/hamming/hamming_impl.hh:68-176): cell (i, j), i > j, at index
 * i*(i-1)/2 + j, each cell min(distance, max_dist) stored as uint8/16/32.
 */
#ifndef HAMMINGDIST_AMD_H
#define HAMMINGDIST_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define HD_OK 0
#define HD_ERR_INVALID 1   /* bad argument (message says which)              */
#define HD_ERR_NO_DEVICE 2 /* no HIP device visible                           */
#define HD_ERR_HIP 3       /* a HIP runtime call failed (message has the text) */
#define HD_ERR_CALLBACK 4  /* a caller callback returned non-zero             */
#define HD_ERR_CAPACITY 5  /* an output or admission limit was exceeded       */

/* ---- per-site distance semantics ---------------------------------------- */
/* Each counts, over all sites, the sites where f(code_a, code_b) == 1.       */
/* DENSE: f = ((a & b) == 0) on 4-bit codes.
 *   = distance_cpp  src/hamming_impl.cc:131-140, distance_avx2/avx512/sse2,
 *     the CUDA Dist kernel src/distance_cuda.cu:57-65, and distance() /
 *     fasta_reference_distances() with include_x = false
 *     (src/hamming.cc:54-74, 104-141).                                        */
#define HD_SEM_DENSE 0
/* SPARSE: f = (a != b) && a != '-' && b != '-', X counted as invalid.
 *   = distance_sparse over to_sparse_data(include_x = false)
 *     src/hamming_impl.cc:93-129, 170-188.                                    */
#define HD_SEM_SPARSE 1
/* SPARSE_X: as SPARSE, X (code 3) is its own class.
 *   = distance_sparse over to_sparse_data(include_x = true).                  */
#define HD_SEM_SPARSE_X 2
/* CHAR_X: f = ((a & b) == 0) || (a != b && a != '-' && b != '-'), X = 3.
 *   = distance() / fasta_reference_distances() with include_x = true
 *     src/hamming.cc:54-74, 104-141.                                          */
#define HD_SEM_CHAR_X 3

/* ---- library / device ---------------------------------------------------- */
/* Message of the last failing call on this thread ("" if none).             */
const char *hd_last_error(void);

/* Number of visible HIP devices (0 when none or the runtime is absent).
 * Replaces distance_cuda_have_device(), src/distance_cuda.cu:232-238.        */
int hd_device_count(void);

/* Name of the gfx target of device 0 written into buf (NUL-terminated).      */
int hd_device_arch(char *buf, size_t buf_len);

/* ---- one pair ----------------------------------------------------------- */
/* Distance of two packed sequences of `bytes` bytes each, clamped to
 * max_dist (>= 0).  Replaces distance_cuda(a, b, max_dist),
 * include/hamming/distance_cuda.hh:14-16, src/distance_cuda.cu:225-230.     */
int hd_distance_pair(const uint8_t *a, const uint8_t *b, uint64_t bytes,
                     int semantics, int32_t max_dist, int32_t *out);

/* ---- all pairs, host buffers -------------------------------------------- */
/* Lower triangle of all n*(n-1)/2 distances of n packed sequences into the
 * caller's host buffer host_out (n*(n-1)/2 cells of out_bytes = 1, 2 or 4
 * bytes).  max_dist must be <= the largest value of the cell type.
 * Replaces distances_cuda_8bit / distances_cuda_16bit,
 * include/hamming/distance_cuda.hh:19-25, src/distance_cuda.cu:91-214.      */
int hd_all_pairs(const uint8_t *packed, uint64_t n, uint64_t bytes_per_seq,
                 int semantics, int out_bytes, uint32_t max_dist,
                 void *host_out);

/* Row-band callback of hd_all_pairs_stream: `cells` holds the n_cells
 * consecutive cells of complete rows [row_begin, row_end) of the linear
 * triangle, the first one at index first_cell.  The pointer is valid only
 * during the call.  Return 0 to continue, non-zero to abort.                 */
typedef int (*hd_band_cb)(void *user, const void *cells, uint64_t first_cell,
                          uint64_t n_cells, uint64_t row_begin,
                          uint64_t row_end);

/* As hd_all_pairs, but the triangle is delivered in row bands of about
 * band_cells cells (0 = library default) through cb, in row order, while
 * the device computes and copies the next bands.  Replaces
 * distances_cuda_to_lower_triangular, include/hamming/distance_cuda.hh:27-29,
 * src/distance_cuda.cu:140-191, 216-223.                                    */
int hd_all_pairs_stream(const uint8_t *packed, uint64_t n,
                        uint64_t bytes_per_seq, int semantics, int out_bytes,
                        uint32_t max_dist, uint64_t band_cells, hd_band_cb cb,
                        void *user);

/* ---- reference distances ------------------------------------------------ */
/* Distance of one packed query against m packed rows (uint32, unclamped).
 * rows == NULL scans rows 0..m-1 of `db`; otherwise rows[k] (< n_db) names
 * the row for output k.  semantics is HD_SEM_DENSE (include_x = false) or
 * HD_SEM_CHAR_X.  Replaces the per-sequence loop of
 * fasta_reference_distances, src/hamming.cc:104-141.                         */
int hd_reference_distances(const uint8_t *query, const uint8_t *db,
                           uint64_t n_db, uint64_t bytes_per_seq,
                           const uint64_t *rows, uint64_t m, int semantics,
                           uint32_t *out);

/* ---- choice of semantics on the reference CPU route --------------------- */
/* Per-site counts of the four letters over n packed sequences of L real
 * sites: the input of the reference's sparse/dense heuristic
 * (get_reference_expression, src/hamming_impl.cc:142-168; the frac_diff
 * rule, include/hamming/hamming_impl.hh:104-118).  Writes n_diff, the number
 * of (sequence, site) entries that differ from the per-site consensus letter,
 * and the 16-bit mask of nibble codes present in the first L sites.          */
int hd_consensus_diff(const uint8_t *packed, uint64_t n, uint64_t bytes_per_seq,
                      uint64_t L, uint64_t *n_diff, uint32_t *nibble_mask);

/* ---- sparse (i, j, d) emit ----------------------------------------------- */
/* All cells of a host lower triangle (n rows, out_bytes per cell) with
 * d <= threshold, in row-major order, as (i, j, d) uint32 triplets into
 * triplets[3*k..3*k+2].  *count receives the number found; at most
 * `capacity` triplets are written (HD_ERR_CAPACITY if more exist; *count
 * still holds the total).  The device pass behind DataSet::dump_sparse,
 * include/hamming/hamming.hh:79-114.                                        */
int hd_sparse_emit(const void *cells, uint64_t n, int out_bytes,
                   int64_t threshold, uint32_t *triplets, uint64_t capacity,
                   uint64_t *count);

/* hd_sparse_emit on a device triangle into device triplets (d_triplets
 * holds `capacity` triplets), ordered on `stream`; returns after the last
 * write (the total is known on the host).  Same order, count and capacity
 * rule as hd_sparse_emit; the device form of the dump_sparse selection
 * (include/hamming/hamming.hh:79-114) for triangles that stay in HBM.      */
int hd_sparse_emit_dev(const void *d_cells, uint64_t n, int out_bytes,
                       int64_t threshold, uint32_t *d_triplets, uint64_t capacity,
                       uint64_t *count, void *stream);

/* hd_sparse_emit_dev without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, order, count and
 * capacity rule as hd_sparse_emit_dev.  Its scratch (per 4096-cell block
 * counts and scan sums) is a static device array of the library sized for
 * HD_EMIT_NOALLOC_MAX_CELLS cells (8 MiB), held by each GPU the library runs
 * on until the process ends; every scratch word is written before it is
 * read, so no clearing is needed.  Triangles of more cells are rejected
 * (HD_ERR_INVALID) before any device work.  All calls on one GPU share its
 * array: calls that overlap on one GPU (other threads, other streams) race
 * and must be serialised by the caller.                                    */
#define HD_EMIT_NOALLOC_MAX_CELLS 8589934592ull /* 2^33 cells: n <= 131,072 */
int hd_sparse_emit_dev_noalloc(const void *d_cells, uint64_t n, int out_bytes,
                               int64_t threshold, uint32_t *d_triplets,
                               uint64_t capacity, uint64_t *count, void *stream);

/* ---- device-resident API (device pointers, hipStream_t passed as void*) -- */
/* Layout of the bit-sliced planes buffer an encoded dataset needs:
 * planes[k][p][n_pad] uint32, k < n_words (32 sites each), p < n_planes.    */
int hd_planes_layout(uint64_t n, uint64_t bytes_per_seq, int semantics,
                     uint32_t nibble_mask, uint64_t *n_pad, uint32_t *n_words,
                     uint32_t *n_planes, uint64_t *total_words);

/* Bit-mask of the nibble codes present in n packed device rows.             */
int hd_nibble_mask_dev(const uint8_t *d_packed, uint64_t n,
                       uint64_t bytes_per_seq, uint32_t *mask, void *stream);

/* Encode n packed device rows into d_planes (hd_planes_layout words).       */
int hd_encode_dev(const uint8_t *d_packed, uint64_t n, uint64_t bytes_per_seq,
                  int semantics, uint32_t nibble_mask, uint32_t *d_planes,
                  void *stream);

/* All-pairs triangle of an encoded dataset into device memory d_out.
 * variant < 0 picks the tuned tile shape; early_exit != 0 stops a tile once
 * every pair in it has reached max_dist (same output, less work).           */
int hd_all_pairs_dev(const uint32_t *d_planes, uint64_t n,
                     uint64_t bytes_per_seq, int semantics,
                     uint32_t nibble_mask, int out_bytes, uint32_t max_dist,
                     int early_exit, int variant, void *d_out, void *stream);

/* Rows [row_begin, row_end) of the same triangle into d_out, cell (i, j) at
 * index i*(i-1)/2 + j - row_begin*(row_begin-1)/2 (the band's cells are a
 * contiguous slice of the linear triangle, so bands concatenate in row
 * order).  row_begin and row_end must be multiples of the tile variant's rows
 * (64 x rows_per_lane of hd_variant_info; 512 suits every variant), or
 * row_end = n.  The unit of one GPU's share when the triangle is split across
 * GPUs by row bands; no counterpart in the reference.                       */
int hd_all_pairs_dev_rows(const uint32_t *d_planes, uint64_t n,
                          uint64_t bytes_per_seq, int semantics,
                          uint32_t nibble_mask, int out_bytes, uint32_t max_dist,
                          int early_exit, int variant, uint64_t row_begin,
                          uint64_t row_end, void *d_out, void *stream);

/* Reference distances of a device query against device rows (see above).   */
int hd_reference_distances_dev(const uint8_t *d_query, const uint8_t *d_db,
                               uint64_t n_db, uint64_t bytes_per_seq,
                               const uint64_t *d_rows, uint64_t m,
                               int semantics, uint32_t *d_out, void *stream);

/* ---- batched all-pairs (no counterpart in the reference) ---------------- */
/* Many independent datasets of uneven n and L computed by one persistent
 * triangular-tile launch (captured with the batched encode in a hipGraph);
 * the per-dataset result equals hd_all_pairs on that dataset.              */
typedef struct hd_batch hd_batch;
int hd_batch_create(hd_batch **out);
int hd_batch_destroy(hd_batch *batch);
/* Uploads one dataset (n packed rows); *job receives its index.  out_bytes
 * is 1 or 2; max_dist must fit the cell type.                               */
int hd_batch_add(hd_batch *batch, const uint8_t *packed, uint64_t n,
                 uint64_t bytes_per_seq, int semantics, int out_bytes,
                 uint32_t max_dist, uint64_t *job);
/* Builds the tile list with batch variant `variant` (-1 = default; see
 * hd_batch_variant_info) and captures the graph.                            */
int hd_batch_prepare(hd_batch *batch, int variant, int early_exit);
/* Replays the captured graph on `stream` (asynchronous).                    */
int hd_batch_run(hd_batch *batch, void *stream);
/* The same work as plain launches, dataset after dataset (the baseline).    */
int hd_batch_run_sequential(hd_batch *batch, void *stream);
/* Copies one job's triangle to host memory (synchronous).                   */
int hd_batch_output(hd_batch *batch, uint64_t job, void *host_out);
/* Device address and cell count of one job's triangle.                      */
int hd_batch_output_dev(hd_batch *batch, uint64_t job, void **d_cells,
                        uint64_t *cells);
int hd_batch_info(hd_batch *batch, uint64_t *jobs, uint64_t *tiles,
                  uint32_t *blocks, int *variant);
/* A batch variant is the tile shape every job of the launch uses, 64 R rows
 * by W C columns (R rows per lane, C columns per wave, W waves), and the
 * schedule: variants below hd_batch_variant_count()/2 run persistent
 * workgroups on an atomic tile queue, the others one workgroup per tile.    */
int hd_batch_variant_count(void);
int hd_batch_variant_info(int variant, int *rows_per_lane, int *cols_per_wave,
                          int *waves);

/* ---- serving (no counterpart in the reference) --------------------------- */
/* A packed database resident in device memory serves reference-distance
 * queries against row subsets and small all-pairs jobs on its rows.  A
 * worker thread admits requests in arrival order as soon as their device
 * footprint fits the device budget (an arena the server owns, so no device
 * allocation happens per request) and launches everything admitted so far
 * as one batch on a free stream; up to `streams` batches run at once, so
 * requests that arrive while others run overlap with them.  Results are
 * kept in a pinned host arena until hd_server_release.  Each result equals
 * the single-request entry point's (hd_reference_distances /
 * hd_all_pairs) on the same rows.                                           */
typedef struct hd_server hd_server;
#define HD_REQ_PENDING 0
#define HD_REQ_ADMITTED 1
#define HD_REQ_RUNNING 2
#define HD_REQ_DONE 3
#define HD_REQ_FAILED 4
/* db: n_db packed rows (host), copied to the device.  device_bytes and
 * host_bytes bound the memory of requests in flight; max_batch bounds the
 * requests of one launch (0 = 64).                                          */
int hd_server_create(const uint8_t *db, uint64_t n_db, uint64_t bytes_per_seq,
                     uint64_t device_bytes, uint64_t host_bytes, int streams,
                     uint64_t max_batch, hd_server **out);
int hd_server_destroy(hd_server *server);
/* Distances (uint32) of query (bytes_per_seq bytes) to rows[0..m), in that
 * order; semantics HD_SEM_DENSE or HD_SEM_CHAR_X.                          */
int hd_server_submit_query(hd_server *server, const uint8_t *query,
                           const uint64_t *rows, uint64_t m, int semantics,
                           uint64_t *request);
/* Lower triangle of the n database rows rows[0..n) (in that order).        */
int hd_server_submit_all_pairs(hd_server *server, const uint64_t *rows,
                               uint64_t n, int semantics, int out_bytes,
                               uint32_t max_dist, uint64_t *request);
/* Blocks until the request is done or failed, or timeout_s passes
 * (timeout_s < 0: no limit); *state receives HD_REQ_*.                     */
int hd_server_wait(hd_server *server, uint64_t request, double timeout_s,
                   int *state);
/* Copies a done request's result; *latency_s = completion - submission.
 * Calls for one request must not overlap hd_server_release of it.         */
int hd_server_result(hd_server *server, uint64_t request, void *out,
                     uint64_t bytes, double *latency_s);
int hd_server_release(hd_server *server, uint64_t request);
int hd_server_stats(hd_server *server, uint64_t *done, uint64_t *batches,
                    uint64_t *overlapped_batches, uint64_t *max_running,
                    uint64_t *device_used);
/* Query groups the server ran as shared scans (one read of a database row for
 * every query of the group that holds it), and the queries in them.         */
int hd_server_shared_stats(hd_server *server, uint64_t *groups, uint64_t *queries);
/* Query rectangles: enabled = 1 when the server keeps its database also as
 * dense 3-op planes (a database of only 0, A, C, G, T and '-' codes, with
 * room in device memory); a rectangle computes a group of up to 64 DENSE
 * queries of those codes against every database row in one all-pairs tile
 * launch and then gathers each query's rows.  groups, queries: the groups
 * run that way and the queries in them.                                     */
int hd_server_rect_stats(hd_server *server, int *enabled, uint64_t *groups,
                         uint64_t *queries);

/* Number of all-pairs tile shapes compiled in, and a description of one.    */
int hd_variant_count(void);
int hd_variant_info(int variant, int *rows_per_lane, int *cols_per_wave,
                    int *waves, int *planes);
/* The tile shape variant = -1 selects for this dataset, cap and early-exit
 * flag (hd_all_pairs, hd_all_pairs_dev and the other all-pairs calls).     */
int hd_default_variant(uint64_t n, uint64_t bytes_per_seq, int semantics,
                       uint32_t nibble_mask, int early_exit, uint32_t max_dist,
                       int *variant);

#ifdef __cplusplus
}
#endif

#endif /* HAMMINGDIST_AMD_H */
