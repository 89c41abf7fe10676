// Host layer of hammingdist_amd: the reference's C++ library interface
// (include/hamming/hamming.hh, hamming_impl.hh, hamming_utils.hh) on top of the
// HIP engine's C ABI (include/hammingdist_amd.h).  Parsing, packing, result
// containers and text output live here; every distance is computed by the
// engine on the GPU.
#pragma once

#include <array>
#include <cstdint>
#include <functional>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "hammingdist_amd.h"

namespace hdhost {

// A roctx range over a scope: the host layer's phases (FASTA ingest, packing,
// text formatting and writing, loading) show up by name in
// `rocprofv3 --marker-trace`; without a profiler a range costs two calls into
// the roctx stub.
class TraceRange {
public:
  explicit TraceRange(const char *name) { roctxRangePushA(name); }
  ~TraceRange() { roctxRangePop(); }
  TraceRange(const TraceRange &) = delete;
  TraceRange &operator=(const TraceRange &) = delete;
};

using GeneBlock = std::uint8_t;
using ReferenceDistIntType = std::uint32_t;
using DefaultDistIntType = std::uint8_t;

// Asks for transparent huge pages on [p, p + bytes) (page-aligned inside):
// first writes into a fresh multi-GB buffer fault 2 MiB pages instead of
// 4 KiB ones (3x the copy rate into a new 10 GB triangle where the system
// grants huge pages on request only).
void advise_huge_pages(void *p, std::size_t bytes);

// Allocator that leaves new elements uninitialised: packed rows are always
// written in full before they are read, so a zero fill would be wasted.
// Allocations of 64 MiB or more ask for huge pages.
template <typename T> struct NoInitAllocator : std::allocator<T> {
  template <typename U> struct rebind {
    using other = NoInitAllocator<U>;
  };
  NoInitAllocator() = default;
  template <typename U> NoInitAllocator(const NoInitAllocator<U> &) noexcept {}
  T *allocate(std::size_t n) {
    T *p = std::allocator<T>::allocate(n);
    if (n * sizeof(T) >= (std::size_t{64} << 20)) advise_huge_pages(p, n * sizeof(T));
    return p;
  }
  template <typename U> void construct(U *) noexcept {}
  template <typename U, typename... A> void construct(U *p, A &&...a) {
    ::new (static_cast<void *>(p)) U(std::forward<A>(a)...);
  }
};
using PackedRows = std::vector<GeneBlock, NoInitAllocator<GeneBlock>>;
// A lower triangle of distances; its cells are written by the engine in full.
template <typename T> using Cells = std::vector<T, NoInitAllocator<T>>;

// ---- encoding (src/hamming_impl.cc:23-44, 244-267) ------------------------
std::array<GeneBlock, 256> lookup_table(bool include_x = false);
// lookup_table for two characters at once: entry (c1 << 8) | c0 is the
// packed byte of sites c0 (low nibble) and c1 (high nibble).
struct PackTable {
  std::array<GeneBlock, 256> lut;
  std::vector<GeneBlock> pair;
};
const PackTable &pack_table(bool include_x);
// Bytes of one packed row of L sites: two sites per byte, padded to 8 bytes.
std::uint64_t bytes_per_seq(std::uint64_t L);
// Packs one sequence (from_string) into bps bytes; bytes past the sequence
// are 0xFF ('-' in both nibbles).
void pack_row(const char *s, std::uint64_t L, const std::array<GeneBlock, 256> &lut,
              GeneBlock *out, std::uint64_t bps);
void pack_row(const char *s, std::uint64_t L, const PackTable &t, GeneBlock *out,
              std::uint64_t bps);
// Packs every sequence into n * bps contiguous bytes (parallel over rows).
PackedRows pack_all(const std::vector<std::string> &data, bool include_x, std::uint64_t bps);

// ---- FASTA ingest (src/hamming_impl.cc:80-91, 200-242; src/hamming.cc:76-102)
struct FastaData {
  std::vector<std::string> data;
  std::vector<std::size_t> sequence_indices;
};
FastaData read_fasta(const std::string &filename, bool remove_duplicates = false,
                     std::size_t n = 0);

// The records of a FASTA file (read_fasta's rules) without copying them:
// the file is mapped, record ranges, lengths and dedup hashes are found in
// parallel, and records are packed straight from the mapping.
class FastaRecords {
public:
  struct Span {
    std::uint64_t begin, end; // bytes of the record's lines in the file
  };
  FastaRecords(const std::string &filename, std::size_t n);
  ~FastaRecords();
  FastaRecords(const FastaRecords &) = delete;
  FastaRecords &operator=(const FastaRecords &) = delete;

  std::size_t size() const { return spans_.size(); }
  // Sequence length of record k (its characters without line ends).
  std::uint64_t length(std::size_t k) const { return lengths_[k]; }
  std::string text(std::size_t k) const;
  // Row k packed as from_string packs it, into bps bytes.
  void pack(std::size_t k, const PackTable &t, GeneBlock *out, std::uint64_t bps,
            std::string &scratch) const;
  // records[r] packed into out + r * bps, in parallel.
  void pack_rows(const std::vector<std::size_t> &records, const PackTable &t, GeneBlock *out,
                 std::uint64_t bps) const;
  // Per record, the row of its first identical record among the unique
  // ones; `unique` receives the record index of every unique row.
  std::vector<std::size_t> dedup(std::vector<std::size_t> &unique) const;

private:
  std::uint64_t hash(std::size_t k) const;
  bool same_text(std::size_t a, std::size_t b) const;

  int fd_ = -1;
  void *map_ = nullptr;
  std::string buffer_;
  const char *data_ = nullptr;
  std::uint64_t size_ = 0;
  std::vector<Span> spans_;
  std::vector<std::uint64_t> lengths_;
};
std::vector<std::size_t> fasta_sequence_indices(const std::string &fasta_file,
                                                std::size_t n = 0);
void validate_data(const std::vector<std::string> &data);

// ---- engine bridge --------------------------------------------------------
bool gpu_available();
// Throws std::runtime_error with the engine's message when status != HD_OK.
void check(int status);

// safe_int_cast<T>(x), include/hamming/hamming_impl.hh:37-45
template <typename T> T safe_int_cast(int x, T max_x = std::numeric_limits<T>::max()) {
  if (x > static_cast<int>(max_x)) return max_x;
  return static_cast<T>(x);
}

// Route of distances<T>: which per-site semantics the reference would use
// for these inputs (include/hamming/hamming_impl.hh:95-118, 147-159) and the
// matching packing.  Throws the reference's errors for invalid flag mixes.
struct Route {
  int semantics = HD_SEM_DENSE;
  bool pack_include_x = false;
  bool sparse = false;
};
// Packs the dataset's rows with the include_x table or not, bps bytes each.
using Packer = std::function<PackedRows(bool include_x, std::uint64_t bps)>;
Route choose_route(const Packer &pack, std::uint64_t n, std::uint64_t L, bool include_x,
                   bool use_gpu, PackedRows &packed, std::uint64_t bps);

// distances<T> on rows a Packer provides (n rows of L sites); after_pack
// runs once the rows are packed.
template <typename T>
Cells<T> distances_packed(const Packer &pack, std::uint64_t n, std::uint64_t L,
                                bool include_x, bool use_gpu, int max_distance,
                                const std::function<void()> &after_pack);

// distances<T>, include/hamming/hamming_impl.hh:68-176
template <typename T>
Cells<T> distances(std::vector<std::string> &data, bool include_x,
                         bool clear_input_data, bool use_gpu, int max_distance);

// ---- result container (include/hamming/hamming.hh:22-143) -----------------
template <typename T> struct DataSet {
  explicit DataSet(std::vector<std::string> &data, bool include_x = false,
                   bool clear_input_data = false, std::vector<std::size_t> &&indices = {},
                   bool use_gpu = false,
                   int max_distance = std::numeric_limits<int>::max());
  // Full-matrix CSV reader (from_csv), include/hamming/hamming.hh:34-55.
  explicit DataSet(const std::string &filename);
  explicit DataSet(Cells<T> &&distances);

  void dump(const std::string &filename) const;
  void dump_lower_triangular(const std::string &filename) const;
  void dump_sparse(const std::string &filename, int threshold) const;
  void dump_sequence_indices(const std::string &filename) const;
  int get(std::size_t i, std::size_t j) const;

  std::size_t nsamples = 0;
  Cells<T> result;
  std::vector<std::size_t> sequence_indices{};
};

DataSet<DefaultDistIntType> from_stringlist(std::vector<std::string> &data, bool include_x,
                                            bool use_gpu, int max_distance);
DataSet<DefaultDistIntType> from_csv(const std::string &filename);
template <typename T> DataSet<T> from_lower_triangular(const std::string &filename);
template <typename T>
DataSet<T> from_fasta(const std::string &filename, bool include_x, bool remove_duplicates,
                      std::size_t n, bool use_gpu, int max_distance);
void from_fasta_to_lower_triangular(const std::string &input_filename,
                                    const std::string &output_filename,
                                    bool remove_duplicates, std::size_t n, bool use_gpu,
                                    int max_distance);
ReferenceDistIntType distance(const std::string &seq0, const std::string &seq1,
                              bool include_x = false);
std::vector<ReferenceDistIntType> fasta_reference_distances(
    const std::string &reference_sequence, const std::string &fasta_file,
    bool include_x = false);

// ---- text output (include/hamming/hamming_utils.hh:14-110) ----------------
// Writes whole rows of a lower triangle to one file, appending row bands in
// order: each band is formatted in parallel chunks, and the chunks are
// written in parallel at their final offsets.
class RowWriter {
public:
  RowWriter(const std::string &filename, bool append);
  ~RowWriter();
  RowWriter(const RowWriter &) = delete;
  RowWriter &operator=(const RowWriter &) = delete;
  // Rows [row_begin, row_end); `cells` starts at (row_begin, 0); every value
  // is <= max_value (values below 10 take a one-digit path).
  template <typename T>
  void write(const T *cells, std::uint64_t row_begin, std::uint64_t row_end,
             std::uint32_t max_value);
  void write_raw(const char *data, std::size_t len);

private:
  // Maps [offset_, offset_ + len) of the file for writing after reserving it
  // with fallocate; null when the file cannot be written that way (not a
  // regular file, no read access, no fallocate), and then pwrite is used.
  char *map_range(std::uint64_t len, char **base, std::uint64_t *base_len);

  std::string filename_;
  int fd_ = -1;
  bool mappable_ = false;
  std::uint64_t map_bytes_ = 1ull << 30; // writes this large go through a mapping
  std::uint64_t offset_ = 0;
  std::unique_ptr<char[]> buf_; // one-digit text of the last band, reused
  std::uint64_t buf_cap_ = 0;
};

// Writes whole rows [row_begin, row_end) of a triangle whose first cell in
// `cells` is (row_begin, 0); truncates the file when `append` is false.
template <typename T>
void write_rows(const std::string &filename, const T *cells, std::uint64_t row_begin,
                std::uint64_t row_end, bool append);
template <typename T>
void write_lower_triangular(const std::string &filename, const Cells<T> &cells);

// ---- small parallel helpers -----------------------------------------------
unsigned worker_count();
template <typename F>
void parallel_for(std::uint64_t n, F &&f, std::uint64_t min_per_worker = 1);

} // namespace hdhost

#include "parallel.hh"
