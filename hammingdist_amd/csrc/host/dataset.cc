// distances<T>, DataSet<T> and the top-level functions of the host layer.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <functional>
#include <iostream>
#include <memory>
#include <sstream>
#include <stdexcept>

#include "hamming.hh"

namespace hdhost {

namespace {

class Timer {
public:
  // "# hammingdist :: ...<event> completed in <ms> ms." lines, as
  // include/hamming/hamming_impl.hh:74-87 prints them.
  void lap(const std::string &event, bool final = false) {
    const auto now = std::chrono::high_resolution_clock::now();
    std::cout << "# hammingdist :: ..." << event << " completed in "
              << std::chrono::duration_cast<std::chrono::milliseconds>(now - start_).count()
              << " ms." << (final ? "" : "..") << std::endl;
    start_ = now;
  }

private:
  std::chrono::high_resolution_clock::time_point start_ =
      std::chrono::high_resolution_clock::now();
};

std::uint64_t common_length(const std::vector<std::string> &data) {
  return data.empty() ? 0 : data[0].size();
}

} // namespace

Route choose_route(const Packer &pack, std::uint64_t n, std::uint64_t L, bool include_x,
                   bool use_gpu, PackedRows &packed, std::uint64_t bps) {
  // include/hamming/hamming_impl.hh:99-102
  if (include_x && use_gpu)
    throw std::runtime_error(
        "use_gpu=True cannot be used if include_x=True, please set use_gpu=False");
  Route r;
  if (use_gpu) {
    // The reference GPU route is always the dense nibble kernel
    // (hamming_impl.hh:147-159, src/distance_cuda.cu:57-65).
    packed = pack(false, bps);
    return r;
  }
  if (include_x) {
    // hamming_impl.hh:104: include_x always takes the sparse route.
    r.semantics = HD_SEM_SPARSE_X;
    r.pack_include_x = true;
    r.sparse = true;
    packed = pack(true, bps);
    return r;
  }
  // hamming_impl.hh:108-118: sparse when fewer than 0.5% of all entries
  // differ from the per-site consensus letter.
  packed = pack(false, bps);
  std::uint64_t n_diff = 0;
  std::uint32_t mask = 0;
  check(hd_consensus_diff(packed.data(), n, bps, L, &n_diff, &mask));
  constexpr double sparse_threshold{0.005};
  const double frac_diff = static_cast<double>(n_diff) / static_cast<double>(n * L);
  if (frac_diff < sparse_threshold) {
    r.semantics = HD_SEM_SPARSE;
    r.sparse = true;
  }
  return r;
}

template <typename T>
Cells<T> distances_packed(const Packer &pack, std::uint64_t n, std::uint64_t L,
                                bool include_x, bool use_gpu, int max_distance,
                                const std::function<void()> &after_pack) {
  Cells<T> result(n < 2 ? 0 : n * (n - 1) / 2); // every cell is written below
  const T max_dist = safe_int_cast<T>(max_distance);
  Timer timer;
  const std::uint64_t bps = bytes_per_seq(L);
  PackedRows packed;
  const Route route = choose_route(pack, n, L, include_x, use_gpu, packed, bps);
  std::cout << "# hammingdist :: Using GPU with "
            << (route.sparse ? "sparse" : "dense") << " distance semantics..."
            << std::endl;
  if (after_pack) after_pack();
  timer.lap("pre-processing");
  if (n >= 2)
    check(hd_all_pairs(packed.data(), n, bps, route.semantics, sizeof(T), max_dist,
                       result.data()));
  timer.lap("distance calculation", true);
  return result;
}

template <typename T>
Cells<T> distances(std::vector<std::string> &data, bool include_x,
                         bool clear_input_data, bool use_gpu, int max_distance) {
  const Packer pack = [&](bool x, std::uint64_t bps) { return pack_all(data, x, bps); };
  return distances_packed<T>(pack, data.size(), common_length(data), include_x, use_gpu,
                             max_distance, [&] {
                               if (clear_input_data) data.clear();
                             });
}

template Cells<std::uint8_t> distances<std::uint8_t>(std::vector<std::string> &, bool, bool,
                                                     bool, int);
template Cells<std::uint16_t> distances<std::uint16_t>(std::vector<std::string> &, bool,
                                                       bool, bool, int);

// ---- DataSet ---------------------------------------------------------------

template <typename T>
DataSet<T>::DataSet(std::vector<std::string> &data, bool include_x, bool clear_input_data,
                    std::vector<std::size_t> &&indices, bool use_gpu, int max_distance)
    : nsamples(data.size()), sequence_indices(std::move(indices)) {
  validate_data(data);
  result = distances<T>(data, include_x, clear_input_data, use_gpu, max_distance);
}

// include/hamming/hamming.hh:57-61
template <typename T> DataSet<T>::DataSet(Cells<T> &&distances) : result{std::move(distances)} {
  const auto x = 8 * result.size() + 1;
  const auto r = static_cast<std::size_t>(std::round(std::sqrt(static_cast<double>(x))));
  nsamples = (r + 1) / 2;
}

template <typename T> int DataSet<T>::get(std::size_t i, std::size_t j) const {
  // include/hamming/hamming.hh:129-138, with a bounds check.
  if (i >= nsamples || j >= nsamples) throw std::out_of_range("index out of range");
  if (i < j) return result[j * (j - 1) / 2 + i];
  if (i > j) return result[i * (i - 1) / 2 + j];
  return 0;
}

// include/hamming/hamming.hh:63-73
template <typename T> void DataSet<T>::dump(const std::string &filename) const {
  std::ofstream stream(filename);
  std::string line;
  for (std::size_t i = 0; i < nsamples; ++i) {
    line.clear();
    for (std::size_t j = 0; j < nsamples; ++j) {
      line += std::to_string(get(i, j));
      if (j != nsamples - 1) line += ", ";
    }
    line += '\n';
    stream << line;
  }
}

template <typename T>
void DataSet<T>::dump_lower_triangular(const std::string &filename) const {
  write_lower_triangular(filename, result);
}

// include/hamming/hamming.hh:116-127
template <typename T>
void DataSet<T>::dump_sequence_indices(const std::string &filename) const {
  std::ofstream stream(filename);
  std::string out;
  if (sequence_indices.empty()) {
    for (std::size_t i = 0; i < nsamples; ++i) out += std::to_string(i) + "\n";
  } else {
    for (auto s : sequence_indices) out += std::to_string(s) + "\n";
  }
  stream << out;
}

template struct DataSet<std::uint8_t>;
template struct DataSet<std::uint16_t>;

// ---- top-level functions (src/hamming.cc) -----------------------------------

DataSet<DefaultDistIntType> from_stringlist(std::vector<std::string> &data, bool include_x,
                                            bool use_gpu, int max_distance) {
  return DataSet<DefaultDistIntType>(data, include_x, false, {}, use_gpu, max_distance);
}

namespace {

// Rows of a FASTA file as read_fasta gives them (with dedup when asked),
// checked as validate_data checks them (src/hamming_impl.cc:80-91).
struct FastaRows {
  std::vector<std::size_t> rows;    // record of each row
  std::vector<std::size_t> indices; // dedup: row of each record
  std::uint64_t L = 0;
};

FastaRows fasta_rows(const FastaRecords &recs, bool remove_duplicates) {
  FastaRows r;
  r.rows.resize(recs.size());
  for (std::size_t k = 0; k < r.rows.size(); ++k) r.rows[k] = k;
  if (remove_duplicates) r.indices = recs.dedup(r.rows);
  return r;
}

void validate_rows(const FastaRecords &recs, FastaRows &r) {
  if (r.rows.empty() || recs.length(r.rows[0]) == 0)
    throw std::runtime_error("Error: Empty sequence");
  r.L = recs.length(r.rows[0]);
  for (std::size_t k : r.rows)
    if (recs.length(k) != r.L)
      throw std::runtime_error("Error: Sequences do not all have the same length");
}

} // namespace

// include/hamming/hamming.hh:168-178.  The records are packed straight from
// the mapped file; no per-sequence strings are made.
template <typename T>
DataSet<T> from_fasta(const std::string &filename, bool include_x, bool remove_duplicates,
                      std::size_t n, bool use_gpu, int max_distance) {
  FastaRecords recs(filename, n);
  FastaRows fr = fasta_rows(recs, remove_duplicates);
  validate_rows(recs, fr);
  const Packer pack = [&](bool x, std::uint64_t bps) {
    PackedRows out(fr.rows.size() * bps);
    recs.pack_rows(fr.rows, pack_table(x), out.data(), bps);
    return out;
  };
  DataSet<T> ds(distances_packed<T>(pack, fr.rows.size(), fr.L, include_x, use_gpu,
                                    max_distance, {}));
  ds.nsamples = fr.rows.size();
  ds.sequence_indices = std::move(fr.indices);
  return ds;
}
template DataSet<std::uint8_t> from_fasta<std::uint8_t>(const std::string &, bool, bool,
                                                        std::size_t, bool, int);
template DataSet<std::uint16_t> from_fasta<std::uint16_t>(const std::string &, bool, bool,
                                                          std::size_t, bool, int);

namespace {

struct StreamWriter {
  std::string filename;
  int out_bytes;
  std::uint32_t max_value;
  std::unique_ptr<RowWriter> writer;
};

int write_band(void *user, const void *cells, std::uint64_t, std::uint64_t,
               std::uint64_t row_begin, std::uint64_t row_end) {
  auto *w = static_cast<StreamWriter *>(user);
  try {
    if (!w->writer) w->writer = std::make_unique<RowWriter>(w->filename, false);
    if (w->out_bytes == 1)
      w->writer->write(static_cast<const std::uint8_t *>(cells), row_begin, row_end,
                       w->max_value);
    else
      w->writer->write(static_cast<const std::uint16_t *>(cells), row_begin, row_end,
                       w->max_value);
    return 0;
  } catch (const std::exception &e) {
    std::cerr << "# hammingdist :: write failed: " << e.what() << std::endl;
    return 1;
  }
}

} // namespace

// src/hamming.cc:27-52 with the GPU streaming loop of
// src/distance_cuda.cu:140-191, 216-223: the triangle goes to the file in
// row bands while the device computes the next bands.
void from_fasta_to_lower_triangular(const std::string &input_filename,
                                    const std::string &output_filename,
                                    bool remove_duplicates, std::size_t n, bool use_gpu,
                                    int max_distance) {
  if (!use_gpu)
    throw std::runtime_error(
        "from_fasta_to_lower_triangular is currently only available on GPU");
  std::cout << "# hammingdist :: Using GPU..." << std::endl;
  Timer timer;
  PackedRows packed;
  std::size_t count = 0;
  std::uint64_t bps = 0;
  {
    FastaRecords recs(input_filename, n);
    FastaRows fr = fasta_rows(recs, remove_duplicates);
    count = fr.rows.size();
    if (count < 2) return; // no cells: the reference writes nothing
    validate_rows(recs, fr);
    bps = bytes_per_seq(fr.L);
    packed.resize(count * bps);
    recs.pack_rows(fr.rows, pack_table(false), packed.data(), bps);
  }
  timer.lap("pre-processing");
  const std::uint16_t max_dist = max_distance > 65535 ? std::uint16_t{65535}
                                                      : static_cast<std::uint16_t>(max_distance);
  // The text is the same for any cell width that holds max_dist.
  StreamWriter w{output_filename, max_dist <= 255 ? 1 : 2, max_dist, nullptr};
  const int st = hd_all_pairs_stream(packed.data(), count, bps, HD_SEM_DENSE, w.out_bytes,
                                     max_dist, 0, write_band, &w);
  check(st);
  timer.lap("distance calculation", true);
}

// distance(), src/hamming.cc:54-74
ReferenceDistIntType distance(const std::string &seq0, const std::string &seq1,
                              bool include_x) {
  if (seq0.size() != seq1.size())
    throw std::runtime_error("Error: Sequences do not all have the same length");
  if (seq0.empty()) return 0;
  const std::uint64_t bps = bytes_per_seq(seq0.size());
  const auto lut = lookup_table(include_x);
  std::vector<GeneBlock> a(bps), b(bps);
  pack_row(seq0.data(), seq0.size(), lut, a.data(), bps);
  pack_row(seq1.data(), seq1.size(), lut, b.data(), bps);
  std::int32_t d = 0;
  check(hd_distance_pair(a.data(), b.data(), bps,
                         include_x ? HD_SEM_CHAR_X : HD_SEM_DENSE,
                         std::numeric_limits<std::int32_t>::max(), &d));
  return static_cast<ReferenceDistIntType>(d);
}

// fasta_reference_distances, src/hamming.cc:104-141.  Every record is
// compared over its own length against the reference prefix; records are
// packed to the reference's row width, and sites past a record's end (packed
// as '-') are corrected for: '-' against an invalid reference site would
// count 1, while the reference does not visit those sites.
std::vector<ReferenceDistIntType> fasta_reference_distances(
    const std::string &reference_sequence, const std::string &fasta_file, bool include_x) {
  // All records, no limit, no dedup: the reference's loop at
  // src/hamming.cc:115-139 reads the file the way read_fasta does.
  FastaRecords recs(fasta_file, 0);
  const std::uint64_t Lref = reference_sequence.size();
  const std::size_t m = recs.size();
  std::vector<ReferenceDistIntType> out(m, 0);
  for (std::size_t k = 0; k < m; ++k)
    if (recs.length(k) > Lref)
      throw std::runtime_error("Error: sequence " + std::to_string(k) +
                               " is longer than the reference sequence");
  if (m == 0 || Lref == 0) return out;
  const auto lut = lookup_table(include_x);
  const std::uint64_t bps = bytes_per_seq(Lref);
  std::vector<GeneBlock> query(bps);
  pack_row(reference_sequence.data(), Lref, lut, query.data(), bps);
  // invalid_after[s] = invalid reference sites in [s, Lref)
  std::vector<std::uint32_t> invalid_after(Lref + 1, 0);
  for (std::uint64_t s = Lref; s-- > 0;)
    invalid_after[s] = invalid_after[s + 1] +
                       (lut[static_cast<unsigned char>(reference_sequence[s])] == 0 ? 1u : 0u);
  std::vector<std::size_t> all(m);
  for (std::size_t k = 0; k < m; ++k) all[k] = k;
  PackedRows db(m * bps);
  recs.pack_rows(all, pack_table(include_x), db.data(), bps);
  check(hd_reference_distances(query.data(), db.data(), m, bps, nullptr, m,
                               include_x ? HD_SEM_CHAR_X : HD_SEM_DENSE, out.data()));
  for (std::size_t k = 0; k < m; ++k) out[k] -= invalid_after[recs.length(k)];
  return out;
}

} // namespace hdhost
