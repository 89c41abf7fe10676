// FASTA ingest, packing and the engine bridge of the host layer.
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <stdexcept>
#include <string_view>
#include <unordered_map>

#include "hamming.hh"

namespace hdhost {

unsigned worker_count() {
  static const unsigned n = [] {
    unsigned hw = std::thread::hardware_concurrency();
    if (const char *e = std::getenv("HD_HOST_THREADS")) {
      const int v = std::atoi(e);
      if (v > 0) return (unsigned)v;
    }
    return std::max(1u, std::min(hw ? hw : 1u, 32u));
  }();
  return n;
}

// lookupTable, src/hamming_impl.cc:23-44: '-' 1111, A 0001, C 0010, G 0100,
// T 1000 in both nibbles, X 0011 only with include_x, anything else 0.
std::array<GeneBlock, 256> lookup_table(bool include_x) {
  std::array<GeneBlock, 256> t{};
  t[static_cast<unsigned char>('-')] = 0xFF;
  t[static_cast<unsigned char>('A')] = 0x11;
  t[static_cast<unsigned char>('C')] = 0x22;
  t[static_cast<unsigned char>('G')] = 0x44;
  t[static_cast<unsigned char>('T')] = 0x88;
  if (include_x) t[static_cast<unsigned char>('X')] = 0x33;
  return t;
}

std::uint64_t bytes_per_seq(std::uint64_t L) { return ((L + 1) / 2 + 7) / 8 * 8; }

namespace {
PackTable make_pack_table(bool include_x) {
  PackTable t;
  t.lut = lookup_table(include_x);
  t.pair.resize(65536);
  for (unsigned c1 = 0; c1 < 256; ++c1)
    for (unsigned c0 = 0; c0 < 256; ++c0)
      t.pair[(c1 << 8) | c0] = static_cast<GeneBlock>((t.lut[c0] & 0x0F) | (t.lut[c1] & 0xF0));
  return t;
}
} // namespace

const PackTable &pack_table(bool include_x) {
  static const PackTable tables[2] = {make_pack_table(false), make_pack_table(true)};
  return tables[include_x ? 1 : 0];
}

// from_string, src/hamming_impl.cc:244-267: site 2k low nibble, 2k+1 high
// nibble; an odd last site gets '-' in the high nibble; 0xFF pads to bps.
void pack_row(const char *s, std::uint64_t L, const std::array<GeneBlock, 256> &lut,
              GeneBlock *out, std::uint64_t bps) {
  const std::uint64_t full = L / 2;
  const unsigned char *u = reinterpret_cast<const unsigned char *>(s);
  for (std::uint64_t b = 0; b < full; ++b)
    out[b] = static_cast<GeneBlock>((lut[u[2 * b]] & 0x0F) | (lut[u[2 * b + 1]] & 0xF0));
  std::uint64_t k = full;
  if (L & 1) out[k++] = static_cast<GeneBlock>((lut[u[L - 1]] & 0x0F) | 0xF0);
  if (k < bps) std::memset(out + k, 0xFF, bps - k);
}

// The same bytes as the per-character form, two sites per table lookup.
void pack_row(const char *s, std::uint64_t L, const PackTable &t, GeneBlock *out,
              std::uint64_t bps) {
  const std::uint64_t full = L / 2;
  const GeneBlock *pair = t.pair.data();
  for (std::uint64_t b = 0; b < full; ++b) {
    std::uint16_t v;
    std::memcpy(&v, s + 2 * b, 2); // little-endian: site 2b in the low byte
    out[b] = pair[v];
  }
  std::uint64_t k = full;
  if (L & 1)
    out[k++] = static_cast<GeneBlock>((t.lut[static_cast<unsigned char>(s[L - 1])] & 0x0F) |
                                      0xF0);
  if (k < bps) std::memset(out + k, 0xFF, bps - k);
}

PackedRows pack_all(const std::vector<std::string> &data, bool include_x, std::uint64_t bps) {
  PackedRows packed(data.size() * bps);
  const PackTable &t = pack_table(include_x);
  parallel_for(
      data.size(),
      [&](std::uint64_t b, std::uint64_t e) {
        for (std::uint64_t i = b; i < e; ++i)
          pack_row(data[i].data(), data[i].size(), t, packed.data() + i * bps, bps);
      },
      items_per_worker(2 * bps));
  return packed;
}

// validate_data, src/hamming_impl.cc:80-91
void validate_data(const std::vector<std::string> &data) {
  if (data.empty() || data[0].empty()) throw std::runtime_error("Error: Empty sequence");
  const auto length = data[0].size();
  for (const auto &d : data)
    if (d.size() != length)
      throw std::runtime_error("Error: Sequences do not all have the same length");
}

namespace {

// Record ranges of a FASTA buffer, as the reference's reader
// (src/hamming_impl.cc:200-242, a std::getline loop) delimits them: the
// first line is skipped; a record is every line up to the next line that
// starts with '>' (that line is consumed); the loop stops at end of file.
// In getline terms the last segment is a record unless the file ends with
// a '>' line that has no '\n' (reading that line already sets eof), and a
// file that ends with a '>' line and '\n' gives one more, empty, record.
std::vector<FastaRecords::Span> record_spans(const char *b, std::uint64_t N, std::size_t n) {
  std::vector<FastaRecords::Span> spans;
  const void *first_nl = N ? std::memchr(b, '\n', N) : nullptr;
  if (!first_nl) return spans; // empty, or one line without '\n'
  const std::uint64_t body = static_cast<const char *>(first_nl) - b + 1;
  // Header starts: positions p >= body with b[p] == '>' that begin a line.
  const std::uint64_t chunks =
      std::max<std::uint64_t>(1, std::min<std::uint64_t>(worker_count() * 4, (N - body) >> 20));
  std::vector<std::vector<std::uint64_t>> found(chunks);
  parallel_for(chunks, [&](std::uint64_t cb, std::uint64_t ce) {
    for (std::uint64_t c = cb; c < ce; ++c) {
      // newlines at k in [c0, c1) announce a line start at k + 1
      const std::uint64_t c0 = body + (N - body) * c / chunks;
      const std::uint64_t c1 = body + (N - body) * (c + 1) / chunks;
      auto &out = found[c];
      if (c == 0 && body < N && b[body] == '>') out.push_back(body);
      std::uint64_t k = c0;
      while (k < c1) {
        const void *nl = std::memchr(b + k, '\n', c1 - k);
        if (!nl) break;
        k = static_cast<const char *>(nl) - b;
        if (k + 1 < N && b[k + 1] == '>') out.push_back(k + 1);
        ++k;
      }
    }
  });
  if (n == 0) n = std::numeric_limits<std::size_t>::max();
  std::uint64_t start = body;
  for (const auto &vec : found)
    for (std::uint64_t h : vec) {
      if (spans.size() >= n) return spans;
      spans.push_back({start, h});
      const void *nl = std::memchr(b + h, '\n', N - h);
      if (!nl) return spans; // the last line is this header, without '\n'
      start = static_cast<const char *>(nl) - b + 1;
    }
  if (spans.size() < n) spans.push_back({start, N});
  return spans;
}

// Calls f(piece, length) for each line of [p, p + len) without its '\n'.
template <typename F> void for_each_piece(const char *p, std::uint64_t len, F &&f) {
  const char *end = p + len;
  while (p < end) {
    const void *nl = std::memchr(p, '\n', static_cast<std::size_t>(end - p));
    const char *stop = nl ? static_cast<const char *>(nl) : end;
    if (stop > p) f(p, static_cast<std::uint64_t>(stop - p));
    p = nl ? stop + 1 : end;
  }
}

inline std::uint64_t mix64(std::uint64_t h, std::uint64_t w) {
  h ^= w * 0x9E3779B97F4A7C15ull;
  h = (h << 31) | (h >> 33);
  return h * 0xBF58476D1CE4E5B9ull;
}

} // namespace

// ---- FastaRecords ----------------------------------------------------------

FastaRecords::FastaRecords(const std::string &filename, std::size_t n) {
  const TraceRange trace("read FASTA");
  fd_ = ::open(filename.c_str(), O_RDONLY);
  if (fd_ < 0) throw std::runtime_error("Error: Failed to open file '" + filename + "'");
  struct stat st {};
  if (::fstat(fd_, &st) == 0 && S_ISREG(st.st_mode) && st.st_size > 0) {
    void *m = ::mmap(nullptr, static_cast<std::size_t>(st.st_size), PROT_READ, MAP_PRIVATE,
                     fd_, 0);
    if (m != MAP_FAILED) {
      map_ = m;
      size_ = static_cast<std::uint64_t>(st.st_size);
      data_ = static_cast<const char *>(m);
    }
  }
  if (!data_) { // pipes and other streams, empty files
    char chunk[1 << 16];
    ssize_t r;
    while ((r = ::read(fd_, chunk, sizeof chunk)) > 0)
      buffer_.append(chunk, static_cast<std::size_t>(r));
    data_ = buffer_.data();
    size_ = buffer_.size();
  }
  spans_ = record_spans(data_, size_, n);
  lengths_.resize(spans_.size());
  const std::uint64_t per_record = items_per_worker(spans_.empty() ? 1 : size_ / spans_.size());
  parallel_for(
      spans_.size(),
      [&](std::uint64_t b, std::uint64_t e) {
        for (std::uint64_t k = b; k < e; ++k) {
          std::uint64_t len = 0;
          for_each_piece(data_ + spans_[k].begin, spans_[k].end - spans_[k].begin,
                         [&](const char *, std::uint64_t l) { len += l; });
          lengths_[k] = len;
        }
      },
      per_record);
}

FastaRecords::~FastaRecords() {
  if (map_) ::munmap(map_, static_cast<std::size_t>(size_));
  if (fd_ >= 0) ::close(fd_);
}

std::string FastaRecords::text(std::size_t k) const {
  std::string s;
  s.reserve(lengths_[k]);
  for_each_piece(data_ + spans_[k].begin, spans_[k].end - spans_[k].begin,
                 [&](const char *p, std::uint64_t l) { s.append(p, static_cast<std::size_t>(l)); });
  return s;
}

void FastaRecords::pack(std::size_t k, const PackTable &t, GeneBlock *out, std::uint64_t bps,
                        std::string &scratch) const {
  const char *p = data_ + spans_[k].begin;
  const std::uint64_t len = lengths_[k];
  // One line (a trailing '\n' at most): pack in place; else join first.
  const void *nl = std::memchr(p, '\n', static_cast<std::size_t>(spans_[k].end - spans_[k].begin));
  if (!nl || static_cast<const char *>(nl) >= p + len) {
    pack_row(p, len, t, out, bps);
    return;
  }
  scratch.clear();
  for_each_piece(p, spans_[k].end - spans_[k].begin,
                 [&](const char *q, std::uint64_t l) { scratch.append(q, static_cast<std::size_t>(l)); });
  pack_row(scratch.data(), len, t, out, bps);
}

std::uint64_t FastaRecords::hash(std::size_t k) const {
  // 64-bit words of the joined text, whatever its line breaks.
  std::uint64_t h = 0x243F6A8885A308D3ull ^ lengths_[k], acc = 0;
  unsigned fill = 0;
  for_each_piece(data_ + spans_[k].begin, spans_[k].end - spans_[k].begin,
                 [&](const char *p, std::uint64_t l) {
                   const char *end = p + l;
                   while (p < end) {
                     if (fill == 0 && end - p >= 8) {
                       std::uint64_t w;
                       std::memcpy(&w, p, 8);
                       h = mix64(h, w);
                       p += 8;
                     } else {
                       acc |= static_cast<std::uint64_t>(static_cast<unsigned char>(*p++))
                              << (8 * fill);
                       if (++fill == 8) {
                         h = mix64(h, acc);
                         acc = 0;
                         fill = 0;
                       }
                     }
                   }
                 });
  return mix64(h, acc ^ (static_cast<std::uint64_t>(fill) << 56));
}

bool FastaRecords::same_text(std::size_t a, std::size_t b) const {
  if (lengths_[a] != lengths_[b]) return false;
  // Two cursors over the pieces of both records.
  struct Cursor {
    const char *p, *end, *piece;
    std::uint64_t left;
    void next() {
      while (left == 0 && p < end) {
        const void *nl = std::memchr(p, '\n', static_cast<std::size_t>(end - p));
        const char *stop = nl ? static_cast<const char *>(nl) : end;
        piece = p;
        left = static_cast<std::uint64_t>(stop - p);
        p = nl ? stop + 1 : end;
      }
    }
  };
  Cursor x{data_ + spans_[a].begin, data_ + spans_[a].end, nullptr, 0};
  Cursor y{data_ + spans_[b].begin, data_ + spans_[b].end, nullptr, 0};
  std::uint64_t todo = lengths_[a];
  while (todo) {
    x.next();
    y.next();
    const std::uint64_t step = std::min(x.left, y.left);
    if (std::memcmp(x.piece, y.piece, static_cast<std::size_t>(step)) != 0) return false;
    x.piece += step;
    y.piece += step;
    x.left -= step;
    y.left -= step;
    todo -= step;
  }
  return true;
}

std::vector<std::size_t> FastaRecords::dedup(std::vector<std::size_t> &unique) const {
  const TraceRange trace("dedup");
  std::vector<std::uint64_t> hashes(spans_.size());
  parallel_for(
      spans_.size(),
      [&](std::uint64_t b, std::uint64_t e) {
        for (std::uint64_t k = b; k < e; ++k) hashes[k] = hash(k);
      },
      items_per_worker(spans_.empty() ? 1 : size_ / spans_.size()));
  std::unordered_multimap<std::uint64_t, std::size_t> rows; // hash -> unique row
  rows.reserve(spans_.size());
  std::vector<std::size_t> index;
  index.reserve(spans_.size());
  unique.clear();
  for (std::size_t k = 0; k < spans_.size(); ++k) {
    std::size_t row = unique.size();
    auto range = rows.equal_range(hashes[k]);
    for (auto it = range.first; it != range.second; ++it)
      if (same_text(unique[it->second], k)) {
        row = it->second;
        break;
      }
    if (row == unique.size()) {
      rows.emplace(hashes[k], row);
      unique.push_back(k);
    }
    index.push_back(row);
  }
  return index;
}

void FastaRecords::pack_rows(const std::vector<std::size_t> &records, const PackTable &t,
                             GeneBlock *out, std::uint64_t bps) const {
  const TraceRange trace("pack rows");
  parallel_for(
      records.size(),
      [&](std::uint64_t b, std::uint64_t e) {
        std::string scratch;
        for (std::uint64_t r = b; r < e; ++r) pack(records[r], t, out + r * bps, bps, scratch);
      },
      items_per_worker(2 * bps));
}

// read_fasta, src/hamming_impl.cc:200-242: skip the first line, read up to n
// records (n = 0: all); with remove_duplicates keep the first occurrence of
// each distinct sequence and record, per input, the row it maps to.
FastaData read_fasta(const std::string &filename, bool remove_duplicates, std::size_t n) {
  FastaRecords recs(filename, n);
  FastaData out;
  std::vector<std::size_t> rows(recs.size());
  for (std::size_t k = 0; k < rows.size(); ++k) rows[k] = k;
  if (remove_duplicates) out.sequence_indices = recs.dedup(rows);
  out.data.resize(rows.size());
  parallel_for(
      rows.size(),
      [&](std::uint64_t b, std::uint64_t e) {
        for (std::uint64_t k = b; k < e; ++k) out.data[k] = recs.text(rows[k]);
      },
      items_per_worker(rows.empty() ? 1 : recs.length(rows[0])));
  return out;
}

// fasta_sequence_indices, src/hamming.cc:76-102
std::vector<std::size_t> fasta_sequence_indices(const std::string &fasta_file,
                                                std::size_t n) {
  FastaRecords recs(fasta_file, n);
  std::vector<std::size_t> unique;
  return recs.dedup(unique);
}

bool gpu_available() { return hd_device_count() > 0; }

void check(int status) {
  if (status != HD_OK) throw std::runtime_error(hd_last_error());
}

} // namespace hdhost
