// FASTA ingest, packing and the engine bridge of the host layer.
#include <cstdio>
#include <cstring>
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

std::vector<GeneBlock> pack_all(const std::vector<std::string> &data, bool include_x,
                                std::uint64_t bps) {
  std::vector<GeneBlock> packed(data.size() * bps);
  const auto lut = lookup_table(include_x);
  parallel_for(data.size(), [&](std::uint64_t b, std::uint64_t e) {
    for (std::uint64_t i = b; i < e; ++i)
      pack_row(data[i].data(), data[i].size(), lut, packed.data() + i * bps, bps);
  });
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

// Whole file in memory.
std::string slurp(const std::string &filename) {
  FILE *f = std::fopen(filename.c_str(), "rb");
  if (!f) throw std::runtime_error("Error: Failed to open file '" + filename + "'");
  std::string buf;
  if (std::fseek(f, 0, SEEK_END) == 0) {
    const long sz = std::ftell(f);
    if (sz > 0) buf.resize(static_cast<std::size_t>(sz));
    std::fseek(f, 0, SEEK_SET);
  }
  std::size_t got = buf.empty() ? 0 : std::fread(buf.data(), 1, buf.size(), f);
  buf.resize(got);
  // Files whose size is not known up front (pipes) are read in chunks.
  char chunk[1 << 16];
  std::size_t r;
  while ((r = std::fread(chunk, 1, sizeof chunk, f)) > 0) buf.append(chunk, r);
  std::fclose(f);
  return buf;
}

// std::getline over an in-memory file, with the stream's end-of-file state:
// a final line without '\n' still succeeds but sets eof; a read at the end
// fails and sets eof.
class LineReader {
public:
  explicit LineReader(const std::string &buf) : buf_(buf) {}
  bool getline(std::string_view &line) {
    if (pos_ >= buf_.size()) {
      eof_ = true;
      return false;
    }
    const char *start = buf_.data() + pos_;
    const void *nl = std::memchr(start, '\n', buf_.size() - pos_);
    if (!nl) {
      line = std::string_view(start, buf_.size() - pos_);
      pos_ = buf_.size();
      eof_ = true;
      return true;
    }
    const std::size_t len = static_cast<const char *>(nl) - start;
    line = std::string_view(start, len);
    pos_ += len + 1;
    return true;
  }
  bool eof() const { return eof_; }

private:
  const std::string &buf_;
  std::size_t pos_ = 0;
  bool eof_ = false;
};

// One record: lines appended until a line whose first char is '>' (an empty
// line's first char reads as '\0'), src/hamming_impl.cc:219-222.
std::string next_record(LineReader &r) {
  std::string seq;
  std::string_view line;
  while (r.getline(line) && (line.empty() || line[0] != '>')) seq.append(line);
  return seq;
}

} // namespace

// read_fasta, src/hamming_impl.cc:200-242: skip the first line, read up to n
// records (n = 0: all); with remove_duplicates keep the first occurrence of
// each distinct sequence and record, per input, the row it maps to.
FastaData read_fasta(const std::string &filename, bool remove_duplicates, std::size_t n) {
  FastaData out;
  const std::string buf = slurp(filename);
  if (n == 0) n = std::numeric_limits<std::size_t>::max();
  LineReader r(buf);
  std::string_view line;
  r.getline(line);
  std::unordered_map<std::string, std::size_t> index;
  std::size_t count = 0;
  while (count < n && !r.eof()) {
    std::string seq = next_record(r);
    if (remove_duplicates) {
      auto res = index.emplace(std::move(seq), out.data.size());
      if (res.second) out.data.push_back(res.first->first);
      out.sequence_indices.push_back(res.first->second);
    } else {
      out.data.push_back(std::move(seq));
    }
    ++count;
  }
  return out;
}

// fasta_sequence_indices, src/hamming.cc:76-102
std::vector<std::size_t> fasta_sequence_indices(const std::string &fasta_file,
                                                std::size_t n) {
  return read_fasta(fasta_file, true, n).sequence_indices;
}

bool gpu_available() { return hd_device_count() > 0; }

void check(int status) {
  if (status != HD_OK) throw std::runtime_error(hd_last_error());
}

} // namespace hdhost
