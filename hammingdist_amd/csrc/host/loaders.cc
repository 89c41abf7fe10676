// Parallel text loaders: from_csv (the DataSet(filename) constructor) and
// from_lower_triangular.
//
// Behaviour kept from the reference (include/hamming/hamming.hh:34-55 and
// :152-166), stated here as the rules this code implements:
//   lines   : the file splits at '\n'; a last piece without '\n' is a line
//             when it holds a character (the std::getline loop).
//   fields  : a line splits at ','.  Field k is the text up to the next ','
//             (or the line's end).  Once a field has run to the line's end,
//             a further field of that line repeats the previous one; a field
//             asked for right after a ',' that ends the line is empty (the
//             istringstream / std::getline(s, d, ',') rules).
//   numbers : each field is read as std::stoi reads it: leading white space,
//             an optional sign, decimal digits, anything after them ignored;
//             no digit throws std::invalid_argument("stoi"), a value outside
//             int throws std::out_of_range("stoi").  The first failing field
//             in file order decides the exception.
//   from_csv: n = the number of '\n' in the file; the result holds
//             n (n + 1) / 2 cells, zero where no line writes; line r
//             (0-based) writes its first r fields, cast to T (wrapping), at
//             cell r (r - 1) / 2.
//   from_lower_triangular: every field of every line (a line of k commas
//             has k + 1 fields, so an empty line is one empty field), in
//             order, each safe_int_cast<T> (capped at T's maximum, negative
//             values wrap); n follows from the cell count.
// A file that cannot be opened reads as empty, as the reference's ifstream
// does.
//
// Implementation: the file is memory-mapped; the lines are found by a
// parallel '\n' scan, split into byte-balanced ranges of lines, and each
// range is parsed by its own thread straight into the result (for
// from_lower_triangular after a parallel count of each line's fields and a
// prefix sum of the counts gives every line its first cell).
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "hamming.hh"

namespace hdhost {

namespace {

// The bytes of a whole file: mapped when it is a regular non-empty file,
// read into memory otherwise; empty when it cannot be opened.
class FileText {
public:
  explicit FileText(const std::string &filename) {
    fd_ = ::open(filename.c_str(), O_RDONLY);
    if (fd_ < 0) return;
    struct stat st {};
    if (::fstat(fd_, &st) == 0 && S_ISREG(st.st_mode) && st.st_size > 0) {
      void *m = ::mmap(nullptr, static_cast<std::size_t>(st.st_size), PROT_READ, MAP_PRIVATE,
                       fd_, 0);
      if (m != MAP_FAILED) {
        map_ = m;
        size_ = static_cast<std::uint64_t>(st.st_size);
        data_ = static_cast<const char *>(m);
        ::madvise(m, static_cast<std::size_t>(size_), MADV_SEQUENTIAL);
        return;
      }
    }
    char chunk[1 << 16];
    ssize_t r;
    while ((r = ::read(fd_, chunk, sizeof chunk)) > 0)
      buffer_.append(chunk, static_cast<std::size_t>(r));
    data_ = buffer_.data();
    size_ = buffer_.size();
  }
  ~FileText() {
    if (map_) ::munmap(map_, static_cast<std::size_t>(size_));
    if (fd_ >= 0) ::close(fd_);
  }
  FileText(const FileText &) = delete;
  FileText &operator=(const FileText &) = delete;
  const char *data() const { return data_; }
  std::uint64_t size() const { return size_; }

private:
  int fd_ = -1;
  void *map_ = nullptr;
  const char *data_ = "";
  std::uint64_t size_ = 0;
  std::string buffer_;
};

// Line starts of the text: starts[k] is the first byte of line k, and
// starts[lines] = size + 1 closes the last line (line k is
// [starts[k], starts[k + 1] - 1) without its '\n').  Also counts the '\n'.
struct Lines {
  std::vector<std::uint64_t> starts;
  std::uint64_t newlines = 0;
  std::uint64_t count() const { return starts.size() - 1; }
  std::uint64_t begin(std::uint64_t k) const { return starts[k]; }
  std::uint64_t end(std::uint64_t k) const { return starts[k + 1] - 1; }
};

Lines find_lines(const char *p, std::uint64_t size) {
  const std::uint64_t workers = std::max<std::uint64_t>(
      1, std::min<std::uint64_t>(worker_count(), size / (1 << 20) + 1));
  std::vector<std::vector<std::uint64_t>> part(workers);
  parallel_for(workers, [&](std::uint64_t w0, std::uint64_t w1) {
    for (std::uint64_t w = w0; w < w1; ++w) {
      const std::uint64_t b = size * w / workers, e = size * (w + 1) / workers;
      const char *q = p + b, *end = p + e;
      while (q < end) {
        const void *nl = std::memchr(q, '\n', static_cast<std::size_t>(end - q));
        if (!nl) break;
        const char *c = static_cast<const char *>(nl);
        part[w].push_back(static_cast<std::uint64_t>(c - p));
        q = c + 1;
      }
    }
  });
  Lines L;
  std::uint64_t total = 0;
  for (const auto &v : part) total += v.size();
  L.newlines = total;
  L.starts.reserve(total + 2);
  L.starts.push_back(0);
  for (const auto &v : part)
    for (std::uint64_t nl : v) L.starts.push_back(nl + 1);
  // the piece after the last '\n' is a line only when it has characters
  if (L.starts.back() < size) L.starts.push_back(size + 1);
  return L;
}

// Splits lines [0, count) into up to `parts` ranges of about equal bytes.
std::vector<std::uint64_t> balanced_line_ranges(const Lines &L, std::uint64_t parts) {
  const std::uint64_t n = L.count();
  std::vector<std::uint64_t> cut{0};
  if (n == 0) {
    cut.push_back(0);
    return cut;
  }
  const std::uint64_t bytes = L.starts[n];
  for (std::uint64_t w = 1; w < parts; ++w) {
    const std::uint64_t target = bytes * w / parts;
    const std::uint64_t k = static_cast<std::uint64_t>(
        std::lower_bound(L.starts.begin(), L.starts.begin() + n, target) - L.starts.begin());
    cut.push_back(std::max(cut.back(), std::min(k, n)));
  }
  cut.push_back(n);
  return cut;
}

// std::getline(istringstream(line), d, ',') step by step over one line.
class Fields {
public:
  Fields(const char *b, const char *e) : p_(b), end_(e) {}
  // The next field into [*fb, *fe); false when the stream had already reached
  // the line's end (the field is then unchanged: the previous one repeats).
  bool next(const char **fb, const char **fe) {
    if (eof_) return false;
    const void *c = std::memchr(p_, ',', static_cast<std::size_t>(end_ - p_));
    if (c) {
      *fb = p_;
      *fe = static_cast<const char *>(c);
      p_ = *fe + 1;
    } else {
      *fb = p_;
      *fe = end_;
      p_ = end_;
      eof_ = true; // this extraction reached the end (empty field included)
    }
    return true;
  }
  bool good() const { return !eof_; }

private:
  const char *p_, *end_;
  bool eof_ = false;
};

enum class NumError { kNone = 0, kInvalid = 1, kRange = 2 };

inline bool is_space(char c) {
  return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r';
}

// std::stoi's reading of [b, e).
inline int read_int(const char *b, const char *e, NumError *err) {
  const char *p = b;
  while (p < e && is_space(*p)) ++p;
  bool neg = false;
  if (p < e && (*p == '+' || *p == '-')) {
    neg = *p == '-';
    ++p;
  }
  if (p >= e || *p < '0' || *p > '9') {
    *err = NumError::kInvalid;
    return 0;
  }
  std::uint64_t v = 0;
  bool over = false;
  for (; p < e && *p >= '0' && *p <= '9'; ++p) {
    v = v * 10 + static_cast<std::uint64_t>(*p - '0');
    if (v > 0x80000000ull) over = true, v = 0x80000000ull;
  }
  if (over || (!neg && v > 0x7FFFFFFFull)) {
    *err = NumError::kRange;
    return 0;
  }
  return neg ? static_cast<int>(-static_cast<std::int64_t>(v)) : static_cast<int>(v);
}

// The first error in file order among per-range errors (range order is
// file order), thrown as the reference's std::stoi would throw it.
struct RangeError {
  NumError err = NumError::kNone;
};

void throw_first(const std::vector<RangeError> &errs) {
  for (const auto &e : errs) {
    if (e.err == NumError::kInvalid) throw std::invalid_argument("stoi");
    if (e.err == NumError::kRange) throw std::out_of_range("stoi");
  }
}

} // namespace

template <typename T> DataSet<T>::DataSet(const std::string &filename) {
  const TraceRange trace("from_csv parse");
  const FileText text(filename);
  const Lines L = find_lines(text.data(), text.size());
  nsamples = L.newlines;
  const std::uint64_t cells = nsamples * (nsamples + 1) / 2;
  result.resize(cells);
  // zero the cells no line writes (line r writes r cells from r (r - 1) / 2)
  const std::uint64_t lines = std::min<std::uint64_t>(L.count(), nsamples + 1);
  const std::uint64_t written = lines * (lines - (lines ? 1 : 0)) / 2;
  std::fill(result.begin() + static_cast<std::ptrdiff_t>(std::min(written, cells)),
            result.end(), T{0});
  const auto cut = balanced_line_ranges(L, worker_count());
  std::vector<RangeError> errs(cut.size() - 1);
  parallel_for(cut.size() - 1, [&](std::uint64_t w0, std::uint64_t w1) {
    for (std::uint64_t w = w0; w < w1; ++w) {
      const std::uint64_t r_end = std::min(cut[w + 1], lines);
      for (std::uint64_t r = cut[w]; r < r_end && errs[w].err == NumError::kNone; ++r) {
        Fields f(text.data() + L.begin(r), text.data() + L.end(r));
        const char *fb = nullptr, *fe = nullptr;
        T *out = result.data() + r * (r - (r ? 1 : 0)) / 2;
        for (std::uint64_t j = 0; j < r; ++j) {
          f.next(&fb, &fe); // past the line's end the previous field repeats
          NumError e = NumError::kNone;
          const int v = read_int(fb, fe, &e);
          if (e != NumError::kNone) {
            errs[w].err = e; // this range stops at its first error
            break;
          }
          out[j] = static_cast<T>(v);
        }
      }
    }
  });
  throw_first(errs);
}

template DataSet<std::uint8_t>::DataSet(const std::string &);
template DataSet<std::uint16_t>::DataSet(const std::string &);

DataSet<DefaultDistIntType> from_csv(const std::string &filename) {
  return DataSet<DefaultDistIntType>(filename);
}

template <typename T> DataSet<T> from_lower_triangular(const std::string &filename) {
  const TraceRange trace("from_lower_triangular parse");
  const FileText text(filename);
  const Lines L = find_lines(text.data(), text.size());
  const std::uint64_t lines = L.count();
  // fields per line: its commas + 1
  std::vector<std::uint64_t> first(lines + 1, 0);
  const auto cut = balanced_line_ranges(L, worker_count());
  parallel_for(cut.size() - 1, [&](std::uint64_t w0, std::uint64_t w1) {
    for (std::uint64_t w = w0; w < w1; ++w)
      for (std::uint64_t r = cut[w]; r < cut[w + 1]; ++r) {
        const char *b = text.data() + L.begin(r), *e = text.data() + L.end(r);
        std::uint64_t commas = 0;
        for (const char *q = b; q < e; ++q) commas += *q == ',';
        first[r + 1] = commas + 1;
      }
  });
  for (std::uint64_t r = 0; r < lines; ++r) first[r + 1] += first[r];
  Cells<T> cells(first[lines]);
  std::vector<RangeError> errs(cut.size() - 1);
  parallel_for(cut.size() - 1, [&](std::uint64_t w0, std::uint64_t w1) {
    for (std::uint64_t w = w0; w < w1; ++w) {
      for (std::uint64_t r = cut[w]; r < cut[w + 1] && errs[w].err == NumError::kNone; ++r) {
        const char *q = text.data() + L.begin(r), *const end = text.data() + L.end(r);
        T *out = cells.data() + first[r];
        while (true) {
          // Fast path: a field of 1-9 digits ending at ',' or the line's end
          // (what the writers write); anything else goes through read_int.
          const char *f0 = q;
          std::uint32_t v = 0;
          while (q < end && static_cast<unsigned>(*q - '0') < 10u && q - f0 < 9)
            v = v * 10 + static_cast<std::uint32_t>(*q++ - '0');
          if (q == f0 || (q < end && *q != ',')) {
            const void *c = std::memchr(f0, ',', static_cast<std::size_t>(end - f0));
            q = c ? static_cast<const char *>(c) : end;
            NumError e = NumError::kNone;
            const int x = read_int(f0, q, &e);
            if (e != NumError::kNone) {
              errs[w].err = e;
              break;
            }
            v = static_cast<std::uint32_t>(x);
            *out++ = safe_int_cast<T>(x);
          } else {
            *out++ = safe_int_cast<T>(static_cast<int>(v));
          }
          if (q == end) break; // the field ran to the line's end
          ++q;                 // past the ',': one more field, maybe empty
        }
      }
    }
  });
  throw_first(errs);
  return DataSet<T>(std::move(cells));
}

template DataSet<std::uint8_t> from_lower_triangular<std::uint8_t>(const std::string &);
template DataSet<std::uint16_t> from_lower_triangular<std::uint16_t>(const std::string &);

} // namespace hdhost
