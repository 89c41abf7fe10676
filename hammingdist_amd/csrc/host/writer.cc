// Text output of the host layer: lower-triangular rows and sparse triplets.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <cstring>
#include <fstream>
#include <memory>
#include <stdexcept>

#include "hamming.hh"

namespace hdhost {

void advise_huge_pages(void *p, std::size_t bytes) {
  const std::uintptr_t page = static_cast<std::uintptr_t>(::sysconf(_SC_PAGESIZE));
  const std::uintptr_t b = (reinterpret_cast<std::uintptr_t>(p) + page - 1) / page * page;
  const std::uintptr_t e = (reinterpret_cast<std::uintptr_t>(p) + bytes) / page * page;
  if (e > b) (void)::madvise(reinterpret_cast<void *>(b), e - b, MADV_HUGEPAGE);
}

namespace {

// Decimal text of 0..65535: digits and length, built once.
struct DecimalTable {
  char digits[65536][5];
  std::uint8_t len[65536];
  DecimalTable() {
    for (unsigned v = 0; v < 65536; ++v) {
      char tmp[8];
      const int l = std::snprintf(tmp, sizeof tmp, "%u", v);
      len[v] = static_cast<std::uint8_t>(l);
      std::memcpy(digits[v], tmp, static_cast<std::size_t>(l));
    }
  }
};

const DecimalTable &decimals() {
  static const DecimalTable t;
  return t;
}

inline char *put_uint(char *p, std::uint32_t v) {
  if (v < 65536) {
    const auto &t = decimals();
    std::memcpy(p, t.digits[v], 5);
    return p + t.len[v];
  }
  char tmp[16];
  const int l = std::snprintf(tmp, sizeof tmp, "%u", v);
  std::memcpy(p, tmp, static_cast<std::size_t>(l));
  return p + l;
}

constexpr std::size_t kMaxDigits = 10;

// Rows [r0, r1) as text; `cells` points at cell (r0, 0).  Row i is
// "d(i,0),...,d(i,i-1)\n" (include/hamming/hamming_utils.hh:19-53).
template <typename T>
std::size_t format_rows(const T *cells, std::uint64_t r0, std::uint64_t r1, char *out) {
  char *p = out;
  const T *c = cells;
  for (std::uint64_t i = r0; i < r1; ++i) {
    for (std::uint64_t j = 0; j < i; ++j) {
      p = put_uint(p, static_cast<std::uint32_t>(*c++));
      *p++ = ',';
    }
    if (i > 0) p[-1] = '\n';
  }
  return static_cast<std::size_t>(p - out);
}

// Rows [r0, r1) whose cells are all single digits: row i is exactly 2i bytes.
template <typename T>
std::size_t format_rows_1digit(const T *cells, std::uint64_t r0, std::uint64_t r1, char *out) {
  char *p = out;
  const T *c = cells;
  for (std::uint64_t i = r0; i < r1; ++i) {
    for (std::uint64_t j = 0; j < i; ++j) {
      p[0] = static_cast<char>('0' + *c++);
      p[1] = ',';
      p += 2;
    }
    if (i > 0) p[-1] = '\n';
  }
  return static_cast<std::size_t>(p - out);
}

void pwrite_all(int fd, const char *data, std::size_t len, std::uint64_t off) {
  while (len) {
    const ssize_t w = ::pwrite(fd, data, len, static_cast<off_t>(off));
    if (w < 0) {
      if (errno == EINTR) continue;
      throw std::runtime_error(std::string("write to lower-triangular file failed: ") +
                               std::strerror(errno));
    }
    data += w;
    len -= static_cast<std::size_t>(w);
    off += static_cast<std::uint64_t>(w);
  }
}

} // namespace

RowWriter::RowWriter(const std::string &filename, bool append) : filename_(filename) {
  // No O_APPEND even when appending: on Linux pwrite on an O_APPEND
  // descriptor ignores its offset, so the end is read once instead.  Read
  // access as well where the file allows it, for the mapped path.
  const int flags = O_CREAT | (append ? 0 : O_TRUNC);
  fd_ = ::open(filename.c_str(), O_RDWR | flags, 0644);
  if (fd_ < 0) fd_ = ::open(filename.c_str(), O_WRONLY | flags, 0644);
  else {
    struct stat st;
    mappable_ = ::fstat(fd_, &st) == 0 && S_ISREG(st.st_mode);
  }
  if (fd_ < 0) throw std::runtime_error("Error: Failed to open file '" + filename + "'");
  // HD_WRITER_MAP_BYTES overrides the mapping threshold (it lets the mapped
  // path be exercised on small files).
  if (const char *m = std::getenv("HD_WRITER_MAP_BYTES")) map_bytes_ = std::strtoull(m, nullptr, 10);
  if (append) {
    const off_t end = ::lseek(fd_, 0, SEEK_END);
    offset_ = end > 0 ? static_cast<std::uint64_t>(end) : 0;
  }
}

// Buffered writes into one file serialise on its inode lock; one large write
// through a shared mapping faults pages in from every thread at once (2.5 GB
// of text on ext4: 0.17 s against 0.38 s for 16 threads of pwrite).  Written
// band after band through mappings the same text took 1.1 s on ext4 (0.24 s
// with pwrite), so only writes of map_bytes_ (1 GiB) or more take this path.
// fallocate first reserves the blocks, so the stores cannot hit a full disk;
// a file system without fallocate keeps the pwrite path.
char *RowWriter::map_range(std::uint64_t len, char **base, std::uint64_t *base_len) {
  if (!mappable_ || len == 0) return nullptr;
  const int fa = ::posix_fallocate(fd_, static_cast<off_t>(offset_), static_cast<off_t>(len));
  if (fa != 0) {
    if (fa == ENOSPC || fa == EFBIG || fa == EIO)
      throw std::runtime_error(std::string("write to lower-triangular file failed: ") +
                               std::strerror(fa));
    mappable_ = false; // e.g. EOPNOTSUPP: posix_fallocate's emulation is not used
    return nullptr;
  }
  const std::uint64_t page = static_cast<std::uint64_t>(::sysconf(_SC_PAGESIZE));
  const std::uint64_t start = offset_ / page * page;
  *base_len = offset_ + len - start;
  void *m = ::mmap(nullptr, *base_len, PROT_READ | PROT_WRITE, MAP_SHARED, fd_,
                   static_cast<off_t>(start));
  if (m == MAP_FAILED) {
    mappable_ = false;
    return nullptr;
  }
  *base = static_cast<char *>(m);
  return *base + (offset_ - start);
}

RowWriter::~RowWriter() {
  if (fd_ >= 0) ::close(fd_);
}

void RowWriter::write_raw(const char *data, std::size_t len) {
  pwrite_all(fd_, data, len, offset_);
  offset_ += len;
}

// Rows [row_begin, row_end) of a triangle whose first cell in `cells` is
// (row_begin, 0): chunks of about equal cell counts are formatted in
// parallel, then written in parallel at their prefix-summed file offsets.
template <typename T>
void RowWriter::write(const T *cells, std::uint64_t row_begin, std::uint64_t row_end,
                      std::uint32_t max_value) {
  const TraceRange trace("format + write rows");
  if (row_begin == 0) row_begin = 1; // row 0 has no cells
  if (row_end <= row_begin) return;
  const std::uint64_t c0 = row_begin * (row_begin - 1) / 2;
  const std::uint64_t total = row_end * (row_end - 1) / 2 - c0;
  const std::uint64_t chunks = std::max<std::uint64_t>(
      1, std::min<std::uint64_t>(worker_count() * 4, total / (1u << 16) + 1));
  std::vector<std::uint64_t> bounds{row_begin};
  for (std::uint64_t k = 1; k < chunks; ++k) {
    const double target = static_cast<double>(c0) + static_cast<double>(total) * k / chunks;
    std::uint64_t r = static_cast<std::uint64_t>(0.5 + std::sqrt(2.0 * target + 0.25));
    r = std::max(r, bounds.back());
    r = std::min(r, row_end);
    bounds.push_back(r);
  }
  bounds.push_back(row_end);
  const std::size_t nb = bounds.size() - 1;
  const bool one_digit = max_value < 10;
  if (one_digit) {
    // Every cell is two bytes ("d," or "d\n"), so each chunk's place in the
    // text is known in advance: chunks are formatted straight into the mapped
    // file (one large write) or into the writer's reused buffer, which is
    // then written by as many threads.
    const std::uint64_t bytes = 2 * total;
    char *base = nullptr;
    std::uint64_t base_len = 0;
    char *dst = bytes >= map_bytes_ ? map_range(bytes, &base, &base_len) : nullptr;
    const bool mapped = dst != nullptr;
    if (!mapped) {
      if (buf_cap_ < bytes) {
        buf_.reset(new char[bytes]); // no zero fill: every byte is formatted
        buf_cap_ = bytes;
      }
      dst = buf_.get();
    }
    parallel_for(nb, [&](std::uint64_t b, std::uint64_t e) {
      for (std::uint64_t k = b; k < e; ++k) {
        const std::uint64_t r0 = bounds[k], r1 = bounds[k + 1];
        if (r1 <= r0) continue;
        const std::uint64_t before = r0 * (r0 - 1) / 2 - c0;
        char *out = dst + 2 * before;
        format_rows_1digit(cells + before, r0, r1, out);
        if (!mapped)
          pwrite_all(fd_, out, 2 * (r1 * (r1 - 1) / 2 - r0 * (r0 - 1) / 2),
                     offset_ + 2 * before);
      }
    });
    if (mapped) ::munmap(base, base_len);
    offset_ += bytes;
    return;
  }
  std::vector<std::string> text(nb);
  parallel_for(nb, [&](std::uint64_t b, std::uint64_t e) {
    for (std::uint64_t k = b; k < e; ++k) {
      const std::uint64_t r0 = bounds[k], r1 = bounds[k + 1];
      if (r1 <= r0) continue;
      const std::uint64_t cells_k = r1 * (r1 - 1) / 2 - r0 * (r0 - 1) / 2;
      const T *src = cells + (r0 * (r0 - 1) / 2 - c0);
      text[k].resize(cells_k * (kMaxDigits + 1) + 1);
      text[k].resize(format_rows(src, r0, r1, text[k].data()));
    }
  });
  std::vector<std::uint64_t> off(nb + 1, 0);
  for (std::size_t k = 0; k < nb; ++k) off[k + 1] = off[k] + text[k].size();
  char *base = nullptr;
  std::uint64_t base_len = 0;
  if (char *dst = off[nb] >= map_bytes_ ? map_range(off[nb], &base, &base_len) : nullptr) {
    parallel_for(nb, [&](std::uint64_t b, std::uint64_t e) {
      for (std::uint64_t k = b; k < e; ++k)
        std::memcpy(dst + off[k], text[k].data(), text[k].size());
    });
    ::munmap(base, base_len);
  } else {
    parallel_for(nb, [&](std::uint64_t b, std::uint64_t e) {
      for (std::uint64_t k = b; k < e; ++k)
        pwrite_all(fd_, text[k].data(), text[k].size(), offset_ + off[k]);
    });
  }
  offset_ += off[nb];
}

template void RowWriter::write<std::uint8_t>(const std::uint8_t *, std::uint64_t, std::uint64_t,
                                             std::uint32_t);
template void RowWriter::write<std::uint16_t>(const std::uint16_t *, std::uint64_t,
                                              std::uint64_t, std::uint32_t);

template <typename T>
void write_rows(const std::string &filename, const T *cells, std::uint64_t row_begin,
                std::uint64_t row_end, bool append) {
  RowWriter w(filename, append);
  w.write(cells, row_begin, row_end, 0xFFFFFFFFu);
}

template void write_rows<std::uint8_t>(const std::string &, const std::uint8_t *,
                                       std::uint64_t, std::uint64_t, bool);
template void write_rows<std::uint16_t>(const std::string &, const std::uint16_t *,
                                        std::uint64_t, std::uint64_t, bool);

// write_lower_triangular, include/hamming/hamming_utils.hh:106-110: nothing
// (no file) for an empty vector; a cell count that is not triangular ends in
// an incomplete row terminated by ',' as partial_write_lower_triangular does.
template <typename T>
void write_lower_triangular(const std::string &filename, const Cells<T> &cells) {
  const std::uint64_t size = cells.size();
  if (size == 0) return;
  std::uint64_t rows = 1; // full rows 1..rows-1
  while ((rows + 1) * rows / 2 <= size) ++rows;
  std::uint32_t mx = 0;
  for (const T v : cells) mx = std::max<std::uint32_t>(mx, v);
  RowWriter w(filename, false);
  w.write(cells.data(), 1, rows, mx);
  const std::uint64_t used = rows * (rows - 1) / 2;
  if (used < size) {
    std::string tail;
    char buf[16];
    for (std::uint64_t x = used; x < size; ++x) {
      char *e = put_uint(buf, static_cast<std::uint32_t>(cells[x]));
      tail.append(buf, static_cast<std::size_t>(e - buf));
      tail.push_back(',');
    }
    w.write_raw(tail.data(), tail.size());
  }
}

template void write_lower_triangular<std::uint8_t>(const std::string &,
                                                   const Cells<std::uint8_t> &);
template void write_lower_triangular<std::uint16_t>(const std::string &,
                                                    const Cells<std::uint16_t> &);

// DataSet::dump_sparse, include/hamming/hamming.hh:79-114: "i j d\n" for each
// cell with d <= threshold, in row-major order.  The cells are selected and
// compacted on the GPU (hd_sparse_emit); the host formats the triplets.
template <typename T>
void DataSet<T>::dump_sparse(const std::string &filename, int threshold) const {
  std::FILE *f = std::fopen(filename.c_str(), "wb");
  if (!f) throw std::runtime_error("Error: Failed to open file '" + filename + "'");
  std::unique_ptr<std::FILE, int (*)(std::FILE *)> guard(f, std::fclose);
  // Whole rows only, as the reference's loop over i < nsamples.
  const std::uint64_t rows = nsamples;
  const std::uint64_t cells = rows < 2 ? 0 : rows * (rows - 1) / 2;
  if (cells == 0 || threshold < 0 || cells > result.size()) {
    if (cells > result.size())
      throw std::runtime_error("Error: distance data smaller than nsamples implies");
    return;
  }
  std::uint64_t count = 0;
  int st = hd_sparse_emit(result.data(), rows, sizeof(T), threshold, nullptr, 0, &count);
  if (st != HD_OK && st != HD_ERR_CAPACITY) check(st);
  if (count == 0) return;
  std::vector<std::uint32_t> trip(3 * count);
  check(hd_sparse_emit(result.data(), rows, sizeof(T), threshold, trip.data(), count, &count));
  const std::uint64_t chunks = std::min<std::uint64_t>(worker_count() * 4, count);
  std::vector<std::string> text(chunks);
  parallel_for(chunks, [&](std::uint64_t b, std::uint64_t e) {
    for (std::uint64_t k = b; k < e; ++k) {
      const std::uint64_t t0 = count * k / chunks, t1 = count * (k + 1) / chunks;
      std::string &s = text[k];
      s.resize((t1 - t0) * 3 * (kMaxDigits + 1));
      char *p = s.data();
      for (std::uint64_t t = t0; t < t1; ++t) {
        p = put_uint(p, trip[3 * t]);
        *p++ = ' ';
        p = put_uint(p, trip[3 * t + 1]);
        *p++ = ' ';
        p = put_uint(p, trip[3 * t + 2]);
        *p++ = '\n';
      }
      s.resize(static_cast<std::size_t>(p - s.data()));
    }
  });
  for (const auto &t : text)
    if (!t.empty() && std::fwrite(t.data(), 1, t.size(), f) != t.size())
      throw std::runtime_error("write to sparse file failed");
}

template void DataSet<std::uint8_t>::dump_sparse(const std::string &, int) const;
template void DataSet<std::uint16_t>::dump_sparse(const std::string &, int) const;

} // namespace hdhost
