// Shared internals of the HIP engine: error reporting, HIP call checking,
// RAII device/pinned buffers and the plane encoding chosen per semantics.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "hammingdist_amd.h"
#include "kernels.h"

namespace hd {

// A roctx range over a scope: the engine's host-side phases (uploads,
// encodes, launches, downloads, callbacks) show up by name in
// `rocprofv3 --marker-trace` beside the kernels; without a profiler a range
// costs two calls into the roctx stub.
class TraceRange {
public:
  explicit TraceRange(const char *name) { roctxRangePushA(name); }
  ~TraceRange() { roctxRangePop(); }
  TraceRange(const TraceRange &) = delete;
  TraceRange &operator=(const TraceRange &) = delete;
};

// Error carried from an internal failure to the C ABI status code.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &msg) : std::runtime_error(msg), code(c) {}
};

void set_last_error(const std::string &msg);

#define HD_HIP(call)                                                           \
  do {                                                                         \
    hipError_t hd_e_ = (call);                                                 \
    if (hd_e_ != hipSuccess)                                                   \
      throw ::hd::Error(HD_ERR_HIP, std::string(#call) + ": " +               \
                                        hipGetErrorString(hd_e_));            \
  } while (0)

#define HD_REQUIRE(cond, msg)                                                  \
  do {                                                                         \
    if (!(cond))                                                               \
      throw ::hd::Error(HD_ERR_INVALID, (msg));                                \
  } while (0)

// Runs f, converting exceptions into a status code plus hd_last_error text.
template <typename F> int guarded(F &&f) {
  try {
    f();
    return HD_OK;
  } catch (const Error &e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc &) {
    set_last_error("host allocation failed");
    return HD_ERR_CAPACITY;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return HD_ERR_INVALID;
  }
}

void ensure_device();

// Device buffer owned for the duration of one call.
// Device blocks of up to kPoolBlockMax bytes are kept for reuse instead of
// freed: hipFree of a 1-16 MiB block took 46-261 us and hipMalloc 12-34 us,
// most of a small call's time (an all-pairs of 3 rows of 29,903 sites took
// 0.58 ms).  A block goes back to the pool after a device synchronisation,
// which hipFree performs as well, so no work still using it can see it
// handed out again.  Sizes are rounded to a power of two up to 1 MiB and to
// a multiple of 2 MiB above; at most kPoolCachedMax bytes stay cached per
// device (blocks beyond that are freed).
constexpr size_t kPoolBlockMax = 64ull << 20;
constexpr size_t kPoolCachedMax = 512ull << 20;
size_t pool_class(size_t bytes);
// A cached block of the class on the current device (whose id goes to *dev),
// or a new one; pool_give returns it to that device's blocks.
void *pool_take(size_t cls, int *dev);
void pool_give(void *p, size_t cls, int dev);

class DeviceBuffer {
public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t bytes) { alloc(bytes); }
  ~DeviceBuffer() { release(); }
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
  void alloc(size_t bytes) {
    release();
    if (bytes && bytes <= kPoolBlockMax) {
      cls_ = pool_class(bytes);
      p_ = pool_take(cls_, &dev_);
    } else if (bytes) {
      HD_HIP(hipMalloc(&p_, bytes));
    }
    n_ = bytes;
  }
  void release() {
    if (p_ && cls_) pool_give(p_, cls_, dev_);
    else if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
    cls_ = 0;
  }
  template <typename T = void> T *get() const { return static_cast<T *>(p_); }
  size_t size() const { return n_; }

private:
  void *p_ = nullptr;
  size_t n_ = 0;
  size_t cls_ = 0; // pooled block's size class, 0 when not pooled
  int dev_ = 0;     // pooled block's device
};

// Pinned host buffer (hipHostMalloc) for asynchronous copies.
class PinnedBuffer {
public:
  PinnedBuffer() = default;
  explicit PinnedBuffer(size_t bytes) { alloc(bytes); }
  ~PinnedBuffer() { release(); }
  PinnedBuffer(const PinnedBuffer &) = delete;
  PinnedBuffer &operator=(const PinnedBuffer &) = delete;
  void alloc(size_t bytes) {
    release();
    if (bytes) HD_HIP(hipHostMalloc(&p_, bytes, hipHostMallocDefault));
    n_ = bytes;
  }
  void release() {
    if (p_) (void)hipHostFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  template <typename T = void> T *get() const { return static_cast<T *>(p_); }
  size_t size() const { return n_; }

private:
  void *p_ = nullptr;
  size_t n_ = 0;
};

class Stream {
public:
  Stream() { HD_HIP(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking)); }
  ~Stream() { (void)hipStreamDestroy(s_); }
  Stream(const Stream &) = delete;
  Stream &operator=(const Stream &) = delete;
  hipStream_t get() const { return s_; }

private:
  hipStream_t s_{};
};

class Event {
public:
  Event() { HD_HIP(hipEventCreateWithFlags(&e_, hipEventDisableTiming)); }
  ~Event() { (void)hipEventDestroy(e_); }
  Event(const Event &) = delete;
  Event &operator=(const Event &) = delete;
  hipEvent_t get() const { return e_; }

private:
  hipEvent_t e_{};
};

// Bit-sliced plane encoding of one dataset: every nibble code maps to a
// `planes`-bit mask, one bit per plane.  `op` names the tile kernel's pair
// function (its template parameter P, kernels.h): op = planes (4, 5) counts a
// site when no plane is set in both sequences; op = 3 is the dense 3-op code
// over 4 planes whose rows also carry per-row correction counts after the
// planes; op = 7, 9 are the sparse codes with '-' as a code (2 or 3 code
// bits and a '-' plane, correction counts after the planes); op = 8 is the
// sparse class code of very long rows (3 code bits and a '-' plane).
struct PlaneCode {
  int planes = 4;
  int op = 4;
  uint8_t table[16] = {};
};

PlaneCode plane_code(int semantics, uint32_t nibble_mask, uint32_t words = 0);
// True for the 4-plane code equal to the nibble (the dense semantics).
bool is_identity(const PlaneCode &c);
// True for the dense 3-op code.
inline bool is_dense3(const PlaneCode &c) { return c.op == 3; }
// encode_tile mode (kernels.h EncodeMode) for a code.
int encode_mode(const PlaneCode &c);

inline uint64_t tri_cells(uint64_t n) { return n < 2 ? 0 : n * (n - 1) / 2; }

// Geometry of one encoded dataset (planes[g][k][p][64], see kernels.h).
struct Geometry {
  uint64_t n = 0, bpr = 0, n_pad = 0;
  uint32_t W = 0; // 32-site words per row
  PlaneCode code;
  uint64_t plane_words() const { return (uint64_t)W * code.planes * n_pad; }
  // Planes, then (codes with correction counts) two whole-row count slots
  // and two per-32-word-chunk slots (kernels.h, dense3_corr_words and
  // gap_corr_words).
  uint64_t total_words() const {
    return plane_words() +
           (has_corr(code.op) ? (2 + 2 * (uint64_t)((W + 31) / 32)) * n_pad : 0);
  }
};
Geometry geometry(uint64_t n, uint64_t bpr, int semantics, uint32_t mask);

// Tile shapes of the all-pairs kernels: rows per lane R, columns per wave C,
// waves per workgroup NW; a tile is 64 R rows by NW C columns.
struct Variant {
  int R, C, NW;
};
int variant_count();
const Variant &variant_of(int v);
// The tile variant run_allpairs uses for geometry g (its n rows) when the
// caller names none.
int default_variant(const Geometry &g, bool early_exit, uint32_t max_dist);

void launch_encode(const Geometry &g, const uint8_t *d_packed, uint32_t *d_planes,
                   hipStream_t s);
void launch_nibble_mask(const uint8_t *d_packed, uint64_t bytes, uint32_t *d_mask,
                        hipStream_t s);
// All-pairs triangle of row blocks [rb0, rb1) (plain, one launch).
void run_allpairs(const Geometry &g, const uint32_t *planes, int out_bytes,
                  uint32_t max_dist, bool early_exit, int variant, uint64_t rb0,
                  uint64_t rb1, void *out, uint64_t cell_offset, hipStream_t s);
uint32_t check_interval(bool early_exit, uint32_t max_dist, uint32_t W);
hipStream_t lib_stream();
hipStream_t lib_copy_stream();

} // namespace hd
