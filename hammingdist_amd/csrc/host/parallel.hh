// Minimal fork-join helper for the host layer (no OpenMP dependency).
#pragma once

#include <algorithm>
#include <cstdint>
#include <thread>
#include <vector>

namespace hdhost {

// Calls f(begin, end) on contiguous chunks of [0, n) from up to
// worker_count() threads, each with at least min_per_worker items (a thread
// costs tens of microseconds to start, more than packing a few rows), and
// joins them.
template <typename F> void parallel_for(std::uint64_t n, F &&f, std::uint64_t min_per_worker) {
  if (n == 0) return;
  const std::uint64_t workers = std::min<std::uint64_t>(
      worker_count(), std::max<std::uint64_t>(1, n / std::max<std::uint64_t>(1, min_per_worker)));
  if (workers <= 1) {
    f(std::uint64_t{0}, n);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve(workers);
  for (std::uint64_t w = 0; w < workers; ++w) {
    const std::uint64_t b = n * w / workers, e = n * (w + 1) / workers;
    pool.emplace_back([&f, b, e] { f(b, e); });
  }
  for (auto &t : pool) t.join();
}

// Items per worker for items of about `bytes` bytes each: 1 MiB of work.
inline std::uint64_t items_per_worker(std::uint64_t bytes) {
  return std::max<std::uint64_t>(1, (std::uint64_t{1} << 20) / std::max<std::uint64_t>(1, bytes));
}

} // namespace hdhost
