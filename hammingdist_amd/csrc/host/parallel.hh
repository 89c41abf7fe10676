// Minimal fork-join helper for the host layer (no OpenMP dependency).
#pragma once

#include <algorithm>
#include <cstdint>
#include <thread>
#include <vector>

namespace hdhost {

// Calls f(begin, end) on contiguous chunks of [0, n) from up to
// worker_count() threads and joins them.
template <typename F> void parallel_for(std::uint64_t n, F &&f) {
  if (n == 0) return;
  const std::uint64_t workers = std::min<std::uint64_t>(worker_count(), n);
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

} // namespace hdhost
