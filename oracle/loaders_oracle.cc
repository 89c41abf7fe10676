// loaders_oracle.cc — CPU restatement of the reference's two text loaders,
// with the same standard-library stream calls, for checking the engine's
// parallel loaders (hammingdist_amd/csrc/host/loaders.cc).
//
// TEST INFRASTRUCTURE ONLY: only tests/ loads it; the product never does.
//
//   oracle_from_csv               DataSet(filename), include/hamming/hamming.hh:34-55
//   oracle_from_lower_triangular  from_lower_triangular, include/hamming/hamming.hh:152-166
//                                 with safe_int_cast, include/hamming/hamming_impl.hh:37-45
//
// Each returns 0, or 1 where the reference throws std::invalid_argument and
// 2 where it throws std::out_of_range; cells are malloc'd (free with
// oracle_free) as out_bytes-wide values.
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

template <typename T> T safe_cast(int x) {
  if (x > static_cast<int>(std::numeric_limits<T>::max())) return std::numeric_limits<T>::max();
  return static_cast<T>(x);
}

template <typename T> int csv(const char *path, void **cells, uint64_t *n_cells, uint64_t *n) {
  std::ifstream count_stream(path);
  const std::size_t samples = std::count(std::istreambuf_iterator<char>(count_stream),
                                         std::istreambuf_iterator<char>(), '\n');
  std::vector<T> result(samples * (samples + 1) / 2);
  std::ifstream stream(path);
  std::string line;
  std::size_t i = 0, row = 0;
  try {
    while (std::getline(stream, line)) {
      std::istringstream fields(line);
      std::string field;
      for (std::size_t j = 0; j < row; ++j) {
        std::getline(fields, field, ',');
        result[i++] = static_cast<T>(std::stoi(field));
      }
      ++row;
    }
  } catch (const std::invalid_argument &) {
    return 1;
  } catch (const std::out_of_range &) {
    return 2;
  }
  *n = samples;
  *n_cells = result.size();
  *cells = std::malloc(std::max<std::size_t>(1, result.size() * sizeof(T)));
  std::memcpy(*cells, result.data(), result.size() * sizeof(T));
  return 0;
}

template <typename T>
int lower_triangular(const char *path, void **cells, uint64_t *n_cells) {
  std::vector<T> result;
  std::ifstream stream(path);
  std::string line;
  try {
    while (std::getline(stream, line)) {
      std::istringstream fields(line);
      std::string field;
      while (fields.good()) {
        std::getline(fields, field, ',');
        result.push_back(safe_cast<T>(std::stoi(field)));
      }
    }
  } catch (const std::invalid_argument &) {
    return 1;
  } catch (const std::out_of_range &) {
    return 2;
  }
  *n_cells = result.size();
  *cells = std::malloc(std::max<std::size_t>(1, result.size() * sizeof(T)));
  std::memcpy(*cells, result.data(), result.size() * sizeof(T));
  return 0;
}

} // namespace

extern "C" {

int oracle_from_csv(const char *path, int out_bytes, void **cells, uint64_t *n_cells,
                    uint64_t *nsamples) {
  return out_bytes == 1 ? csv<uint8_t>(path, cells, n_cells, nsamples)
                        : csv<uint16_t>(path, cells, n_cells, nsamples);
}

int oracle_from_lower_triangular(const char *path, int out_bytes, void **cells,
                                 uint64_t *n_cells) {
  return out_bytes == 1 ? lower_triangular<uint8_t>(path, cells, n_cells)
                        : lower_triangular<uint16_t>(path, cells, n_cells);
}

void oracle_free(void *p) { std::free(p); }
}
