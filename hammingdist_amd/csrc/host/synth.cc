// Deterministic synthetic aligned-sequence generator (see synth.hh).
#include "synth.hh"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>

#include "hamming.hh"

namespace hdhost {

namespace {

constexpr char kBases[4] = {'A', 'C', 'G', 'T'};

inline std::uint64_t splitmix64(std::uint64_t &x) {
  std::uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// xoshiro256** seeded from (seed, stream) through splitmix64.
class Rng {
public:
  Rng(std::uint64_t seed, std::uint64_t stream) {
    std::uint64_t x = seed * 0x9E3779B97F4A7C15ull ^ (stream + 0x632BE59BD9B4E019ull);
    for (auto &v : s_) v = splitmix64(x);
  }
  std::uint64_t next() {
    const std::uint64_t r = rotl(s_[1] * 5, 7) * 9;
    const std::uint64_t t = s_[1] << 17;
    s_[2] ^= s_[0];
    s_[3] ^= s_[1];
    s_[1] ^= s_[2];
    s_[0] ^= s_[3];
    s_[2] ^= t;
    s_[3] = rotl(s_[3], 45);
    return r;
  }
  // uniform in [0, m)
  std::uint64_t below(std::uint64_t m) {
    return static_cast<std::uint64_t>((static_cast<unsigned __int128>(next()) * m) >> 64);
  }
  // uniform in (0, 1]
  double unit() { return (static_cast<double>(next() >> 11) + 1.0) * 0x1.0p-53; }
  // Gap before the next success of Bernoulli(p) trials (0 = the next trial).
  std::uint64_t skip(double p) {
    if (p <= 0.0) return UINT64_MAX;
    if (p >= 1.0) return 0;
    const double g = std::floor(std::log(unit()) / std::log1p(-p));
    return g >= 1.8e19 ? UINT64_MAX : static_cast<std::uint64_t>(g);
  }

private:
  static std::uint64_t rotl(std::uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  std::uint64_t s_[4];
};

inline int base_index(char c) {
  switch (c) {
  case 'A': return 0;
  case 'C': return 1;
  case 'G': return 2;
  default: return 3;
  }
}

// Substitutes each site of s[0..L) with probability p.
void substitute(char *s, std::uint64_t L, double p, Rng &r) {
  for (std::uint64_t i = r.skip(p); i < L;) {
    s[i] = kBases[(base_index(s[i]) + 1 + r.below(3)) & 3];
    const std::uint64_t g = r.skip(p);
    if (g >= L - i) break;
    i += g + 1;
  }
}

void overwrite(char *s, std::uint64_t L, double p, char c, Rng &r) {
  for (std::uint64_t i = r.skip(p); i < L;) {
    s[i] = c;
    const std::uint64_t g = r.skip(p);
    if (g >= L - i) break;
    i += g + 1;
  }
}

void gap_runs(char *s, std::uint64_t L, double frac, double mean, Rng &r) {
  if (frac <= 0.0) return;
  const double start_p = frac / mean;
  const double stop_p = 1.0 / mean; // run length ~ Geometric(mean) on {1, 2, ...}
  for (std::uint64_t i = r.skip(start_p); i < L;) {
    const std::uint64_t len = 1 + r.skip(stop_p);
    const std::uint64_t end = len >= L - i ? L : i + len;
    std::memset(s + i, '-', end - i);
    const std::uint64_t g = r.skip(start_p);
    if (g >= L - i) break;
    i += g + 1;
  }
}

struct Roots {
  std::string root;
  std::vector<std::string> clades;
};

constexpr std::uint64_t kRootStream = 1ull << 62;
constexpr std::uint64_t kCladeStream = 1ull << 61;
constexpr std::uint64_t kDupStream = 1ull << 60;
constexpr std::uint64_t kMutantStream = 1ull << 59;

Roots make_roots(const SynthSpec &spec) {
  Roots roots;
  roots.root.resize(spec.L);
  Rng r(spec.seed, kRootStream);
  for (std::uint64_t i = 0; i < spec.L; ++i) roots.root[i] = kBases[r.next() >> 62];
  for (std::uint32_t c = 0; c < spec.clades; ++c) {
    Rng rc(spec.seed, kCladeStream + c);
    std::string s = roots.root;
    substitute(s.data(), spec.L, spec.p_clade, rc);
    roots.clades.push_back(std::move(s));
  }
  return roots;
}

// Sequence i (before duplication) into out[0..L).
void make_sequence(const SynthSpec &spec, const Roots &roots, std::uint64_t i, char *out) {
  Rng r(spec.seed, i);
  const std::string &base =
      roots.clades.empty() ? roots.root : roots.clades[r.below(roots.clades.size())];
  std::memcpy(out, base.data(), spec.L);
  substitute(out, spec.L, spec.p_sub, r);
  gap_runs(out, spec.L, spec.gap_frac, spec.gap_mean, r);
  overwrite(out, spec.L, spec.p_n, 'N', r);
  overwrite(out, spec.L, spec.p_x, 'X', r);
}

// (target, source) copies of the duplication step, in application order.
std::vector<std::pair<std::uint64_t, std::uint64_t>> duplicates(const SynthSpec &spec) {
  std::vector<std::pair<std::uint64_t, std::uint64_t>> out;
  if (spec.n < 2 || spec.dup_frac <= 0.0) return out;
  std::uint64_t k = static_cast<std::uint64_t>(std::llround(spec.dup_frac * spec.n));
  k = std::min<std::uint64_t>(k, spec.n - 1);
  Rng r(spec.seed, kDupStream);
  std::vector<std::uint64_t> idx(spec.n - 1);
  for (std::uint64_t i = 0; i < idx.size(); ++i) idx[i] = i + 1;
  for (std::uint64_t i = 0; i < k; ++i) std::swap(idx[i], idx[i + r.below(idx.size() - i)]);
  idx.resize(k);
  std::sort(idx.begin(), idx.end());
  for (auto t : idx) out.emplace_back(t, r.below(t));
  return out;
}

void check_spec(const SynthSpec &spec) {
  if (spec.L == 0) throw std::runtime_error("synthetic sequences need L > 0");
}

} // namespace

std::vector<char> synth_chars(const SynthSpec &spec) {
  check_spec(spec);
  const Roots roots = make_roots(spec);
  std::vector<char> out(spec.n * spec.L);
  parallel_for(spec.n, [&](std::uint64_t b, std::uint64_t e) {
    for (std::uint64_t i = b; i < e; ++i) make_sequence(spec, roots, i, out.data() + i * spec.L);
  });
  for (auto [t, s] : duplicates(spec))
    std::memcpy(out.data() + t * spec.L, out.data() + s * spec.L, spec.L);
  return out;
}

std::vector<std::uint8_t> synth_packed(const SynthSpec &spec, bool include_x) {
  check_spec(spec);
  const Roots roots = make_roots(spec);
  const std::uint64_t bps = bytes_per_seq(spec.L);
  std::vector<std::uint8_t> out(spec.n * bps);
  const auto lut = lookup_table(include_x);
  parallel_for(spec.n, [&](std::uint64_t b, std::uint64_t e) {
    std::string buf(spec.L, 'A');
    for (std::uint64_t i = b; i < e; ++i) {
      make_sequence(spec, roots, i, buf.data());
      pack_row(buf.data(), spec.L, lut, out.data() + i * bps, bps);
    }
  });
  for (auto [t, s] : duplicates(spec))
    std::memcpy(out.data() + t * bps, out.data() + s * bps, bps);
  return out;
}

void synth_fasta(const SynthSpec &spec, const std::string &path) {
  const std::vector<char> chars = synth_chars(spec);
  std::FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("Error: Failed to open file '" + path + "'");
  std::unique_ptr<std::FILE, int (*)(std::FILE *)> guard(f, std::fclose);
  std::string head;
  for (std::uint64_t i = 0; i < spec.n; ++i) {
    head = ">seq" + std::to_string(i) + "\n";
    if (std::fwrite(head.data(), 1, head.size(), f) != head.size() ||
        std::fwrite(chars.data() + i * spec.L, 1, spec.L, f) != spec.L ||
        std::fputc('\n', f) == EOF)
      throw std::runtime_error("write to FASTA file failed");
  }
}

std::string synth_mutant_of_root(const SynthSpec &spec, double p_sub, std::uint64_t stream) {
  check_spec(spec);
  const Roots roots = make_roots(spec);
  std::string s = roots.root;
  Rng r(spec.seed, kMutantStream + stream);
  substitute(s.data(), spec.L, p_sub, r);
  overwrite(s.data(), spec.L, spec.p_n, 'N', r);
  overwrite(s.data(), spec.L, spec.p_x, 'X', r);
  return s;
}

} // namespace hdhost
