// Deterministic synthetic aligned-sequence generator (benchmarks and tests).
//
// A dataset is a root sequence (uniform A/C/G/T), optional clade roots (the
// root with per-site substitution probability p_clade), and n sequences:
// each copies its clade root (or the root), then
//   - substitutes each site with probability p_sub by one of the other 3
//     letters,
//   - starts a gap run at each site with probability gap_frac / gap_mean,
//     of geometric length with mean gap_mean, clipped at L,
//   - overwrites each site with 'N' with probability p_n and with 'X' with
//     probability p_x.
// Finally round(dup_frac * n) distinct sequences (never the first) become
// copies of a uniformly chosen earlier sequence, in increasing index order.
// Every sequence draws from its own counter-based stream, so the output does
// not depend on the thread count.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace hdhost {

struct SynthSpec {
  std::uint64_t n = 0, L = 0, seed = 0;
  double p_sub = 0.0;
  std::uint32_t clades = 0;
  double p_clade = 0.0;
  double gap_frac = 0.0, gap_mean = 20.0;
  double p_n = 0.0, p_x = 0.0;
  double dup_frac = 0.0;
};

// n sequences of L chars, back to back.
std::vector<char> synth_chars(const SynthSpec &spec);
// The same sequences, packed (two sites per byte) to bytes_per_seq(L) each.
std::vector<std::uint8_t> synth_packed(const SynthSpec &spec, bool include_x);
// The same sequences as a FASTA file with records ">seq<i>".
void synth_fasta(const SynthSpec &spec, const std::string &path);
// One random mutant of the dataset root: substitutions with p_sub, then
// 'N'/'X' with p_n/p_x (no gaps), from stream `stream`.
std::string synth_mutant_of_root(const SynthSpec &spec, double p_sub, std::uint64_t stream);

} // namespace hdhost
