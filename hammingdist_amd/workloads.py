"""Synthetic workloads of the benchmark and the tests.

Each workload names the generator settings (hammingdist_amd/csrc/host/synth.hh)
of one benchmark case.  Sequences come from a seeded generator: a uniform
A/C/G/T root, optional clade roots, per-site substitutions, gap runs of '-'
(geometric length, mean 20), and 'N' / 'X' sites; a fraction of sequences are
copies of earlier ones.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace

import numpy as np


@dataclass(frozen=True)
class Spec:
    n: int
    L: int
    seed: int
    p_sub: float = 0.0
    clades: int = 0
    p_clade: float = 0.0
    gap_frac: float = 0.0
    gap_mean: float = 20.0
    p_n: float = 0.0
    p_x: float = 0.0
    dup_frac: float = 0.0

    def kwargs(self) -> dict:
        return dict(n=self.n, L=self.L, seed=self.seed, p_sub=self.p_sub, clades=self.clades,
                    p_clade=self.p_clade, gap_frac=self.gap_frac, gap_mean=self.gap_mean,
                    p_n=self.p_n, p_x=self.p_x, dup_frac=self.dup_frac)


def plain(n: int = 2000, L: int = 29903, seed: int = 1000) -> Spec:
    """Plain all-pairs input: root + 1e-3 substitutions, 0.5% of sites in gap
    runs, 0.2% 'N', 15% exact duplicates (run with remove_duplicates=True,
    max_distance=255, uint8 output)."""
    return Spec(n=n, L=L, seed=seed, p_sub=1e-3, gap_frac=0.005, p_n=0.002, dup_frac=0.15)


def large(n: int = 50_000, L: int = 29903, seed: int = 1001) -> Spec:
    """Large drop-in input: 64 clade roots (2e-3 from the root), each sequence
    3e-4 from its clade root, plus the gaps, 'N' and duplicates of plain()."""
    return replace(plain(n, L, seed), p_sub=3e-4, clades=64, p_clade=2e-3)


def wide(n: int = 100_000, L: int = 10_000, seed: int = 1002) -> Spec:
    """Wide-distance input: root + 5% substitutions and 0.1% 'X' sites
    (include_x=True, uint16 output)."""
    return Spec(n=n, L=L, seed=seed, p_sub=0.05, p_x=0.001)


def batched(count: int = 256, base_seed: int = 1003, n_min: int = 50, n_max: int = 5000,
            L_min: int = 1000, L_max: int = 30000) -> list[tuple[Spec, int]]:
    """Independent datasets: n log-uniform in [n_min, n_max], L uniform in
    [L_min, L_max], plain() generator each with its own seed; out_bytes
    alternates 1 (max 255) and 2 (max 65535)."""
    rng = np.random.default_rng(base_seed)
    out = []
    for k in range(count):
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        L = int(rng.integers(L_min, L_max + 1))
        out.append((plain(n, L, base_seed + k), 1 if k % 2 == 0 else 2))
    return out


def serving_db(n: int = 1_000_000, L: int = 29903, seed: int = 1004) -> Spec:
    """Serving database: the large() generator at n rows."""
    return large(n, L, seed)


def generate_packed(spec: Spec, include_x: bool = False) -> np.ndarray:
    """(n, bytes_per_seq) uint8 packed rows."""
    from . import _native

    flat = _native._synth_packed(**spec.kwargs(), include_x=include_x)
    return flat.reshape(spec.n, -1)


def generate_chars(spec: Spec) -> np.ndarray:
    """(n, L) uint8 chars."""
    from . import _native

    return _native._synth_chars(**spec.kwargs()).reshape(spec.n, spec.L)


def generate_strings(spec: Spec) -> list[str]:
    c = generate_chars(spec)
    return [bytes(row).decode("latin-1") for row in c]


def write_fasta(spec: Spec, path: str) -> None:
    from . import _native

    _native._synth_fasta(path, **spec.kwargs())
