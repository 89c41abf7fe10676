"""Row-band split of an all-pairs triangle across GPUs.

The linear lower triangle (cell (i, j), i > j, at i(i-1)/2 + j) splits into
contiguous row bands whose cells are contiguous slices of it, so the bands of
G GPUs concatenate in row order with no exchange.  Band k holds rows
[n sqrt(k/G), n sqrt((k+1)/G)) rounded to the tile alignment, which gives
every band about the same number of cells (the work of a band is its cells
times the words per row).  Each GPU computes its band with
hd_all_pairs_dev_rows over the planes of rows 0 .. band end (one broadcast of
the packed rows, then a local encode).  This build runs one GPU; the bands
here are computed on it one after another by the tests.
"""
from __future__ import annotations

import math

# every tile variant's rows (64 x rows per lane) divide this
ALIGN = 512


def row_bands(n: int, gpus: int, align: int = ALIGN) -> list[tuple[int, int]]:
    """[(row_begin, row_end)] of `gpus` bands of about equal cell counts;
    every boundary but n is a multiple of `align`.  Bands may be empty when
    n is small."""
    if gpus < 1:
        raise ValueError("gpus must be >= 1")
    total = band_cells(0, n)
    cuts = [0]
    for k in range(1, gpus):
        # the aligned row nearest (in cells) to the k/G point of the triangle
        r = n * math.sqrt(k / gpus)
        lo = int(r // align) * align
        hi = lo + align
        target = total * k / gpus
        best = min((c for c in (lo, hi) if c <= n),
                   key=lambda c: abs(band_cells(0, c) - target), default=lo)
        cuts.append(min(max(best, cuts[-1]), n))
    cuts.append(n)
    return [(cuts[k], cuts[k + 1]) for k in range(gpus)]


def band_cells(row_begin: int, row_end: int) -> int:
    """Cells of rows [row_begin, row_end) of the triangle."""
    tri = lambda r: r * (r - 1) // 2 if r > 0 else 0  # noqa: E731
    return tri(row_end) - tri(row_begin)
