"""Autotuner of the all-pairs tile shapes.

The tuned space is each kernel's compiled tile variants: rows per lane (the
32-bit words of the row planes each lane holds per step, R), columns per
wavefront (C) and, for batches, the schedule (persistent workgroups on an
atomic tile queue, or one workgroup per tile).  The kernels keep no tile in
LDS (columns are wave-uniform scalar loads), so there is no LDS depth to tune.

For a batch, tune_batch times every batch variant (hd_batch_variant_info) on
the batch itself and keeps the fastest; for a plain dataset resident on the
device, tune_plain times every plain variant (hd_variant_info) on a leading
band of rows of the dataset (a contiguous slice of its triangle).  Winners are
cached in a JSON file keyed by a signature of the work (plane code, cell type,
cap, early exit, and log2 buckets of the job count, rows, words per row and
pairs), so later runs of similar work skip the sweep.

Cache file: $HD_TUNE_CACHE, else ~/.cache/hammingdist_amd/tuning.json.
"""
from __future__ import annotations

import json
import math
import os
import time

from . import capi


def cache_path() -> str:
    return os.environ.get("HD_TUNE_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "hammingdist_amd", "tuning.json")


def _load() -> dict:
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(cache: dict) -> None:
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def _bucket(x: float) -> int:
    return int(round(math.log2(max(x, 1.0))))


def batch_signature(jobs: list[tuple[int, int, int]], early_exit: bool = True) -> str:
    """jobs: (n, bytes_per_seq, out_bytes) per dataset."""
    pairs = sum(n * (n - 1) // 2 for n, _, _ in jobs)
    work = sum(n * (n - 1) // 2 * bps for n, bps, _ in jobs)
    mean_n = sum(n for n, _, _ in jobs) / max(1, len(jobs))
    mean_w = sum(bps / 16 for _, bps, _ in jobs) / max(1, len(jobs))
    widths = "".join(sorted({str(ob) for _, _, ob in jobs}))
    return (f"batch:j{_bucket(len(jobs))}:n{_bucket(mean_n)}:W{_bucket(mean_w)}"
            f":p{_bucket(pairs)}:w{_bucket(work)}:o{widths}:e{int(bool(early_exit))}")


def plain_signature(n: int, bps: int, semantics: int, mask: int, out_bytes: int,
                    max_dist: int, early_exit: bool) -> str:
    """Signature of one plain all-pairs launch: the plane code the engine picks
    (its stored planes), the cell type, the cap (exact below 64, where the
    early exit stops tiles within the first chunks, else its log2 bucket), and
    log2 buckets of the rows and words per row."""
    _, W, P, total = capi.planes_layout(n, bps, semantics, mask)
    code = f"s{semantics}p{P}{'c' if total > W * P * ((n + 511) // 512 * 512) else ''}"
    cap = str(max_dist) if max_dist < 64 else f"b{_bucket(max_dist)}"
    return (f"plain:{code}:n{_bucket(n)}:W{_bucket(W)}:o{out_bytes}:m{cap}"
            f":e{int(bool(early_exit))}")


def _time(fn, sync, reps: int) -> float:
    fn()
    sync()
    t = time.perf_counter()
    for _ in range(reps):
        fn()
    sync()
    return (time.perf_counter() - t) / reps


def tune_batch(batch: "capi.Batch", jobs: list[tuple[int, int, int]], stream: int = 0,
               sync=None, reps: int = 2, early_exit: bool = True, use_cache: bool = True):
    """Prepares `batch` with the fastest batch tile shape; returns
    (variant, {variant: seconds})."""
    if sync is None:
        import torch

        sync = torch.cuda.synchronize
    sig = batch_signature(jobs, early_exit)
    cache = _load() if use_cache else {}
    if use_cache and sig in cache:
        v = int(cache[sig]["variant"])
        batch.prepare(v, early_exit)
        return v, cache[sig].get("times", {})
    times = {}
    for v in range(len(capi.batch_variants())):
        batch.prepare(v, early_exit)
        times[v] = _time(lambda: batch.run(stream), sync, reps)
    best = min(times, key=times.get)
    batch.prepare(best, early_exit)
    cache[sig] = {"variant": best, "times": {str(k): t for k, t in times.items()}}
    if use_cache:
        _store(cache)
    return best, times


def tune_plain(d_planes, n: int, bps: int, semantics: int, mask: int, out_bytes: int,
               max_dist: int, early_exit: bool, d_out, stream: int = 0, sync=None,
               reps: int = 2, sample_rows: int = 8192, use_cache: bool = True):
    """The fastest plain tile variant for this launch; returns (variant,
    {variant: seconds}).  Each variant is timed on the triangle of the first
    min(n, sample_rows) rows (the leading band of the same encoded planes,
    written into the head of d_out), so the sweep costs a fraction of one
    full launch; d_out must hold at least that band's cells."""
    if sync is None:
        import torch

        sync = torch.cuda.synchronize
    sig = plain_signature(n, bps, semantics, mask, out_bytes, max_dist, early_exit)
    cache = _load() if use_cache else {}
    if use_cache and sig in cache:
        return int(cache[sig]["variant"]), cache[sig].get("times", {})
    m = min(n, sample_rows)
    times = {}
    for v in range(len(capi.variants())):
        times[v] = _time(lambda: capi.all_pairs_dev_rows(d_planes, n, bps, semantics, mask,
                                                         out_bytes, max_dist, early_exit, v,
                                                         0, m, d_out, stream),
                         sync, reps)
    best = min(times, key=times.get)
    cache[sig] = {"variant": best, "times": {str(k): t for k, t in times.items()}}
    if use_cache:
        _store(cache)
    return best, times
