"""Autotuner of the all-pairs tile shapes.

For a batch it times every compiled batch tile shape (rows per lane, columns
per wave, waves per workgroup; hd_batch_variant_info) on the batch itself and
keeps the fastest; for a plain dataset it does the same over the plain tile
shapes (hd_variant_info) on a sample of rows.  Winners are cached in a JSON
file keyed by a coarse signature of the work (plane count, cell type, size
buckets), so later runs of similar work skip the sweep.

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


def batch_signature(jobs: list[tuple[int, int, int]]) -> str:
    """jobs: (n, bytes_per_seq, out_bytes) per dataset."""
    pairs = sum(n * (n - 1) // 2 for n, _, _ in jobs)
    work = sum(n * (n - 1) // 2 * bps for n, bps, _ in jobs)
    mean_n = sum(n for n, _, _ in jobs) / max(1, len(jobs))
    return f"batch:j{_bucket(len(jobs))}:n{_bucket(mean_n)}:p{_bucket(pairs)}:w{_bucket(work)}"


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
    sig = batch_signature(jobs)
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
