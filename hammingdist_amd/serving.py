"""Request generation and load driving for the serving loop (hd_server_*).

A request is either a reference-distance query — a fresh mutant of a random
database row against a random subset of rows (size log-uniform in
[m_min, m_max], drawn without replacement) — or a small all-pairs job on
n (uniform in [n_min, n_max]) random database rows.  `closed_loop` measures
capacity (requests/s with everything submitted at once); `open_loop` offers a
Poisson stream at a fixed rate and reports requests/s and latency percentiles.
"""
from __future__ import annotations

import math
import threading
import time
from dataclasses import dataclass

import numpy as np

from . import capi

LETTER_NIBBLES = np.array([1, 2, 4, 8], np.uint8)


@dataclass
class Request:
    kind: str  # "query" or "all_pairs"
    rows: np.ndarray
    query: np.ndarray | None = None
    out_bytes: int = 1
    max_dist: int = 255


def mutant_row(row: np.ndarray, L: int, p_sub: float, rng: np.random.Generator) -> np.ndarray:
    """Packed row with each of its L sites substituted (to another letter)
    with probability p_sub."""
    out = row.copy()
    k = rng.binomial(L, p_sub)
    if k == 0:
        return out
    sites = rng.choice(L, size=k, replace=False)
    for s in sites:
        b, hi = divmod(int(s), 2)
        cur = (out[b] >> 4) if hi else (out[b] & 15)
        choices = LETTER_NIBBLES[LETTER_NIBBLES != cur]
        new = int(choices[rng.integers(0, len(choices))])
        out[b] = (out[b] & 0x0F) | (new << 4) if hi else (out[b] & 0xF0) | new
    return out


def make_requests(db: np.ndarray, L: int, count: int, seed: int, query_frac: float = 0.9,
                  m_min: int = 10_000, m_max: int = 1_000_000, n_min: int = 100,
                  n_max: int = 3000, p_sub: float = 3e-4) -> list[Request]:
    rng = np.random.default_rng(seed)
    n_db = db.shape[0]
    perm = rng.permutation(n_db).astype(np.uint64)
    reqs = []
    for _ in range(count):
        if rng.random() < query_frac:
            m = int(round(10 ** rng.uniform(math.log10(m_min), math.log10(min(m_max, n_db)))))
            m = max(1, min(m, n_db))
            off = int(rng.integers(0, n_db))
            rows = np.concatenate([perm[off:off + m], perm[:max(0, off + m - n_db)]])
            src = int(rng.integers(0, n_db))
            reqs.append(Request("query", rows, mutant_row(db[src], L, p_sub, rng)))
        else:
            n = int(rng.integers(n_min, min(n_max, n_db) + 1))
            rows = rng.choice(n_db, size=n, replace=False).astype(np.uint64)
            reqs.append(Request("all_pairs", rows))
    return reqs


def submit(server: "capi.Server", r: Request) -> int:
    if r.kind == "query":
        return server.submit_query(r.query, r.rows, capi.HD_SEM_DENSE)
    return server.submit_all_pairs(r.rows, capi.HD_SEM_DENSE, r.out_bytes, r.max_dist)


def run_alone(server: "capi.Server", reqs: list[Request]):
    """Each request submitted, waited for and read before the next."""
    out = []
    for r in reqs:
        rid = submit(server, r)
        assert server.wait(rid) == capi.HD_REQ_DONE
        out.append(server.result(rid)[0])
    return out


def run_sequential_plain(capi_mod, torch, d_db, n_db: int, bpr: int, reqs: list[Request],
                         stream: int):
    """The requests one at a time through the plain device entry points, as
    a client without the server would run them: a query is one
    hd_reference_distances_dev launch over its row list; an all-pairs job
    gathers its rows, takes their nibble mask, encodes and runs
    hd_all_pairs_dev (early exit on, default tile shape).  Each result is
    downloaded before the next request starts.  d_db: the packed database
    on the device (n_db x bpr bytes)."""
    dev = d_db.device
    rows2d = d_db.view(n_db, bpr)
    out = []
    for r in reqs:
        if r.kind == "query":
            d_q = torch.from_numpy(np.ascontiguousarray(r.query)).to(dev)
            d_rows = torch.from_numpy(r.rows.astype(np.int64)).to(dev)
            d_out = torch.empty(r.rows.size, dtype=torch.int32, device=dev)
            capi_mod.reference_distances_dev(d_q, d_db, n_db, bpr, d_rows, r.rows.size,
                                             capi_mod.HD_SEM_DENSE, d_out, stream)
            out.append(d_out.cpu().numpy().view(np.uint32))
        else:
            n = r.rows.size
            idx = torch.from_numpy(r.rows.astype(np.int64)).to(dev)
            d_packed = rows2d.index_select(0, idx).reshape(-1)
            sem = capi_mod.HD_SEM_DENSE
            mask = capi_mod.nibble_mask_dev(d_packed, n, bpr, stream)
            _, _, _, total = capi_mod.planes_layout(n, bpr, sem, mask)
            d_planes = torch.empty(total, dtype=torch.int32, device=dev)
            capi_mod.encode_dev(d_packed, n, bpr, sem, mask, d_planes, stream)
            cells = n * (n - 1) // 2
            dt = {1: torch.uint8, 2: torch.int16, 4: torch.int32}[r.out_bytes]
            d_out = torch.empty(cells, dtype=dt, device=dev)
            if n >= 2:
                capi_mod.all_pairs_dev(d_planes, n, bpr, sem, mask, r.out_bytes, r.max_dist,
                                       True, -1, d_out, stream)
            host = d_out.cpu().numpy()
            out.append(host.view({1: np.uint8, 2: np.uint16, 4: np.uint32}[r.out_bytes]))
    return out


def closed_loop(server: "capi.Server", reqs: list[Request], submitters: int = 4):
    """All requests at once, submitted by `submitters` client threads (each
    submit stages its inputs in the caller's thread, so several clients
    submit in parallel); the results are read in submission order after
    that.  Returns (results, seconds, latencies); seconds runs from the first
    submission until the last request is done, i.e. its result is in the
    server's pinned host memory: each request's done time is the time its
    submit call returned plus the latency the server measured from inside
    that call (an upper bound).  Copying the results out is the client's
    work and not counted."""
    from concurrent.futures import ThreadPoolExecutor

    def one(r):
        rid = submit(server, r)
        return rid, time.perf_counter()

    t = time.perf_counter()
    with ThreadPoolExecutor(max_workers=submitters) as ex:
        subs = list(ex.map(one, reqs))
    closed_loop.submit_seconds = time.perf_counter() - t
    res, lat, done = [], [], 0.0
    for rid, t_ret in subs:
        assert server.wait(rid) == capi.HD_REQ_DONE
        a, l = server.result(rid)
        res.append(a)
        lat.append(l)
        done = max(done, t_ret + l)
    return res, done - t, lat


def open_loop(server: "capi.Server", reqs: list[Request], rate: float, seed: int = 0):
    """Poisson arrivals at `rate` requests/s from a submitter thread while the
    caller's thread collects results in submission order.  Returns
    (results, seconds from first submission to last completion, latencies)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=len(reqs))
    ids: list[int] = [None] * len(reqs)  # type: ignore[list-item]
    ready = [threading.Event() for _ in reqs]
    errors = []

    def producer():
        t0 = time.perf_counter()
        due = 0.0
        try:
            for k, r in enumerate(reqs):
                due += gaps[k]
                delay = t0 + due - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                ids[k] = submit(server, r)
                ready[k].set()
        except Exception as exc:  # surfaced by the collector
            errors.append(exc)
            for e in ready:
                e.set()

    th = threading.Thread(target=producer, daemon=True)
    t = time.perf_counter()
    th.start()
    res, lat = [], []
    for k in range(len(reqs)):
        ready[k].wait()
        if errors:
            raise errors[0]
        assert server.wait(ids[k]) == capi.HD_REQ_DONE
        a, l = server.result(ids[k])
        res.append(a)
        lat.append(l)
    th.join()
    return res, time.perf_counter() - t, lat


def percentiles(lat: list[float]) -> dict:
    a = np.sort(np.asarray(lat)) * 1e3
    return {"p50_ms": float(np.percentile(a, 50)), "p99_ms": float(np.percentile(a, 99)),
            "mean_ms": float(a.mean()), "max_ms": float(a.max())}
