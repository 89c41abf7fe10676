"""Serving loop (hd_server_*): every result equals the oracle's, and the same
request gives the same result alone and in a concurrent mix, including when
a small device budget forces admission to wait for memory."""
import numpy as np
import pytest

import oracle_lib as O
from hammingdist_amd import capi, serving, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def db():
    spec = workloads.large(n=4000, L=3001, seed=77)
    return workloads.generate_packed(spec), spec.L


def small_requests(db, L, count, seed):
    return serving.make_requests(db, L, count, seed, query_frac=0.7, m_min=50, m_max=4000,
                                 n_min=2, n_max=300, p_sub=3e-3)


def oracle_result(db, r):
    if r.kind == "query":
        return np.array([O.distance_dense(r.query, db[int(k)]) for k in r.rows], np.uint32)
    return O.all_pairs_dense(np.ascontiguousarray(db[r.rows.astype(np.int64)]), r.out_bytes,
                             r.max_dist)


def test_results_match_oracle_and_are_identical_alone_and_mixed(db):
    packed, L = db
    reqs = small_requests(packed, L, 60, seed=5)
    srv = capi.Server(packed, device_bytes=1 << 30, host_bytes=256 << 20, streams=3)
    alone = serving.run_alone(srv, reqs)
    for r, a in zip(reqs, alone):
        np.testing.assert_array_equal(a, oracle_result(packed, r))
    mixed, secs, lat = serving.closed_loop(srv, reqs)
    for a, b in zip(alone, mixed):
        np.testing.assert_array_equal(a, b)
    st = srv.stats()
    assert st["done"] == 120
    assert st["batches"] < 120  # requests were batched together
    open_res, _, _ = serving.open_loop(srv, reqs, rate=2000.0, seed=3)
    for a, b in zip(alone, open_res):
        np.testing.assert_array_equal(a, b)
    srv.close()


def test_small_budget_queues_by_footprint(db):
    packed, L = db
    reqs = small_requests(packed, L, 30, seed=9)
    # room for only a few requests at a time
    biggest = max(r.rows.size for r in reqs)
    budget = max(4 << 20, biggest * (packed.shape[1] + 64) * 3)
    srv = capi.Server(packed, device_bytes=budget, host_bytes=64 << 20, streams=2)
    mixed, _, _ = serving.closed_loop(srv, reqs)
    for r, a in zip(reqs, mixed):
        np.testing.assert_array_equal(a, oracle_result(packed, r))
    srv.close()


def test_oversized_request_is_rejected(db):
    packed, L = db
    srv = capi.Server(packed, device_bytes=1 << 20, host_bytes=1 << 20, streams=1)
    rows = np.arange(3000, dtype=np.uint64)
    with pytest.raises(capi.EngineError, match="budget"):
        srv.submit_all_pairs(rows)
    with pytest.raises(capi.EngineError, match="out of range"):
        srv.submit_query(packed[0], np.array([packed.shape[0]], np.uint64))
    srv.close()


@pytest.mark.parametrize("min_entries", [None, "1"])
def test_shared_scan_matches_oracle(db, min_entries, monkeypatch):
    # Queries of overlapping row sets (rows repeated within a query too),
    # DENSE and CHAR_X, submitted together so that they batch: groups of a
    # semantics read each database row once for all of their queries
    # (HD_SHARED_MIN_ENTRIES=1 takes every group of two or more that way).
    packed, L = db
    if min_entries:
        monkeypatch.setenv("HD_SHARED_MIN_ENTRIES", min_entries)
    rng = np.random.default_rng(17)
    n_db = packed.shape[0]
    hot = rng.choice(n_db, size=300, replace=False)
    reqs = []
    for k in range(24):
        m = int(rng.integers(1, 900))
        rows = np.concatenate([rng.choice(hot, size=m // 2), rng.integers(0, n_db, m - m // 2)])
        rng.shuffle(rows)
        src = int(rng.integers(0, n_db))
        sem = capi.HD_SEM_DENSE if k % 3 else capi.HD_SEM_CHAR_X
        reqs.append((src, rows.astype(np.uint64), sem))
    srv = capi.Server(packed, device_bytes=1 << 30, host_bytes=256 << 20, streams=1,
                      max_batch=64)
    ids = [srv.submit_query(packed[src].copy(), rows, sem) for src, rows, sem in reqs]
    chars = workloads.generate_chars(workloads.large(n=n_db, L=L, seed=77))
    text = [bytes(c).decode() for c in chars]
    for rid, (src, rows, sem) in zip(ids, reqs):
        assert srv.wait(rid) == capi.HD_REQ_DONE
        got, _ = srv.result(rid)
        if sem == capi.HD_SEM_DENSE:
            want = [O.distance_dense(packed[src], packed[int(r)]) for r in rows]
        else:
            want = [O.distance_char(text[src], text[int(r)]) for r in rows]
        np.testing.assert_array_equal(got, np.array(want, np.uint32))
    st = srv.stats()
    if min_entries:
        assert st["shared_groups"] >= 1 and st["shared_queries"] >= 2
    srv.close()
