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


@pytest.mark.parametrize("mode", ["default", "shared", "rect"])
def test_shared_scan_matches_oracle(db, mode, monkeypatch):
    # Queries of overlapping row sets (rows repeated within a query too),
    # DENSE and CHAR_X, submitted together so that they batch: groups of a
    # semantics read each database row once for all of their queries
    # (shared: HD_SHARED_MIN_ENTRIES=1 takes every group of two or more that
    # way, with the query rectangles off), or the DENSE ones run as one query
    # rectangle over the whole database (rect: HD_RECT_FORCE=1).
    packed, L = db
    if mode == "shared":
        monkeypatch.setenv("HD_SHARED_MIN_ENTRIES", "1")
        monkeypatch.setenv("HD_SERVER_RECT", "0")
    elif mode == "rect":
        monkeypatch.setenv("HD_RECT_FORCE", "1")
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
    if mode == "shared":
        assert st["shared_groups"] >= 1 and st["shared_queries"] >= 2
        assert not st["rect_enabled"] and st["rect_groups"] == 0
    elif mode == "rect":
        assert st["rect_enabled"] and st["rect_groups"] >= 1 and st["rect_queries"] >= 2
    srv.close()


def test_query_rectangles_uneven_groups(db, monkeypatch):
    # 70 DENSE queries of the dense 3-op codes in one batch (a rectangle of
    # 64 and one of 6: partial wavefronts and a partial last row block of the
    # 4,000-row database), beside DENSE queries holding X (not of the code:
    # their own scans) and CHAR_X queries; every result against the oracle,
    # and each the same as the request alone.
    packed, L = db
    monkeypatch.setenv("HD_RECT_FORCE", "1")
    rng = np.random.default_rng(23)
    n_db, bps = packed.shape
    reqs = []
    srcs = []
    for k in range(84):
        m = int(rng.integers(1, 2 * n_db))  # rows repeat within a query
        rows = rng.integers(0, n_db, m).astype(np.uint64)
        src = int(rng.integers(0, n_db))
        q = packed[src].copy()
        sem = capi.HD_SEM_DENSE
        if k % 12 == 5:
            q[int(rng.integers(0, bps))] = 0x33  # X at two sites
        elif k % 12 == 11:
            sem = capi.HD_SEM_CHAR_X
        reqs.append((q, rows, sem))
        srcs.append(src)
    srv = capi.Server(packed, device_bytes=2 << 30, host_bytes=512 << 20, streams=1,
                      max_batch=128)
    ids = [srv.submit_query(q, rows, sem) for q, rows, sem in reqs]
    got = []
    for rid in ids:
        assert srv.wait(rid) == capi.HD_REQ_DONE
        got.append(srv.result(rid)[0])
    st = srv.stats()
    assert st["rect_enabled"] and st["rect_groups"] >= 2 and st["rect_queries"] == 70
    for (q, rows, sem), g in zip(reqs[:20], got[:20]):
        alone = serving.run_alone(srv, [serving.Request("query", rows, q)])[0] \
            if sem == capi.HD_SEM_DENSE else g
        np.testing.assert_array_equal(g, alone)
    chars = workloads.generate_chars(workloads.large(n=n_db, L=L, seed=77))
    text = [bytes(c).decode() for c in chars]
    uniq_cache = {}
    for (q, rows, sem), src, g in zip(reqs, srcs, got):
        if sem == capi.HD_SEM_DENSE:
            key = q.tobytes()
            if key not in uniq_cache:
                uniq_cache[key] = np.array([O.distance_dense(q, packed[r]) for r in range(n_db)],
                                           np.uint32)
            np.testing.assert_array_equal(g, uniq_cache[key][rows.astype(np.int64)])
        else:
            want = [O.distance_char(text[src], text[int(r)]) for r in rows[:200]]
            np.testing.assert_array_equal(g[:200], np.array(want, np.uint32))
    srv.close()


def test_query_rectangles_off_for_codes_outside_the_dense_code(db, monkeypatch):
    # A database holding X (nibble 3) keeps the per-query and shared scans.
    packed, L = db
    monkeypatch.setenv("HD_RECT_FORCE", "1")
    with_x = packed.copy()
    with_x[5, 3] = 0x33
    srv = capi.Server(with_x, device_bytes=1 << 30, host_bytes=256 << 20, streams=1)
    rows = np.arange(with_x.shape[0], dtype=np.uint64)
    ids = [srv.submit_query(with_x[k], rows) for k in (0, 5, 9)]
    for k, rid in zip((0, 5, 9), ids):
        assert srv.wait(rid) == capi.HD_REQ_DONE
        got, _ = srv.result(rid)
        want = [O.distance_dense(with_x[k], with_x[r]) for r in range(with_x.shape[0])]
        np.testing.assert_array_equal(got, np.array(want, np.uint32))
    st = srv.stats()
    assert not st["rect_enabled"] and st["rect_groups"] == 0
    srv.close()
