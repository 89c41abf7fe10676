"""Parity on the benchmark workloads themselves, at the sizes they run at.

large  n=50,000, L=29,903, 64-clade generator: from_fasta_to_lower_triangular
       with max_distance=2 and the threshold-3 sparse triplets (the reference
       pins the streamed file against its CPU route in src/hamming_t.cc:411-451
       and the sparse dump in python/tests/test_hammingdist.py:149-204).
wide   n=100,000, L=10,000, 5% substitutions + 0.1% X: from_fasta_large with
       include_x=True (the sparse route, src/hamming_impl_t.cc:106-180) and
       fasta_reference_distances against one fresh mutant.
batched 256 datasets of uneven n and L in one tuned launch: every job's
       triangle against the oracle.
serving the 1,000,000-row database resident in HBM under the bench's request
       mix: alone == mixed, sampled results against the oracle.

Each workload runs once at a size the oracle finishes in seconds (every cell,
every byte of the text compared) and once at full size, where cells sampled
over the whole triangle (the first and last rows, the cells around 2^31 and
2^32) are compared with the oracle's pair functions and size-independent
properties tie the rest together (the capped triangle equals the clamped
uncapped one; the sparse emit selects exactly the cells under the threshold).
A short-row triangle of 100,000 rows (5.0e9 cells, indices past 2^32) is
compared with the oracle on every cell.
"""
import os

import numpy as np
import pytest

import hammingdist_amd as hd
import oracle_lib as O
from hammingdist_amd import capi, workloads

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SAMPLE = 4096


def stream():
    return torch.cuda.current_stream().cuda_stream


def device_triangle(packed, sem, out_bytes, max_dist, early, variant=-1):
    """Device planes + triangle of packed host rows; returns (d_out, mask)."""
    n, bps = packed.shape
    sh = stream()
    d_packed = torch.from_numpy(np.ascontiguousarray(packed).reshape(-1)).cuda()
    mask = capi.nibble_mask_dev(d_packed, n, bps, sh)
    _, _, _, total = capi.planes_layout(n, bps, sem, mask)
    d_planes = torch.empty(total, dtype=torch.int32, device="cuda")
    capi.encode_dev(d_packed, n, bps, sem, mask, d_planes, sh)
    d_out = torch.empty(n * (n - 1) // 2 * out_bytes, dtype=torch.uint8, device="cuda")
    capi.all_pairs_dev(d_planes, n, bps, sem, mask, out_bytes, max_dist, early, variant, d_out,
                       sh)
    torch.cuda.synchronize()
    del d_planes, d_packed
    return d_out


def gather(d_cells, out_bytes, idx):
    """Cells idx (uint64 array) of a device triangle, as int64 on the host."""
    tdt = {1: torch.uint8, 2: torch.int16, 4: torch.int32}[out_bytes]
    t = torch.from_numpy(idx.astype(np.int64)).cuda()
    v = d_cells.view(tdt)[t].cpu().numpy()
    mask = (1 << (8 * out_bytes)) - 1
    return v.astype(np.int64) & mask


def log(msg):
    print(msg, flush=True)


# ---- large drop-in (64-clade generator) --------------------------------------

def test_large_reduced_streamed_triangle_and_sparse(tmp_path):
    spec = workloads.large(n=3000)
    fa = str(tmp_path / "large.fa")
    workloads.write_fasta(spec, fa)
    chars = workloads.generate_chars(spec)
    packed = O.pack(chars)
    n = spec.n
    log(f"large reduced: n={n} L={spec.L}")
    # the reference's streaming path: dense cells, uint16 at min(max_distance, 65535)
    want16 = O.all_pairs_dense_simd(packed, 0, n, 2, 2)
    np.testing.assert_array_equal(want16, O.all_pairs_dense(packed, 2, 2))
    out = str(tmp_path / "lt.txt")
    hd.from_fasta_to_lower_triangular(fa, out, max_distance=2)
    ref_out = str(tmp_path / "lt_ref.txt")
    O.lib().oracle_write_lower_triangular(ref_out.encode(), want16.ctypes.data, n, 2)
    assert open(out, "rb").read() == open(ref_out, "rb").read()
    # the same triangle as a DataSet on both routes, then its sparse dumps
    # (the CPU route's sparse/dense heuristic picks dense on this data, so
    # both routes give the dense cells)
    assert not O.use_sparse(chars, False)
    want = O.all_pairs_dense_simd(packed, 0, n, 1, 2)
    for use_gpu in (True, False):
        ds = hd.from_fasta(fa, use_gpu=use_gpu, max_distance=2)
        np.testing.assert_array_equal(ds.lt_array, want, err_msg=f"use_gpu={use_gpu}")
        for thr in (3, 1, 0):
            got_file = str(tmp_path / f"sp{thr}.txt")
            ref_file = str(tmp_path / f"sp{thr}_ref.txt")
            ds.dump_sparse(got_file, thr)
            O.lib().oracle_write_sparse(ref_file.encode(), want.ctypes.data, n, 1, thr)
            assert open(got_file, "rb").read() == open(ref_file, "rb").read(), thr


def test_large_full_size_sampled_cells():
    spec = workloads.large()
    packed = workloads.generate_packed(spec)
    n = spec.n
    cells = n * (n - 1) // 2
    log(f"large full: n={n} cells={cells}")
    # every site compared (early exit off), uncapped uint16
    d16 = device_triangle(packed, capi.HD_SEM_DENSE, 2, 65535, False)
    idx = O.sample_cells(np.random.default_rng(50_000), n, SAMPLE)
    I, J = O.cell_rows(idx)
    got = gather(d16, 2, idx)
    want = np.array([O.distance_dense(packed[i], packed[j], 65535) for i, j in zip(I, J)])
    np.testing.assert_array_equal(got, want)
    # the product path: uint8 at max_distance 2 with the early exit on equals
    # the clamped uncapped triangle on every cell
    d8 = device_triangle(packed, capi.HD_SEM_DENSE, 1, 2, True)
    clamped = torch.clamp(d16.view(torch.int16).to(torch.int32) & 0xFFFF, max=2).to(torch.uint8)
    assert torch.equal(d8, clamped)
    del d16, clamped
    torch.cuda.empty_cache()
    # threshold 3 selects every cell: triplet k is cell k
    d_trip = torch.empty(cells * 3, dtype=torch.int32, device="cuda")
    total = capi.sparse_emit_dev(d8, n, 1, 3, d_trip, cells, stream())
    torch.cuda.synchronize()
    assert total == cells
    t = torch.from_numpy(idx.astype(np.int64)).cuda()
    trip = d_trip.view(-1, 3)[t].cpu().numpy().astype(np.int64)
    np.testing.assert_array_equal(trip[:, 0], I)
    np.testing.assert_array_equal(trip[:, 1], J)
    np.testing.assert_array_equal(trip[:, 2], np.minimum(want, 2))
    # threshold 1: exactly the cells <= 1, in row-major order
    total1 = capi.sparse_emit_dev(d8, n, 1, 1, d_trip, cells, stream())
    torch.cuda.synchronize()
    host8 = d8.cpu().numpy()
    sel = np.flatnonzero(host8 <= 1)
    assert total1 == sel.size
    trip1 = d_trip[: 3 * total1].view(-1, 3).cpu().numpy().astype(np.int64)
    I1, J1 = O.cell_rows(sel.astype(np.uint64))
    np.testing.assert_array_equal(trip1[:, 0], I1)
    np.testing.assert_array_equal(trip1[:, 1], J1)
    np.testing.assert_array_equal(trip1[:, 2], host8[sel])
    for k in range(0, total1, max(1, total1 // 64)):
        assert host8[sel[k]] == min(O.distance_dense(packed[I1[k]], packed[J1[k]], 2), 2)


def test_large_full_size_streamed_text(tmp_path):
    spec = workloads.large()
    fa = str(tmp_path / "large.fa")
    workloads.write_fasta(spec, fa)
    packed = workloads.generate_packed(spec)
    n = spec.n
    out = str(tmp_path / "lt.txt")
    log(f"large full streamed: n={n}")
    hd.from_fasta_to_lower_triangular(fa, out, max_distance=2)
    # every cell is one digit: row i (1..n-1) is exactly 2i bytes, at offset i(i-1)
    assert os.path.getsize(out) == n * (n - 1)
    rng = np.random.default_rng(50_001)
    rows = sorted({1, 2, 3, 64, 255, 256, 257, n // 2, n - 2, n - 1,
                   *rng.integers(1, n, size=40).tolist()})
    with open(out, "rb") as f:
        for i in rows:
            f.seek(i * (i - 1))
            line = f.read(2 * i)
            want = O.all_pairs_dense_simd(packed, i, i + 1, 1, 2)
            text = ",".join(str(v) for v in want).encode() + b"\n"
            assert line == text, f"row {i}"
    os.remove(out)
    os.remove(fa)


# ---- wide distances (include_x, sparse route, uint16) ------------------------

def wide_reference():
    from hammingdist_amd import _native

    spec = workloads.wide()
    return _native._synth_mutant(spec.L, 1012, 0.05, 0.0, 0.001, 0)


def test_wide_reduced_every_cell(tmp_path):
    spec = workloads.wide(n=3000)
    fa = str(tmp_path / "wide.fa")
    workloads.write_fasta(spec, fa)
    chars = workloads.generate_chars(spec)
    log(f"wide reduced: n={spec.n} L={spec.L}")
    ds = hd.from_fasta_large(fa, include_x=True)
    want = O.all_pairs_sparse(chars, True, 2, 65535)
    np.testing.assert_array_equal(ds.lt_array, want)
    # the routes of distances<T> agree with the oracle's (sparse: include_x)
    cells, route = O.distances(chars[:300], True, False, 2, 65535)
    assert route == 1
    np.testing.assert_array_equal(hd.from_fasta_large(fa, include_x=True, n=300).lt_array,
                                  cells)
    ref = wide_reference()
    got = hd.fasta_reference_distances(ref, fa, include_x=True)
    assert got.dtype == np.uint32 and got.shape == (spec.n,)
    want_rd = [O.distance_char(ref, bytes(r).decode("latin-1"), True) for r in chars]
    np.testing.assert_array_equal(got, np.array(want_rd, np.uint32))


def test_wide_full_size_sampled_cells(tmp_path):
    spec = workloads.wide()
    fa = str(tmp_path / "wide.fa")
    workloads.write_fasta(spec, fa)
    chars = workloads.generate_chars(spec)
    n = spec.n
    cells = n * (n - 1) // 2
    assert cells > 2**32
    log(f"wide full: n={n} L={spec.L} cells={cells}")
    ds = hd.from_fasta_large(fa, include_x=True)
    assert ds.nsamples == n
    cons = O.consensus(chars, True)
    idx = O.sample_cells(np.random.default_rng(100_000), n, SAMPLE)
    I, J = O.cell_rows(idx)
    for i, j in zip(I.tolist(), J.tolist()):
        want = O.distance_sparse_pair(chars[i], chars[j], cons, True, 65535)
        assert ds[i, j] == want and ds[j, i] == want, (i, j)
    del ds
    ref = wide_reference()
    got = hd.fasta_reference_distances(ref, fa, include_x=True)
    assert got.shape == (n,)
    rows = np.unique(np.concatenate([np.random.default_rng(100_001).integers(0, n, SAMPLE),
                                     [0, 1, n - 2, n - 1]]))
    for r in rows.tolist():
        assert got[r] == O.distance_char(ref, bytes(chars[r]).decode("latin-1"), True), r
    os.remove(fa)


# ---- a triangle past 2^32 cells, compared on every cell ----------------------

def short_rows(n, L, seed):
    rng = np.random.default_rng(seed)
    chars = np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, size=(n, L))].copy()
    # some '-' and 'N' (the dense code's per-row corrections), and near-copies
    # of earlier rows spread over the whole triangle (small distances late in it)
    sel = rng.integers(0, n, size=n // 50)
    chars[sel, rng.integers(0, L, size=sel.size)] = ord("-")
    sel = rng.integers(0, n, size=n // 50)
    chars[sel, rng.integers(0, L, size=sel.size)] = ord("N")
    dst = rng.integers(1, n, size=n // 20)
    src = (dst * rng.random(dst.size)).astype(np.int64)
    chars[dst] = chars[src]
    chars[dst, rng.integers(0, L, size=dst.size)] = ord("A")
    return chars


def test_short_rows_past_2_32_every_cell():
    n, L = 100_000, 16
    chars = short_rows(n, L, 2**32)
    packed = O.pack(chars)
    cells = n * (n - 1) // 2
    log(f"short rows: n={n} L={L} cells={cells}")
    want = O.all_pairs_dense(packed, 1, 255)
    d8 = device_triangle(packed, capi.HD_SEM_DENSE, 1, 255, False)
    assert np.array_equal(d8.cpu().numpy(), want)
    # every tile shape, early exit on and off, gives the same triangle
    for variant in range(len(capi.variants())):
        for early in (False, True):
            other = device_triangle(packed, capi.HD_SEM_DENSE, 1, 255, early, variant)
            assert torch.equal(other, d8), (variant, early)
            del other
    # sparse emit over the whole triangle: exactly the cells under the threshold
    for thr in (2, 3):
        sel = np.flatnonzero(want <= thr)
        cap = sel.size + 16
        d_trip = torch.empty(cap * 3, dtype=torch.int32, device="cuda")
        total = capi.sparse_emit_dev(d8, n, 1, thr, d_trip, cap, stream())
        torch.cuda.synchronize()
        assert total == sel.size
        trip = d_trip[: 3 * total].view(-1, 3).cpu().numpy().astype(np.int64)
        I, J = O.cell_rows(sel.astype(np.uint64))
        np.testing.assert_array_equal(trip[:, 0], I)
        np.testing.assert_array_equal(trip[:, 1], J)
        np.testing.assert_array_equal(trip[:, 2], want[sel])
        assert sel[-1] > 2**32, "the selection reaches past cell 2^32"


def test_short_rows_past_2_32_sparse_sampled():
    n, L = 100_000, 16
    chars = short_rows(n, L, 2**33)
    chars[::97, 3] = ord("X")
    log(f"short rows sparse-X: n={n} L={L}")
    packed = O.pack(chars, include_x=True)
    d16 = device_triangle(packed, capi.HD_SEM_SPARSE_X, 2, 65535, False)
    cons = O.consensus(chars, True)
    idx = O.sample_cells(np.random.default_rng(7), n, SAMPLE)
    I, J = O.cell_rows(idx)
    got = gather(d16, 2, idx)
    want = [O.distance_sparse_pair(chars[i], chars[j], cons, True, 65535) for i, j in zip(I, J)]
    np.testing.assert_array_equal(got, np.array(want))
    d16e = device_triangle(packed, capi.HD_SEM_SPARSE_X, 2, 65535, True)
    assert torch.equal(d16e, d16)


# ---- batched: 256 uneven datasets, every job against the oracle --------------

def test_batched_workload_every_job(tmp_path, monkeypatch):
    # 256 datasets of the plain generator, n log-uniform in [50, 5,000], L
    # uniform in [1,000, 30,000], alternating uint8 (cap 255) and uint16 (cap
    # 65,535), each on the reference CPU route's semantics (the oracle's
    # sparse/dense heuristic); one tuned, graph-captured launch and the
    # sequential plain launches both give the oracle's triangle for every job
    from hammingdist_amd import autotune

    monkeypatch.setenv("HD_TUNE_CACHE", str(tmp_path / "tune.json"))
    specs = workloads.batched()
    b = capi.Batch()
    jobs, wants = [], []
    for spec, ob in specs:
        chars = workloads.generate_chars(spec)
        packed = O.pack(chars)
        n, bps = packed.shape
        md = 255 if ob == 1 else 65535
        sparse = O.use_sparse(chars, False)
        # the engine's consensus count gives the same route
        n_diff, _ = capi.consensus_diff(packed, n, bps, spec.L)
        assert (n_diff / (n * spec.L) < 0.005) == sparse
        if sparse:
            want = O.all_pairs_sparse(chars, False, ob, md)
        else:
            want = O.all_pairs_dense_simd(packed, 0, n, ob, md)
        b.add(packed, capi.HD_SEM_SPARSE if sparse else capi.HD_SEM_DENSE, ob, md)
        jobs.append((n, bps, ob))
        wants.append(want)
    log(f"batched: {len(jobs)} datasets, {sum(w.size for w in wants)} cells")
    sh = stream()
    variant, _ = autotune.tune_batch(b, jobs, sh, torch.cuda.synchronize, reps=1)
    for run in (b.run, b.run_sequential):
        run(sh)
        torch.cuda.synchronize()
        for k, want in enumerate(wants):
            np.testing.assert_array_equal(b.output(k), want,
                                          err_msg=f"job {k} variant {variant} {run.__name__}")
    b.close()


# ---- serving: the 1,000,000-row database resident in HBM ---------------------

def test_serving_workload_full_database():
    # the serving database at its real size (1,000,000 x 29,903 sites, 15 GB)
    # and the bench's request mix (90 % queries against 10^4-10^6 random rows,
    # 10 % all-pairs jobs of 100-3,000 rows): every result the same alone and
    # in a concurrent mix, sampled rows and cells against the oracle, and the
    # last database rows reached
    from hammingdist_amd import serving

    spec = workloads.serving_db()
    db = workloads.generate_packed(spec)
    log(f"serving full: db {db.shape}")
    srv = capi.Server(db, device_bytes=24 << 30, host_bytes=4 << 30, streams=3, max_batch=32)
    reqs = serving.make_requests(db, spec.L, 60, seed=4004)
    tail = db.shape[0] - np.arange(1, 65, dtype=np.uint64)
    reqs.append(serving.Request("query", tail, db[5].copy()))
    alone = serving.run_alone(srv, reqs)
    mixed, _, _ = serving.closed_loop(srv, reqs)
    for a, b in zip(alone, mixed):
        np.testing.assert_array_equal(a, b)
    rng = np.random.default_rng(4005)
    for r, got in zip(reqs, alone):
        if r.kind == "query":
            assert got.shape == (r.rows.size,)
            idx = np.unique(np.concatenate([rng.integers(0, r.rows.size, 64),
                                            [0, r.rows.size - 1]]))
            for k in idx:
                assert got[k] == O.distance_dense(r.query, db[int(r.rows[k])]), (r.rows.size, k)
        else:
            n = r.rows.size
            assert got.shape == (n * (n - 1) // 2,)
            x = O.sample_cells(rng, n, 64)
            I, J = O.cell_rows(x)
            for xi, i, j in zip(x, I, J):
                want = min(O.distance_dense(db[int(r.rows[i])], db[int(r.rows[j])]), r.max_dist)
                assert got[int(xi)] == want, (n, i, j)
    srv.close()
