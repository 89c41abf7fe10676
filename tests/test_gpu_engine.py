"""Parity of the HIP engine against the oracle, through the C ABI.

Every comparison is exact (integer cells).  Sizes follow the reference's
CPU==GPU tests (src/hamming_t.cc:346-451: lengths 1..1092, 2..703 samples
around the 127/128/255/256/257 boundaries, max_dist {0,1,2,3,89,497,9999999})
plus sizes that cross this engine's own tile boundaries (64-row groups,
128/256/512-row tiles, 32-site words).
"""
import os

import numpy as np
import pytest

import oracle_lib as O
from conftest import random_strings
from hammingdist_amd import capi

pytestmark = pytest.mark.gpu

DT = {1: np.uint8, 2: np.uint16, 4: np.uint32}


def engine_all_pairs(packed, semantics, out_bytes, max_dist, both_paths=True):
    """hd_all_pairs on host rows.  Datasets of up to 512 rows take the
    one-pair-per-wavefront kernel; with both_paths they also run with that
    path switched off (HD_SMALL_PAIRS_MAX_ROWS=0: the tile kernels, as for
    larger datasets), and the two triangles must be equal."""
    n, bps = packed.shape

    def run():
        out = np.zeros(n * (n - 1) // 2, DT[out_bytes])
        capi.all_pairs(np.ascontiguousarray(packed), n, bps, semantics, out_bytes, max_dist, out)
        return out

    got = run()
    if both_paths and n <= 512:
        old = os.environ.get("HD_SMALL_PAIRS_MAX_ROWS")
        os.environ["HD_SMALL_PAIRS_MAX_ROWS"] = "0"
        try:
            tiles = run()
        finally:
            if old is None:
                del os.environ["HD_SMALL_PAIRS_MAX_ROWS"]
            else:
                os.environ["HD_SMALL_PAIRS_MAX_ROWS"] = old
        np.testing.assert_array_equal(got, tiles, err_msg="one pair per wavefront vs tiles")
    return got


def test_device_present():
    assert capi.device_count() >= 1
    print("device arch:", capi.device_arch())


# Dense inputs of three kinds: letters and '-' only, with invalid characters
# (both take the 3-op code with per-row corrections), and packed with the
# include_x table so that X (code 3) forces the 4-plane nibble code.
DENSE_INPUTS = [("ACGT-", False), ("ACGT-N?", False), ("ACGTX-N", True)]


@pytest.mark.parametrize("alphabet,include_x", DENSE_INPUTS)
@pytest.mark.parametrize("n,L", [(2, 1), (3, 17), (7, 88), (11, 381), (127, 1023), (128, 33),
                                 (255, 64), (256, 95), (257, 1092), (513, 130), (703, 381),
                                 (1100, 700)])
def test_dense_all_pairs_matches_oracle(n, L, alphabet, include_x, rng):
    seqs = random_strings(rng, n, L, alphabet)
    packed = O.pack(O.as_chars(seqs), include_x)
    for out_bytes, max_dist in [(1, 255), (2, 65535), (4, 2**32 - 1), (1, 3), (2, 89),
                                (1, 0), (4, 497)]:
        want = O.all_pairs_dense(packed, out_bytes, max_dist)
        got = engine_all_pairs(packed, capi.HD_SEM_DENSE, out_bytes, max_dist)
        np.testing.assert_array_equal(got, want, err_msg=f"n={n} L={L} md={max_dist}")


def test_dense_raw_nibbles_match_oracle(rng):
    # Arbitrary nibble codes (not from the lookup table) take the 4-plane code.
    n, bps = 300, 136
    packed = rng.integers(0, 256, size=(n, bps), dtype=np.uint8)
    for out_bytes, md in [(2, 65535), (1, 9)]:
        want = O.all_pairs_dense(packed, out_bytes, md)
        np.testing.assert_array_equal(
            engine_all_pairs(packed, capi.HD_SEM_DENSE, out_bytes, md), want)


@pytest.mark.parametrize("L", [1, 31, 32, 33, 100, 1025])
def test_dense_invalid_and_gap_rows(L):
    # Whole rows of invalid characters and of '-': the per-row corrections
    # (invalid count of the row, '-' count of the column, padding included)
    # at their extremes, around every clamp.
    seqs = ["N" * L, "-" * L, "A" * L, "N" * L, "-" * L, "G" * L, "A" * (L - 1) + "N"]
    packed = O.pack(O.as_chars(seqs))
    for md in sorted({0, 1, max(L - 1, 0), L, L + 1, 2**31 - 1}):
        want = O.all_pairs_dense(packed, 4, md)
        got = engine_all_pairs(packed, capi.HD_SEM_DENSE, 4, md)
        np.testing.assert_array_equal(got, want, err_msg=f"L={L} md={md}")


def test_dense_early_exit_with_late_gaps_and_invalid(rng):
    # Several 1024-site chunks; gap runs and invalid runs at varying places,
    # few substitutions: distances straddle small caps, so the early-exit
    # votes (taken on exact partial distances) decide for some tiles only.
    n, L = 400, 6000
    root = rng.choice(list("ACGT"), size=L)
    seqs = []
    for _ in range(n):
        s = root.copy()
        s[rng.integers(0, L, size=rng.integers(0, 6))] = rng.choice(list("ACGT"))
        a = int(rng.integers(0, L))
        s[a:a + int(rng.integers(0, 900))] = "-"
        b = int(rng.integers(0, L))
        s[b:b + int(rng.integers(0, 300))] = "N"
        seqs.append("".join(s))
    packed = O.pack(O.as_chars(seqs))
    for out_bytes, md in [(1, 0), (1, 1), (1, 2), (1, 5), (2, 300), (2, 1200)]:
        want = O.all_pairs_dense(packed, out_bytes, md)
        got = engine_all_pairs(packed, capi.HD_SEM_DENSE, out_bytes, md)
        np.testing.assert_array_equal(got, want, err_msg=f"md={md}")


@pytest.mark.parametrize("alphabet", ["ACGT-NX", "ACGT-", "AAAAAAAACGT-N", "ACGTXn*-?"])
@pytest.mark.parametrize("n,L", [(50, 25), (300, 513), (700, 97)])
def test_sparse_semantics_match_oracle(alphabet, n, L, rng):
    seqs = random_strings(rng, n, L, alphabet)
    chars = O.as_chars(seqs)
    for include_x in (False, True):
        packed = O.pack(chars, include_x)
        sem = capi.HD_SEM_SPARSE_X if include_x else capi.HD_SEM_SPARSE
        for out_bytes, md in [(1, 255), (2, 65535), (1, 5)]:
            want = O.all_pairs_sparse(chars, include_x, out_bytes, md)
            got = engine_all_pairs(packed, sem, out_bytes, md)
            np.testing.assert_array_equal(got, want,
                                          err_msg=f"{alphabet} x={include_x} md={md}")


@pytest.mark.parametrize("alphabet", ["ACGT-NX", "ACGTXn*-?", "ACGT-"])
def test_char_x_all_pairs_match_distance(alphabet, rng):
    n, L = 60, 150
    seqs = random_strings(rng, n, L, alphabet)
    packed = O.pack(O.as_chars(seqs), include_x=True)
    got = engine_all_pairs(packed, capi.HD_SEM_CHAR_X, 4, 2**32 - 1)
    I, J = O.tri_index(n)
    want = np.array([O.distance_char(seqs[i], seqs[j], True) for i, j in zip(I, J)], np.uint32)
    np.testing.assert_array_equal(got, want)


def test_a_versus_g_closed_form():
    # src/hamming_impl_t.cc:28-47: n x 'A' against n x 'G' -> min(max_dist, n)
    for L in [1, 2, 31, 32, 33, 63, 64, 65, 255, 256, 257, 1023, 1024, 1025, 8193, 65537,
              131073, 1048577]:
        packed = O.pack(O.as_chars(["A" * L, "G" * L]))
        for md in [0, 1, 2, 11, 999, 9876544]:
            got = engine_all_pairs(packed, capi.HD_SEM_DENSE, 4, md)
            assert int(got[0]) == min(md, L), (L, md)


def test_identical_rows_are_zero(rng):
    # src/hamming_impl_t.cc:6-26 and the CUDA variant src/distance_cuda_t.cc:6-26
    for L in [1, 7, 8, 9, 16, 17, 4095, 4097, 32769]:
        s = random_strings(rng, 1, L, "ACGT-")[0]
        packed = O.pack(O.as_chars([s, s, s]))
        for sem in (capi.HD_SEM_DENSE, capi.HD_SEM_SPARSE):
            assert engine_all_pairs(packed, sem, 4, 2**31 - 1).tolist() == [0, 0, 0]


def test_distance_pair_matches_oracle(rng):
    # src/distance_cuda_t.cc:50-72: pair kernel == distance_cpp on random rows
    for L in [1, 5, 17, 64, 65, 1000, 4097]:
        a, b = random_strings(rng, 2, L, "ACGT-")
        pa, pb = O.pack(O.as_chars([a])), O.pack(O.as_chars([b]))
        for md in [0, 1, 2, 11, 999, 9876544]:
            got = capi.distance_pair(pa[0], pb[0], pa.shape[1], capi.HD_SEM_DENSE, md)
            assert got == O.distance_dense(pa[0], pb[0], md)


# (2, 65535): no early-exit votes (the cap is out of reach of 200 sites);
# the others vote, and their bands start at a cell offset inside the
# triangle, stored through LDS: 128 x 32 tiles under a cap below 64,
# 256 x 64 above.
@pytest.mark.parametrize("out_bytes,md", [(2, 65535), (1, 3), (1, 100), (2, 150)])
def test_stream_bands_concatenate_to_triangle(rng, out_bytes, md):
    import ctypes

    n, L = 1500, 200
    seqs = random_strings(rng, n, L, "ACGT-N")
    packed = np.ascontiguousarray(O.pack(O.as_chars(seqs)))
    want = O.all_pairs_dense(packed, out_bytes, md)
    got = np.zeros_like(want)
    seen = []
    ctype = ctypes.c_uint16 if out_bytes == 2 else ctypes.c_uint8

    def cb(addr, first, count, r0, r1):
        buf = (ctype * count).from_address(addr)
        got[first:first + count] = np.frombuffer(buf, want.dtype)
        assert first == r0 * (r0 - 1) // 2
        assert first + count == r1 * (r1 - 1) // 2
        seen.append((r0, r1))

    capi.all_pairs_stream(packed, n, packed.shape[1], capi.HD_SEM_DENSE, out_bytes, md, 100_000,
                          cb)
    assert len(seen) > 3
    assert seen[0][0] == 1 and seen[-1][1] == n
    np.testing.assert_array_equal(got, want)


def test_reference_distances_match_oracle(rng):
    m, L = 400, 1234
    for alphabet, include_x in [("ACGT-", False), ("ACGT-NX", False), ("ACGT-NX", True),
                                ("ACGTXn*-?", True)]:
        seqs = random_strings(rng, m + 1, L, alphabet)
        ref, rows = seqs[0], seqs[1:]
        db = np.ascontiguousarray(O.pack(O.as_chars(rows), include_x))
        q = np.ascontiguousarray(O.pack(O.as_chars([ref]), include_x)[0])
        sem = capi.HD_SEM_CHAR_X if include_x else capi.HD_SEM_DENSE
        out = np.zeros(m, np.uint32)
        capi.reference_distances(q, db, m, db.shape[1], None, m, sem, out)
        want = [O.distance_char(ref, s, include_x) for s in rows]
        assert out.tolist() == want
        # gathered subset, with repeats
        sel = rng.integers(0, m, size=257).astype(np.uint64)
        out2 = np.zeros(sel.size, np.uint32)
        capi.reference_distances(q, db, m, db.shape[1], sel, sel.size, sem, out2)
        assert out2.tolist() == [want[int(k)] for k in sel]


@pytest.mark.parametrize("units", [1, 63, 64, 65, 256, 257, 512, 513, 1024, 1025, 2048, 2049,
                                   3001])
def test_reference_distances_row_lengths(rng, units):
    # Row lengths around the register-resident query sizes (64 T units of
    # 8 bytes for T = 4, 8, 16, 32) and past them (query read from memory).
    m, bps = 300, 8 * units
    db = rng.integers(0, 256, size=(m, bps), dtype=np.uint8)
    q = rng.integers(0, 256, size=bps, dtype=np.uint8)
    db[::7] = q  # some exact matches (distance 0)
    out = np.zeros(m, np.uint32)
    capi.reference_distances(q, db, m, bps, None, m, capi.HD_SEM_DENSE, out)
    assert out.tolist() == [O.distance_dense(q, db[k]) for k in range(m)]


@pytest.mark.parametrize("L", [4095, 4096, 4097, 32767, 32768, 32769])
def test_reference_distances_char_x_lengths(rng, L):
    m = 70
    seqs = random_strings(rng, m + 1, L, "ACGTXN-")
    ref, rows = seqs[0], seqs[1:]
    db = np.ascontiguousarray(O.pack(O.as_chars(rows), True))
    q = np.ascontiguousarray(O.pack(O.as_chars([ref]), True)[0])
    out = np.zeros(m, np.uint32)
    capi.reference_distances(q, db, m, db.shape[1], None, m, capi.HD_SEM_CHAR_X, out)
    assert out.tolist() == [O.distance_char(ref, s, True) for s in rows]


def test_consensus_diff_matches_oracle(rng):
    for alphabet in ["ACGT-", "AAAAAAAAAAAAAAAACGTN-", "ACGTX-"]:
        n, L = 333, 1001
        seqs = random_strings(rng, n, L, alphabet)
        chars = O.as_chars(seqs)
        packed = np.ascontiguousarray(O.pack(chars))
        n_diff, mask = capi.consensus_diff(packed, n, packed.shape[1], L)
        cons = np.frombuffer(O.consensus(chars, False), np.uint8)
        assert n_diff == int((chars != cons[None, :]).sum())
        codes = set()
        for ch in set(alphabet):
            v = {"A": 1, "C": 2, "G": 4, "T": 8, "-": 15}.get(ch, 0)
            codes.add(v)
        assert mask == sum(1 << c for c in codes)


@pytest.mark.parametrize("out_bytes", [1, 2, 4])
def test_sparse_emit_row_major(out_bytes, rng):
    n = 900
    cells = rng.integers(0, 40, size=n * (n - 1) // 2).astype(DT[out_bytes])
    I, J = O.tri_index(n)
    for thr in [-1, 0, 3, 39, 1000]:
        sel = cells <= thr
        count = capi.sparse_emit(cells, n, out_bytes, thr, None, 0)
        assert count == int(sel.sum())
        if count:
            trip = np.zeros((count, 3), np.uint32)
            capi.sparse_emit(cells, n, out_bytes, thr, trip, count)
            np.testing.assert_array_equal(trip[:, 0], I[sel])
            np.testing.assert_array_equal(trip[:, 1], J[sel])
            np.testing.assert_array_equal(trip[:, 2], cells[sel])


@pytest.mark.parametrize("out_bytes", [1, 2])
def test_large_triangle_download_in_chunks(out_bytes, rng):
    # 12,000 rows: 72e6 cells (69 MB as uint8, 137 MB as uint16) come back
    # to the host through the chunked staging copy (64 MiB chunks).
    n, L = 12_000, 60
    seqs = random_strings(rng, n, L, "ACGT-N")
    packed = O.pack(O.as_chars(seqs))
    md = 255 if out_bytes == 1 else 65535
    got = engine_all_pairs(packed, capi.HD_SEM_DENSE, out_bytes, md)
    want = O.all_pairs_dense(packed, out_bytes, md)
    np.testing.assert_array_equal(got, want)


def test_concurrent_small_calls_match_oracle():
    # Many short host calls from four threads at once: device blocks come
    # from and go back to the library's pool between calls, datasets of up to
    # 512 rows take the one-pair-per-wavefront kernel and larger ones the
    # tile kernels.  Every result equals the oracle's.
    from concurrent.futures import ThreadPoolExecutor

    rng = np.random.default_rng(606)
    jobs = []
    for k in range(48):
        n = int(rng.choice([2, 3, 17, 64, 130, 511, 512, 513, 700]))
        L = int(rng.choice([1, 31, 100, 999, 3000]))
        alphabet = ["ACGT-N", "ACGT-", "ACGTX-N?"][k % 3]
        chars = O.as_chars(random_strings(rng, n, L, alphabet))
        sem = [capi.HD_SEM_DENSE, capi.HD_SEM_SPARSE, capi.HD_SEM_SPARSE_X][k % 3]
        include_x = sem == capi.HD_SEM_SPARSE_X
        packed = O.pack(chars, include_x)
        ob, md = [(1, 255), (2, 65535), (4, 7), (1, 2)][k % 4]
        if sem == capi.HD_SEM_DENSE:
            want = O.all_pairs_dense(packed, ob, md)
        else:
            want = O.all_pairs_sparse(chars, include_x, ob, md)
        jobs.append((packed, sem, ob, md, want))

    def run(job):
        packed, sem, ob, md, _ = job
        return engine_all_pairs(packed, sem, ob, md, both_paths=False)

    with ThreadPoolExecutor(max_workers=4) as ex:
        got = list(ex.map(run, jobs + jobs))
    for k, g in enumerate(got):
        np.testing.assert_array_equal(g, jobs[k % len(jobs)][4], err_msg=f"job {k % len(jobs)}")


def test_concurrent_host_calls_match_sequential(rng):
    # Host threads calling the library at once (ctypes releases the GIL):
    # the calls share the library streams and the pinned staging buffers
    # (inputs and triangles over 64 MiB go through them).  Every result must
    # equal the same call made alone; the small job is also checked against
    # the oracle.
    from concurrent.futures import ThreadPoolExecutor

    table = np.frombuffer(b"ACGT-N", np.uint8)
    jobs = []
    for n, L, out_bytes, md in [(5000, 30000, 1, 255), (6000, 2000, 4, 2**32 - 1),
                                (3000, 12000, 2, 65535), (700, 1000, 1, 3), (2500, 29903, 1, 2)]:
        chars = table[rng.integers(0, table.size, size=(n, L), dtype=np.uint8)]
        # a few near-duplicates so that the small caps are not all passed at once
        chars[1::7] = chars[0::7][: chars[1::7].shape[0]]
        jobs.append((O.pack(chars), out_bytes, md))
    alone = [engine_all_pairs(p, capi.HD_SEM_DENSE, ob, md, both_paths=False)
             for p, ob, md in jobs]
    small = jobs[3]
    np.testing.assert_array_equal(alone[3], O.all_pairs_dense(small[0], small[1], small[2]))
    with ThreadPoolExecutor(max_workers=4) as ex:
        futures = [ex.submit(engine_all_pairs, p, capi.HD_SEM_DENSE, ob, md, False)
                   for _ in range(2) for p, ob, md in jobs]
        got = [f.result() for f in futures]
    for k, g in enumerate(got):
        np.testing.assert_array_equal(g, alone[k % len(jobs)], err_msg=f"job {k % len(jobs)}")
