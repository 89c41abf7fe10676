"""Device-resident entry points (torch device memory, torch's stream): every
tile variant, early exit on and off, and every output width give the same
triangle as the oracle."""
import numpy as np
import pytest

import oracle_lib as O
from conftest import random_strings
from hammingdist_amd import capi

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

DT = {1: np.uint8, 2: np.uint16, 4: np.uint32}


def run_dev(packed, sem, out_bytes, max_dist, early, variant):
    n, bps = packed.shape
    stream = torch.cuda.current_stream().cuda_stream
    d_packed = torch.from_numpy(np.ascontiguousarray(packed)).cuda()
    mask = capi.nibble_mask_dev(d_packed, n, bps, stream)
    _, _, _, total = capi.planes_layout(n, bps, sem, mask)
    d_planes = torch.empty(total, dtype=torch.int32, device="cuda")
    capi.encode_dev(d_packed, n, bps, sem, mask, d_planes, stream)
    d_out = torch.zeros(n * (n - 1) // 2 * out_bytes, dtype=torch.uint8, device="cuda")
    capi.all_pairs_dev(d_planes, n, bps, sem, mask, out_bytes, max_dist, early, variant, d_out,
                       stream)
    torch.cuda.synchronize()
    return d_out.cpu().numpy().view(DT[out_bytes])


@pytest.mark.parametrize("variant", range(5))
@pytest.mark.parametrize("alphabet,include_x", [("ACGT-N", False), ("ACGTX-N", True)])
def test_every_variant_matches_oracle(variant, alphabet, include_x, rng):
    # ACGT-N: the dense 3-op code; X packed as code 3: the 4-plane code.
    for n, L in [(2, 3), (65, 64), (600, 300), (1100, 97)]:
        seqs = random_strings(rng, n, L, alphabet)
        packed = O.pack(O.as_chars(seqs), include_x)
        # (4, 9): 32-bit cells under early-exit votes, stored through LDS
        # one row group of 64 at a time
        for out_bytes, md in [(1, 255), (2, 7), (4, 2**32 - 1), (4, 9)]:
            want = O.all_pairs_dense(packed, out_bytes, md)
            for early in (False, True):
                got = run_dev(packed, capi.HD_SEM_DENSE, out_bytes, md, early, variant)
                np.testing.assert_array_equal(got, want, err_msg=f"v={variant} n={n} L={L}")


@pytest.mark.parametrize("variant", range(5))
def test_every_variant_sparse_x(variant, rng):
    n, L = 700, 150
    seqs = random_strings(rng, n, L, "ACGTX-N")
    chars = O.as_chars(seqs)
    packed = O.pack(chars, include_x=True)
    want = O.all_pairs_sparse(chars, True, 2, 65535)
    got = run_dev(packed, capi.HD_SEM_SPARSE_X, 2, 65535, True, variant)
    np.testing.assert_array_equal(got, want)


def gap_heavy_rows(rng, n, L, alphabet):
    """Rows over `alphabet` with '-' in runs (up to a third of the row), some
    rows all '-', some with a '-' in every 32-site word, some without any."""
    seqs = []
    for r in range(n):
        s = rng.choice(list(alphabet), size=L)
        kind = r % 7
        if kind == 0:
            s[:] = "-"
        elif kind == 1:
            s[rng.integers(0, 32, size=1)[0]::32] = "-"
        elif kind in (2, 3, 4):
            for _ in range(int(rng.integers(1, 4))):
                a = int(rng.integers(0, L))
                s[a:a + int(rng.integers(1, max(2, L // 3)))] = "-"
        seqs.append("".join(s))
    return seqs


GAP_CASES = [(600, 1290, "ACGTNX"), (300, 2100, "ACG"), (70, 33, "AC")]
GAP_CAPS = [(2, 65535), (1, 3), (4, 2**32 - 1), (2, 40), (1, 0), (4, 50)]
_gap_oracle = {}


def gap_case(case, include_x):
    """Rows and oracle triangles of one gap case, shared by every variant."""
    key = (case, include_x)
    if key not in _gap_oracle:
        n, L, alphabet = GAP_CASES[case]
        chars = O.as_chars(gap_heavy_rows(np.random.default_rng(2024 + case), n, L, alphabet))
        want = {cap: O.all_pairs_sparse(chars, include_x, *cap) for cap in GAP_CAPS}
        _gap_oracle[key] = (O.pack(chars, include_x), want)
    return _gap_oracle[key]


@pytest.mark.parametrize("variant", range(5))
@pytest.mark.parametrize("include_x", [False, True])
def test_gap_code_every_variant_with_gap_runs(variant, include_x):
    # The sparse semantics' '-' correction: per-row '-' counts, the bitmap of
    # words holding a '-' and 2 |g_i & g_j| over them; L = 1290 leaves 16
    # padding sites in the last word, L = 2100 spans three early-exit chunks.
    sem = capi.HD_SEM_SPARSE_X if include_x else capi.HD_SEM_SPARSE
    for case, (n, L, _) in enumerate(GAP_CASES):
        packed, wants = gap_case(case, include_x)
        for out_bytes, md in GAP_CAPS:
            want = wants[(out_bytes, md)]
            for early in (False, True):
                got = run_dev(packed, sem, out_bytes, md, early, variant)
                np.testing.assert_array_equal(
                    got, want, err_msg=f"v={variant} n={n} L={L} md={md} early={early}")


@pytest.mark.parametrize("variant", [0, 2, 4])
def test_long_row_class_code(variant, monkeypatch):
    # kClass3 serves rows of 2^28 sites or more; HD_SPARSE_CLASS_CODE=1
    # selects it for these small rows.
    monkeypatch.setenv("HD_SPARSE_CLASS_CODE", "1")
    for include_x in (False, True):
        sem = capi.HD_SEM_SPARSE_X if include_x else capi.HD_SEM_SPARSE
        for case in range(len(GAP_CASES)):
            packed, wants = gap_case(case, include_x)
            n, bps = packed.shape
            assert capi.planes_layout(n, bps, sem, 0xFFFF)[3] == \
                capi.planes_layout(n, bps, sem, 0xFFFF)[0] * ((2 * bps + 31) // 32) * 4
            for (out_bytes, md) in [(2, 65535), (1, 3)]:
                for early in (False, True):
                    got = run_dev(packed, sem, out_bytes, md, early, variant)
                    np.testing.assert_array_equal(got, wants[(out_bytes, md)],
                                                  err_msg=f"case={case} md={md}")


def test_gap_code_early_exit_waits_for_shared_gaps(rng):
    # Rows share long '-' runs in the first chunks, so popcount - G_i - G_j
    # runs far below the partial distance there, and differ only later: the
    # votes use that lower bound, so no tile may stop before its distances
    # reach the cap.
    n, L = 500, 5000
    root = rng.choice(list("ACGT"), size=L)
    seqs = []
    for _ in range(n):
        s = root.copy()
        s[: int(rng.integers(1500, 2500))] = "-"
        s[int(rng.integers(2600, L - 10)) + rng.integers(0, 10, size=int(rng.integers(0, 6)))] = "A"
        seqs.append("".join(s))
    chars = O.as_chars(seqs)
    packed = O.pack(chars)
    for md in (0, 1, 2, 4, 9):
        want = O.all_pairs_sparse(chars, False, 1, md)
        got = run_dev(packed, capi.HD_SEM_SPARSE, 1, md, True, 0)
        np.testing.assert_array_equal(got, want, err_msg=f"md={md}")


def test_early_exit_identical_on_saturating_data():
    # Every pair far above max_dist: the early exit fires in the first chunk.
    rng = np.random.default_rng(7)
    n, L = 900, 5000
    seqs = random_strings(rng, n, L, "ACGT")
    packed = O.pack(O.as_chars(seqs))
    want = O.all_pairs_dense(packed, 1, 2)
    assert (want == 2).all()
    for v in range(5):
        got = run_dev(packed, capi.HD_SEM_DENSE, 1, 2, True, v)
        np.testing.assert_array_equal(got, want)


def invalid_count_rows(rng, n, L, blocks):
    """Rows of one root with few substitutions and '-' runs, and exactly z
    invalid sites ('N' or '?', scattered or in a run) per row, z taken per
    block of rows from `blocks` ((row count, z choices) pairs, cycled)."""
    root = rng.choice(list("ACGT"), size=L)
    seqs, k = [], 0
    while len(seqs) < n:
        count, zs = blocks[k % len(blocks)]
        k += 1
        for _ in range(min(count, n - len(seqs))):
            s = root.copy()
            s[rng.integers(0, L, size=rng.integers(0, 3))] = rng.choice(list("ACGT"))
            a = int(rng.integers(0, L))
            s[a:a + int(rng.integers(0, 200))] = "-"
            z = int(rng.choice(zs))
            if z and rng.random() < 0.3:
                b = int(rng.integers(0, L - z + 1))
                s[b:b + z] = "N"
            elif z:
                s[rng.choice(L, size=z, replace=False)] = rng.choice(list("N?"))
            seqs.append("".join(s))
    return seqs


@pytest.mark.parametrize("variant", range(5))
def test_invalid_site_bound_settles_exactly(variant):
    # With early-exit votes the dense 3-op kernel settles pairs whose bound
    # max(Z_i, Z_j) (invalid sites of either row) reaches the cap before its
    # loop.  Row blocks with every row's Z >= the cap, with Z just under it,
    # with no invalid site and with mixed counts put waves of every tile
    # shape fully settled, partly settled and open; a pair of rows under the
    # cap whose invalid sites share positions has a distance below Z_i + Z_j.
    rng = np.random.default_rng(4242)
    n, L = 1300, 2500
    blocks = [(300, [2, 3, 40]), (70, [0]), (200, [1]), (130, [0, 1, 2, 3, 4, 5]),
              (256, [6, 9]), (64, [0, 2])]
    seqs = invalid_count_rows(rng, n, L, blocks)
    packed = O.pack(O.as_chars(seqs))
    for out_bytes, md in [(1, 1), (1, 2), (1, 3), (2, 5), (4, 7), (1, 40)]:
        want = O.all_pairs_dense(packed, out_bytes, md)
        got = run_dev(packed, capi.HD_SEM_DENSE, out_bytes, md, True, variant)
        np.testing.assert_array_equal(got, want, err_msg=f"v={variant} md={md}")


@pytest.mark.parametrize("variant", [0, 4])
def test_invalid_site_fill_at_shifted_outputs_and_in_bands(variant):
    # Row blocks whose rows all reach the cap are filled with 16-byte stores
    # from the first 16-byte boundary of the output on: outputs that start 1
    # cell past an aligned address, and row bands (band-relative cells), move
    # every block's run against that boundary.  The guard cells around the
    # output keep their pattern.
    rng = np.random.default_rng(77)
    n, L = 1400, 1500
    blocks = [(512, [3, 8]), (256, [0, 1]), (600, [2, 30])]
    packed = O.pack(O.as_chars(invalid_count_rows(rng, n, L, blocks)))
    stream = torch.cuda.current_stream().cuda_stream
    d_packed = torch.from_numpy(np.ascontiguousarray(packed)).cuda()
    mask = capi.nibble_mask_dev(d_packed, n, packed.shape[1], stream)
    _, _, _, total = capi.planes_layout(n, packed.shape[1], capi.HD_SEM_DENSE, mask)
    d_planes = torch.empty(total, dtype=torch.int32, device="cuda")
    capi.encode_dev(d_packed, n, packed.shape[1], capi.HD_SEM_DENSE, mask, d_planes, stream)
    cells = n * (n - 1) // 2
    TI = 64 * capi.variants()[variant]["rows_per_lane"]
    for out_bytes, md in [(1, 2), (2, 3), (4, 2)]:
        want = O.all_pairs_dense(packed, out_bytes, md)
        d_out = torch.full(((cells + 2) * out_bytes,), 0xA5, dtype=torch.uint8, device="cuda")
        capi.all_pairs_dev(d_planes, n, packed.shape[1], capi.HD_SEM_DENSE, mask, out_bytes,
                           md, True, variant, d_out[out_bytes:], stream)
        torch.cuda.synchronize()
        got = d_out.cpu().numpy()
        assert (got[:out_bytes] == 0xA5).all() and (got[-out_bytes:] == 0xA5).all()
        np.testing.assert_array_equal(got[out_bytes:-out_bytes].view(DT[out_bytes]), want,
                                      err_msg=f"v={variant} out_bytes={out_bytes} md={md}")
        for r0, r1 in [(TI, 3 * TI), (2 * TI, n)]:
            c0, c1 = r0 * (r0 - 1) // 2, r1 * (r1 - 1) // 2
            d_band = torch.full(((c1 - c0 + 2) * out_bytes,), 0xA5, dtype=torch.uint8,
                                device="cuda")
            capi.all_pairs_dev_rows(d_planes, n, packed.shape[1], capi.HD_SEM_DENSE, mask,
                                    out_bytes, md, True, variant, r0, r1, d_band[out_bytes:],
                                    stream)
            torch.cuda.synchronize()
            got = d_band.cpu().numpy()
            assert (got[:out_bytes] == 0xA5).all() and (got[-out_bytes:] == 0xA5).all()
            np.testing.assert_array_equal(got[out_bytes:-out_bytes].view(DT[out_bytes]),
                                          want[c0:c1], err_msg=f"band {r0}-{r1}")


def test_reference_distances_dev_matches_host(rng):
    m, L = 1000, 3000
    seqs = random_strings(rng, m + 1, L, "ACGT-N")
    db = O.pack(O.as_chars(seqs[1:]))
    q = O.pack(O.as_chars(seqs[:1]))[0]
    d_db = torch.from_numpy(db).cuda()
    d_q = torch.from_numpy(q).cuda()
    d_out = torch.zeros(m, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    capi.reference_distances_dev(d_q, d_db, m, db.shape[1], None, m, capi.HD_SEM_DENSE, d_out,
                                 stream)
    torch.cuda.synchronize()
    want = [O.distance_char(seqs[0], s) for s in seqs[1:]]
    assert d_out.cpu().numpy().astype(np.uint32).tolist() == want


@pytest.mark.parametrize("out_bytes", [1, 2, 4])
def test_sparse_emit_dev_on_resident_triangle(out_bytes, rng):
    # all-pairs into HBM, then the (i, j, d) emit of cells <= threshold from
    # the resident triangle, row-major (include/hamming/hamming.hh:79-114)
    n, L = 1300, 400
    seqs = random_strings(rng, n, L, "ACGT-N")
    base = seqs[0]
    seqs = [base if k % 5 == 0 else s for k, s in enumerate(seqs)]  # some exact duplicates
    packed = O.pack(O.as_chars(seqs))
    md = {1: 255, 2: 65535, 4: 2**32 - 1}[out_bytes]
    want_cells = O.all_pairs_dense(packed, out_bytes, md)
    n_, bps = packed.shape
    stream = torch.cuda.current_stream().cuda_stream
    d_packed = torch.from_numpy(np.ascontiguousarray(packed)).cuda()
    mask = capi.nibble_mask_dev(d_packed, n, bps, stream)
    _, _, _, total = capi.planes_layout(n, bps, capi.HD_SEM_DENSE, mask)
    d_planes = torch.empty(total, dtype=torch.int32, device="cuda")
    capi.encode_dev(d_packed, n, bps, capi.HD_SEM_DENSE, mask, d_planes, stream)
    d_cells = torch.zeros(want_cells.size * out_bytes, dtype=torch.uint8, device="cuda")
    capi.all_pairs_dev(d_planes, n, bps, capi.HD_SEM_DENSE, mask, out_bytes, md, True, -1,
                       d_cells, stream)
    I, J = O.tri_index(n)
    for thr in [0, 3, 150]:
        sel = want_cells <= thr
        count = capi.sparse_emit_dev(d_cells, n, out_bytes, thr, None, 0, stream)
        assert count == int(sel.sum())
        d_trip = torch.zeros(max(count, 1) * 3, dtype=torch.int32, device="cuda")
        assert capi.sparse_emit_dev(d_cells, n, out_bytes, thr, d_trip, count, stream) == count
        torch.cuda.synchronize()
        trip = d_trip[:count * 3].cpu().numpy().view(np.uint32).reshape(-1, 3)
        np.testing.assert_array_equal(trip[:, 0], I[sel])
        np.testing.assert_array_equal(trip[:, 1], J[sel])
        np.testing.assert_array_equal(trip[:, 2], want_cells[sel])
        # an output 4 bytes past a 16-byte boundary (word stores instead of
        # 16-byte stores) holds the same triplets
        d_off = torch.zeros(max(count, 1) * 3 + 1, dtype=torch.int32, device="cuda")
        assert capi.sparse_emit_dev(d_cells, n, out_bytes, thr, d_off[1:], count, stream) == count
        torch.cuda.synchronize()
        np.testing.assert_array_equal(d_off[1:count * 3 + 1].cpu().numpy().view(np.uint32),
                                      trip.reshape(-1))
        assert int(d_off[0]) == 0
        if count > 1:
            with pytest.raises(capi.EngineError, match="capacity"):
                capi.sparse_emit_dev(d_cells, n, out_bytes, thr, d_trip, count - 1, stream)
        if count > 100:
            # a capacity inside a workgroup's run: the first `cap` triplets
            # are written and nothing past them
            cap = count // 2 + 7
            d_cap = torch.full((count * 3,), 0x7F7F7F7F, dtype=torch.int32, device="cuda")
            with pytest.raises(capi.EngineError, match="capacity"):
                capi.sparse_emit_dev(d_cells, n, out_bytes, thr, d_cap, cap, stream)
            torch.cuda.synchronize()
            got_cap = d_cap.cpu().numpy().view(np.uint32)
            np.testing.assert_array_equal(got_cap[:cap * 3], trip.reshape(-1)[:cap * 3])
            assert (got_cap[cap * 3:] == 0x7F7F7F7F).all()


_long_rows = {}


def long_rows(L):
    """Three rows of L sites: substitutions, '-' runs, 'N' (shared by variants)."""
    if L not in _long_rows:
        rng = np.random.default_rng(5)
        base = rng.choice(np.frombuffer(b"ACGT", np.uint8), size=L)
        rows = np.tile(base, (3, 1))
        for k in range(3):
            rows[k, rng.integers(0, L, size=1000 * (k + 1))] = \
                rng.choice(np.frombuffer(b"ACGT-N", np.uint8))
            rows[k, L // 2 + 7 * k: L // 2 + 7 * k + 5000] = ord("-")
        _long_rows[L] = (rows, O.pack(rows))
    return _long_rows[L]


@pytest.mark.parametrize("variant", [0, 1])
def test_longest_rows_of_the_offset_codes(variant):
    # Rows of 2^18 - 1 words (8,388,576 sites) are the longest the dense
    # 3-op and sparse gap codes take: their tiles address rows through 31-bit
    # buffer offsets, which for 8 rows per lane (variant 1) come within 8 KB
    # of 2^31.  One word more selects the nibble and class codes.
    present = (1 << 0) | (1 << 1) | (1 << 2) | (1 << 4) | (1 << 8) | (1 << 15)
    for L, offset_code in [(((1 << 18) - 1) * 32, True), ((1 << 18) * 32, False)]:
        chars, packed = long_rows(L)
        n, bps = packed.shape
        dense_total = capi.planes_layout(n, bps, capi.HD_SEM_DENSE, present)[3]
        assert (dense_total > 512 * (2 * bps // 32) * 4) == offset_code
        for md, early in [(2**31 - 1, False), (1500, True)]:
            want = O.all_pairs_dense(packed, 4, md)
            got = run_dev(packed, capi.HD_SEM_DENSE, 4, md, early, variant)
            np.testing.assert_array_equal(got, want, err_msg=f"dense L={L} md={md}")
            want = O.all_pairs_sparse(chars, False, 4, md)
            got = run_dev(packed, capi.HD_SEM_SPARSE, 4, md, early, variant)
            np.testing.assert_array_equal(got, want, err_msg=f"sparse L={L} md={md}")


def encoded(packed, sem):
    n, bps = packed.shape
    stream = torch.cuda.current_stream().cuda_stream
    d_packed = torch.from_numpy(np.ascontiguousarray(packed)).cuda()
    mask = capi.nibble_mask_dev(d_packed, n, bps, stream)
    _, _, _, total = capi.planes_layout(n, bps, sem, mask)
    d_planes = torch.empty(total, dtype=torch.int32, device="cuda")
    capi.encode_dev(d_packed, n, bps, sem, mask, d_planes, stream)
    return d_planes, mask


@pytest.mark.parametrize("variant", range(5))
def test_row_bands_concatenate_to_the_triangle(variant, rng):
    # hd_all_pairs_dev_rows: bands on 512-row boundaries (every variant's tile
    # rows divide 512) concatenate to the whole triangle; a band that starts
    # off a tile boundary is rejected before any launch
    n, L = 1500, 333
    seqs = random_strings(rng, n, L, "ACGT-N")
    packed = O.pack(O.as_chars(seqs))
    bps = packed.shape[1]
    stream = torch.cuda.current_stream().cuda_stream
    d_planes, mask = encoded(packed, capi.HD_SEM_DENSE)
    want = O.all_pairs_dense(packed, 2, 300)
    parts = []
    for r0, r1 in [(0, 512), (512, 1024), (1024, n)]:
        cells = r1 * (r1 - 1) // 2 - (r0 * (r0 - 1) // 2 if r0 else 0)
        d_band = torch.zeros(cells * 2, dtype=torch.uint8, device="cuda")
        for early in (False, True):
            capi.all_pairs_dev_rows(d_planes, n, bps, capi.HD_SEM_DENSE, mask, 2, 300, early,
                                    variant, r0, r1, d_band, stream)
            torch.cuda.synchronize()
            got = d_band.cpu().numpy().view(np.uint16)
            first = r0 * (r0 - 1) // 2 if r0 else 0
            np.testing.assert_array_equal(got, want[first:first + cells], err_msg=f"{r0}:{r1}")
        parts.append(got.copy())
    np.testing.assert_array_equal(np.concatenate(parts), want)
    with pytest.raises(capi.EngineError, match="row band"):
        capi.all_pairs_dev_rows(d_planes, n, bps, capi.HD_SEM_DENSE, mask, 2, 300, False,
                                variant, 100, 512, d_band, stream)


def test_tune_plain_picks_a_variant_that_matches(rng, tmp_path, monkeypatch):
    from hammingdist_amd import autotune

    monkeypatch.setenv("HD_TUNE_CACHE", str(tmp_path / "tune.json"))
    n, L = 2600, 700
    seqs = random_strings(rng, n, L, "ACGT-N")
    packed = O.pack(O.as_chars(seqs))
    bps = packed.shape[1]
    stream = torch.cuda.current_stream().cuda_stream
    d_planes, mask = encoded(packed, capi.HD_SEM_DENSE)
    d_out = torch.zeros(n * (n - 1) // 2, dtype=torch.uint8, device="cuda")
    v, times = autotune.tune_plain(d_planes, n, bps, capi.HD_SEM_DENSE, mask, 1, 255, True,
                                   d_out, stream, sample_rows=1024)
    assert 0 <= v < len(capi.variants()) and len(times) == len(capi.variants())
    # cached: the second call returns the same variant without a sweep
    assert autotune.tune_plain(d_planes, n, bps, capi.HD_SEM_DENSE, mask, 1, 255, True,
                               d_out, stream, sample_rows=1024)[0] == v
    capi.all_pairs_dev(d_planes, n, bps, capi.HD_SEM_DENSE, mask, 1, 255, True, v, d_out,
                       stream)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d_out.cpu().numpy(), O.all_pairs_dense(packed, 1, 255))


@pytest.mark.parametrize("gpus", [2, 3, 8])
def test_shard_bands_on_one_gpu_equal_the_triangle(gpus, rng):
    # each band as one GPU would compute it (hd_all_pairs_dev_rows over the
    # same planes), run here one after another: their concatenation is the
    # triangle
    from hammingdist_amd import shard

    n, L = 2100, 200
    seqs = random_strings(rng, n, L, "ACGT-N")
    packed = O.pack(O.as_chars(seqs))
    bps = packed.shape[1]
    stream = torch.cuda.current_stream().cuda_stream
    d_planes, mask = encoded(packed, capi.HD_SEM_DENSE)
    parts = []
    for r0, r1 in shard.row_bands(n, gpus):
        cells = shard.band_cells(r0, r1)
        if cells == 0:
            continue
        d_band = torch.empty(cells, dtype=torch.uint8, device="cuda")
        capi.all_pairs_dev_rows(d_planes, n, bps, capi.HD_SEM_DENSE, mask, 1, 255, True, -1,
                                r0, r1, d_band, stream)
        torch.cuda.synchronize()
        parts.append(d_band.cpu().numpy())
    np.testing.assert_array_equal(np.concatenate(parts), O.all_pairs_dense(packed, 1, 255))
