"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute calls: these run without a GPU)."""
import os
import re

from hammingdist_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(hd_[a-z_0-9]+)\s*\(", text,
                             flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    assert {"hd_all_pairs", "hd_all_pairs_stream", "hd_reference_distances",
            "hd_distance_pair", "hd_device_count", "hd_last_error"} <= names


def test_library_exports_every_declared_symbol():
    lib = capi.lib()
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_signatures_cover_header():
    assert declared_functions() <= set(capi.SIGNATURES)


def test_no_device_is_a_clean_error():
    if capi.device_count() > 0:
        return
    import numpy as np
    import pytest

    packed = np.full((2, 8), 0xFF, np.uint8)
    out = np.zeros(1, np.uint8)
    with pytest.raises(capi.EngineError, match="no HIP device"):
        capi.all_pairs(packed, 2, 8, capi.HD_SEM_DENSE, 1, 255, out)


def test_invalid_arguments_are_rejected_before_device_use():
    import numpy as np
    import pytest

    packed = np.full((2, 8), 0xFF, np.uint8)
    out = np.zeros(1, np.uint8)
    with pytest.raises(capi.EngineError, match="max_dist"):
        capi.all_pairs(packed, 2, 8, capi.HD_SEM_DENSE, 1, 256, out)
    with pytest.raises(capi.EngineError, match="out_bytes"):
        capi.all_pairs(packed, 2, 8, capi.HD_SEM_DENSE, 3, 1, out)
    # planes layout is pure host arithmetic
    n_pad, W, P, total = capi.planes_layout(1000, 14952, capi.HD_SEM_DENSE, 0xFFFF)
    assert (n_pad, W, P) == (1024, 935, 4) and total == 1024 * 935 * 4
    # sparse semantics: every class present and '-' get a code, 2 bits for up
    # to 4 codes, else 3, plus the '-' plane and the correction counts
    letters_gap = (1 << 1) | (1 << 2) | (1 << 4) | (1 << 8) | (1 << 15)
    assert capi.planes_layout(10, 8, capi.HD_SEM_SPARSE, letters_gap)[2] == 4
    assert capi.planes_layout(10, 8, capi.HD_SEM_SPARSE, (1 << 1) | (1 << 2) | (1 << 4))[2] == 3
    assert capi.planes_layout(10, 8, capi.HD_SEM_SPARSE_X, 0xFFFF)[2] == 4
    assert capi.planes_layout(10, 8, capi.HD_SEM_SPARSE_X, (1 << 3) | (1 << 0))[2] == 3
    assert capi.planes_layout(10, 8, capi.HD_SEM_SPARSE_X, (1 << 1) | (1 << 15))[2] == 3
    n_pad, W, P, total = capi.planes_layout(1000, 5000, capi.HD_SEM_SPARSE_X, 0xFFFF)
    # (whole-row counts twice, per 32-word chunk counts and bitmaps: 2 + 2 * 10)
    assert (n_pad, W, P) == (1024, 313, 4) and total == 1024 * 313 * 4 + 22 * 1024
    assert capi.planes_layout(10, 8, capi.HD_SEM_CHAR_X, 1 << 3)[2] == 5
    # dense data holding only A C G T '-' and invalid codes: the 3-op code,
    # 4 planes plus per-row counts; any other code (X = 3): nibble planes
    acgt_gap_inv = (1 << 0) | (1 << 1) | (1 << 2) | (1 << 4) | (1 << 8) | (1 << 15)
    n_pad, W, P, total = capi.planes_layout(1000, 14952, capi.HD_SEM_DENSE, acgt_gap_inv)
    # (whole-row counts Z, G and per 32-word chunk counts: 2 + 2 * 30 rows)
    assert (n_pad, W, P) == (1024, 935, 4) and total == 1024 * 935 * 4 + 62 * 1024
    assert capi.planes_layout(1000, 14952, capi.HD_SEM_DENSE,
                              acgt_gap_inv | (1 << 3))[3] == 1024 * 935 * 4


def test_long_row_sparse_code_layout(monkeypatch):
    from hammingdist_amd import capi

    # Rows of 2^23 sites or more take the class code: planes only, no counts
    # (bytes_per_seq 2^22 = 2^23 sites = 2^18 words); so does the dense
    # semantics' nibble code.
    n_pad, W, P, total = capi.planes_layout(2, 1 << 22, capi.HD_SEM_SPARSE, 0xFFFF)
    assert (n_pad, W, P) == (512, 1 << 18, 4) and total == 512 * (1 << 18) * 4
    n_pad, W, P, total = capi.planes_layout(2, (1 << 22) - 16, capi.HD_SEM_SPARSE, 0xFFFF)
    assert total > 512 * W * 4  # one word shorter: the gap code and its counts
    acgt_gap = (1 << 1) | (1 << 2) | (1 << 4) | (1 << 8) | (1 << 15)
    assert capi.planes_layout(2, 1 << 22, capi.HD_SEM_DENSE, acgt_gap)[3] == 512 * (1 << 18) * 4
    assert capi.planes_layout(2, (1 << 22) - 16, capi.HD_SEM_DENSE, acgt_gap)[3] > \
        512 * ((1 << 18) - 1) * 4
    monkeypatch.setenv("HD_SPARSE_CLASS_CODE", "1")
    n_pad, W, P, total = capi.planes_layout(1000, 5000, capi.HD_SEM_SPARSE_X, 0xFFFF)
    assert (n_pad, W, P) == (1024, 313, 4) and total == 1024 * 313 * 4


def test_default_tile_shape():
    from hammingdist_amd import capi

    shapes = {v["variant"]: (v["rows_per_lane"], v["cols_per_wave"]) for v in capi.variants()}
    acgt_gap_inv = (1 << 0) | (1 << 1) | (1 << 2) | (1 << 4) | (1 << 8) | (1 << 15)

    def shape(n, bps, sem, mask, ee, md):
        return shapes[capi.default_variant(n, bps, sem, mask, ee, md)]

    # small datasets: 128 x 32 tiles
    assert shape(300, 14952, capi.HD_SEM_DENSE, acgt_gap_inv, False, 255) == (2, 8)
    # dense 3-op code comparing every site: 128 x 64 tiles
    assert shape(50000, 14952, capi.HD_SEM_DENSE, acgt_gap_inv, False, 2) == (2, 16)
    # early exit under a small cap: 128 x 32 tiles stop soonest
    assert shape(50000, 14952, capi.HD_SEM_DENSE, acgt_gap_inv, True, 2) == (2, 8)
    # early exit with a cap of 255: 256 x 64 tiles from 24,576 rows, 128 x 64
    # from 8,192 rows
    assert shape(50000, 14952, capi.HD_SEM_DENSE, acgt_gap_inv, True, 255) == (4, 16)
    assert shape(10000, 14952, capi.HD_SEM_DENSE, acgt_gap_inv, True, 255) == (2, 16)
    assert shape(6000, 14952, capi.HD_SEM_DENSE, acgt_gap_inv, True, 255) == (2, 8)
    # a cap out of reach of the row length turns the votes off
    assert shape(50000, 14952, capi.HD_SEM_DENSE, acgt_gap_inv, True, 65535) == (2, 16)
    # fewer rows: smaller tiles, so that the launch holds enough of them
    assert shape(1000, 14952, capi.HD_SEM_DENSE, acgt_gap_inv, False, 255) == (2, 8)
    assert shape(8191, 14952, capi.HD_SEM_DENSE, acgt_gap_inv, False, 65535) == (2, 8)
    assert shape(8192, 14952, capi.HD_SEM_DENSE, acgt_gap_inv, False, 65535) == (2, 16)
    assert shape(100000, 14952, capi.HD_SEM_DENSE, acgt_gap_inv, False, 65535) == (2, 16)
    assert shape(24575, 14952, capi.HD_SEM_DENSE, acgt_gap_inv, True, 255) == (2, 16)
    assert shape(24576, 14952, capi.HD_SEM_DENSE, acgt_gap_inv, True, 255) == (4, 16)
    # gap codes without votes: 256 x 64 tiles from 32,768 rows, 128 x 64 from
    # 4,096, else 128 x 32; with votes 256 x 32 from 32,768 rows, else 128 x 32
    assert shape(100000, 5000, capi.HD_SEM_SPARSE_X, 0xFFFF, False, 65535) == (4, 16)
    assert shape(20000, 5000, capi.HD_SEM_SPARSE_X, 0xFFFF, False, 65535) == (2, 16)
    assert shape(3000, 5000, capi.HD_SEM_SPARSE_X, 0xFFFF, False, 65535) == (2, 8)
    assert shape(100000, 5000, capi.HD_SEM_SPARSE_X, 0xFFFF, True, 255) == (4, 8)
    assert shape(100000, 5000, capi.HD_SEM_SPARSE_X, 0xFFFF, True, 3) == (4, 8)
    assert shape(20000, 5000, capi.HD_SEM_SPARSE_X, 0xFFFF, True, 255) == (2, 8)
    # the nibble code (dense data holding X) and the one-hot code: 256 x 64
    # from 8,192 rows, 128 x 64 from 4,096, else 128 x 32
    assert shape(30000, 5000, capi.HD_SEM_DENSE, acgt_gap_inv | (1 << 3), False, 65535) == \
        (4, 16)
    assert shape(30000, 5000, capi.HD_SEM_CHAR_X, 0xFFFF, False, 65535) == (4, 16)
    assert shape(6000, 5000, capi.HD_SEM_CHAR_X, 0xFFFF, False, 65535) == (2, 16)
    assert shape(3000, 5000, capi.HD_SEM_CHAR_X, 0xFFFF, False, 65535) == (2, 8)


def test_tuning_signatures_separate_the_work_they_describe():
    # host-only: the cache keys tell apart plane codes, caps, cell types,
    # early exit and sizes (hd_planes_layout needs no device)
    from hammingdist_amd import autotune

    dense = (1 << 0) | (1 << 1) | (1 << 2) | (1 << 4) | (1 << 8) | (1 << 15)
    base = autotune.plain_signature(50000, 14952, capi.HD_SEM_DENSE, dense, 1, 2, True)
    assert base != autotune.plain_signature(50000, 14952, capi.HD_SEM_DENSE, dense, 1, 3, True)
    assert base != autotune.plain_signature(50000, 14952, capi.HD_SEM_DENSE, dense, 2, 2, True)
    assert base != autotune.plain_signature(50000, 14952, capi.HD_SEM_DENSE, dense, 1, 2, False)
    assert base != autotune.plain_signature(50000, 14952, capi.HD_SEM_DENSE,
                                            dense | (1 << 3), 1, 2, True)  # X: nibble code
    assert base != autotune.plain_signature(50000, 14952, capi.HD_SEM_SPARSE_X, dense, 1, 2,
                                            True)
    assert base == autotune.plain_signature(50001, 14952, capi.HD_SEM_DENSE, dense, 1, 2, True)
    jobs = [(100, 800, 1), (2000, 15000, 2)]
    assert autotune.batch_signature(jobs, True) != autotune.batch_signature(jobs, False)
    assert autotune.batch_signature(jobs) != autotune.batch_signature([(100, 800, 1)] * 2)


def test_row_bands_partition_the_triangle():
    # the sharding rule: contiguous, aligned, covering, about equal cells
    from hammingdist_amd import shard

    for n in (2, 511, 512, 5000, 50_000, 100_000):
        for g in (1, 2, 3, 8):
            bands = shard.row_bands(n, g)
            assert bands[0][0] == 0 and bands[-1][1] == n
            assert all(a[1] == b[0] for a, b in zip(bands, bands[1:]))
            assert all(r0 % shard.ALIGN == 0 or r0 == n for r0, _ in bands)
            cells = [shard.band_cells(*b) for b in bands]
            assert sum(cells) == n * (n - 1) // 2
            if n >= 50_000:
                # 512-row boundaries: a block near the end is a few % of a band
                assert max(cells) / (sum(cells) / g) < (1.05 if g <= 3 else 1.10)
