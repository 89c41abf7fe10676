"""The reference's Python tests (python/tests/test_hammingdist.py) and its
C++ CPU==GPU consistency tests (src/hamming_t.cc:346-451), run against the
package API, plus the plain all-pairs workload at full size; the oracle is
the checker throughout."""
import json
import os

import numpy as np
import pytest

import hammingdist_amd as hd
import oracle_lib as O
from conftest import random_strings
from oracle.fasta_oracle import fasta_indices, read_fasta

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
KATS = json.load(open(os.path.join(HERE, "golden", "reference_kats.json")))


def write_fasta(path, seqs):
    with open(path, "w") as f:
        for i, s in enumerate(seqs):
            f.write(f">seq{i}\n{s}\n")


def expected_cells(seqs, include_x, use_gpu, out_bytes, max_distance):
    return O.distances(O.as_chars(seqs), include_x, use_gpu, out_bytes, max_distance)[0]


def check_output_sizes(dat, n_in, n_out, tmp_out_file, fasta_sequence_indices=None):
    # python/tests/test_hammingdist.py:13-36
    assert dat.lt_array.shape == (n_out * (n_out - 1) // 2,)
    dat.dump(tmp_out_file)
    dump = np.loadtxt(tmp_out_file, delimiter=",")
    assert len(dump) == n_out and len(dump[0]) == n_out
    dat.dump_lower_triangular(tmp_out_file)
    with open(tmp_out_file) as f:
        data = f.read().splitlines()
    assert len(np.array(data[0].split(","), dtype=float)) == 1
    assert len(np.array(data[-1].split(","), dtype=float)) == n_out - 1
    dat.dump_sequence_indices(tmp_out_file)
    indices = np.loadtxt(tmp_out_file, delimiter=",")
    assert len(indices) == n_in and indices[0] == 0
    if fasta_sequence_indices is not None:
        assert np.allclose(indices, fasta_sequence_indices)


@pytest.mark.parametrize("from_fasta_func", [hd.from_fasta, hd.from_fasta_large])
@pytest.mark.parametrize("use_gpu", [False, True])
def test_from_fasta(from_fasta_func, use_gpu, tmp_path):
    kat = KATS["from_fasta"]
    fasta_file, out = str(tmp_path / "fasta.txt"), str(tmp_path / "out.txt")
    write_fasta(fasta_file, kat["sequences"])
    check_output_sizes(from_fasta_func(fasta_file, use_gpu=use_gpu), 6, 6, out)
    check_output_sizes(from_fasta_func(fasta_file, n=5, use_gpu=use_gpu), 5, 5, out)
    check_output_sizes(from_fasta_func(fasta_file, include_x=True, use_gpu=False), 6, 6, out)
    idx = hd.fasta_sequence_indices(fasta_file)
    check_output_sizes(from_fasta_func(fasta_file, remove_duplicates=True, use_gpu=use_gpu),
                       6, 4, out, idx)
    check_output_sizes(from_fasta_func(fasta_file, remove_duplicates=True, include_x=True,
                                       use_gpu=False), 6, 4, out, idx)
    idx = hd.fasta_sequence_indices(fasta_file, n=2)
    check_output_sizes(from_fasta_func(fasta_file, include_x=True, n=2, use_gpu=False), 2, 2,
                       out, idx)
    idx = hd.fasta_sequence_indices(fasta_file, n=3)
    check_output_sizes(from_fasta_func(fasta_file, remove_duplicates=True, n=3,
                                       use_gpu=use_gpu), 3, 2, out, idx)
    idx = hd.fasta_sequence_indices(fasta_file, n=5)
    check_output_sizes(from_fasta_func(fasta_file, remove_duplicates=True, n=5,
                                       include_x=True, use_gpu=False), 5, 3, out, idx)
    if use_gpu:
        with pytest.raises(RuntimeError, match="use_gpu=True cannot be used if include_x=True"):
            from_fasta_func(fasta_file, include_x=True, use_gpu=True)


@pytest.mark.parametrize("chars,include_x", [(["A"], False), (list("ACGT-"), False),
                                             (list("ACGT-"), True), (list("ACGT-X"), False),
                                             (list("ACGT-X"), True)])
@pytest.mark.parametrize("max_distance", [0, 1, 2, 3, 89, 497, 9999999])
def test_fasta_reference_distances(chars, include_x, max_distance, tmp_path, rng):
    # python/tests/test_hammingdist.py:90-136 (seeded here)
    n_seq, n_chars = 50, 25
    sequences = random_strings(rng, n_seq, n_chars, "".join(chars))
    fasta_file = str(tmp_path / "fasta.txt")
    write_fasta(fasta_file, sequences)
    data = hd.from_fasta(fasta_file, remove_duplicates=False, include_x=include_x,
                         max_distance=max_distance)
    lt_array = data.lt_array
    assert lt_array.shape == (n_seq * (n_seq - 1) // 2,)
    np.testing.assert_array_equal(
        lt_array, expected_cells(sequences, include_x, False, 1, max_distance))
    lt_matrix = np.zeros((n_seq, n_seq), dtype=np.uint8)
    lt_matrix[np.tril_indices(n_seq, -1)] = lt_array
    for i, sequence in enumerate(sequences):
        vec = hd.fasta_reference_distances(sequence, fasta_file, include_x=include_x)
        assert len(vec) == len(sequences) and vec.dtype == np.uint32
        for j, dist in enumerate(vec):
            if include_x or i != j:
                assert data[i, j] == min(max_distance, dist)
                assert lt_matrix[max(i, j), min(i, j)] == min(max_distance, dist)
            assert dist == hd.distance(sequences[i], sequences[j], include_x=include_x)
            assert dist == O.distance_char(sequences[i], sequences[j], include_x)


def test_distance():
    for a, b, x, want in KATS["distance_kats"]:
        assert hd.distance(a, b, include_x=x) == want
    with pytest.raises(RuntimeError):
        hd.distance("ACGT", "ACC")
    assert hd.distance("", "") == 0


def test_distance_two_pairs_and_char_rules():
    for a, b in KATS["distance_two_pairs"]:
        d = hd.from_stringlist([a, b], False, True, 255)
        assert d[0, 1] == 2 and d[1, 0] == 2 and d[0, 0] == 0
        d = hd.from_stringlist([a, b], False, False, 255)
        assert d[0, 1] == 2
        assert hd.distance(a, b) == 2 and hd.distance(a, a) == 0
    valid, invalid = KATS["valid_chars"], KATS["invalid_chars"]
    for c1 in valid + invalid + ["-"]:
        for c2 in valid + invalid + ["-"]:
            want = O.distance_char(c1, c2)
            assert hd.distance(c1, c2) == want
            assert hd.from_stringlist([c1, c2], False, False, 255)[0, 1] == \
                int(expected_cells([c1, c2], False, False, 1, 255)[0])


@pytest.mark.parametrize("max_distance", [0, 1, 2, 3, 89, 497, 9999999])
def test_from_fasta_to_lower_triangular(tmp_path, max_distance):
    lt = KATS["lower_triangular"]
    fasta_file, out = str(tmp_path / "fasta.txt"), str(tmp_path / "out.txt")
    write_fasta(fasta_file, lt["sequences"])
    hd.from_fasta_to_lower_triangular(fasta_file, out, max_distance=max_distance)
    with open(out) as f:
        data = f.read().splitlines()
    assert len(data) == 2
    want = [[min(v, max_distance) for v in r] for r in lt["rows"]]
    assert [list(map(int, line.split(","))) for line in data] == want


@pytest.mark.parametrize("samples", [2, 3, 5, 11, 54, 120, 532, 981, 1568])
@pytest.mark.parametrize("threshold", [0, 1, 2, 3, 4, 9, 89, 497])
def test_dump_sparse(tmp_path, samples, threshold, rng):
    # python/tests/test_hammingdist.py:177-204, plus exact row-major text
    lt_file = str(tmp_path / "lt.txt")
    rows = [rng.integers(low=0, high=256, size=n) for n in range(1, samples)]
    with open(lt_file, "w") as f:
        for r in rows:
            f.write(",".join(str(x) for x in r) + "\n")
    data = hd.from_lower_triangular(lt_file)
    cells = np.concatenate(rows).astype(np.uint8)
    sparse_file = str(tmp_path / "sparse.txt")
    data.dump_sparse(sparse_file, threshold)
    ref_file = str(tmp_path / "ref.txt")
    O.lib().oracle_write_sparse(ref_file.encode(), cells.ctypes.data, samples, 1, threshold)
    assert open(sparse_file).read() == open(ref_file).read()


def test_from_fasta_duplicates_kat(tmp_path):
    d = KATS["dedup"]
    p = str(tmp_path / "f.fa")
    write_fasta(p, d["sequences"])
    for use_gpu in (False, True):
        for n in d["ns"]:
            for md in d["max_dists"]:
                ds = hd.from_fasta(p, False, True, n, use_gpu, md)
                assert ds.nsamples == 3
                out = str(tmp_path / "idx.txt")
                ds.dump_sequence_indices(out)
                assert list(map(int, open(out).read().split())) == d["sequence_indices"]
                for key, want in d["unique_distances"].items():
                    i, j = map(int, key.split(","))
                    assert ds[i, j] == min(md, want) and ds[j, i] == min(md, want)


@pytest.mark.parametrize("ob", [1, 2])
def test_saturation(ob):
    sat = KATS["saturation"]
    tmax = sat["type_max"][str(ob)]
    fn = hd.from_fasta if ob == 1 else hd.from_fasta_large
    for use_gpu in (False, True):
        for off in sat["n_offsets"]:
            L = tmax + off
            for md in sat["max_dists"]:
                if ob == 1:
                    ds = hd.from_stringlist(["A" * L, "T" * L], False, use_gpu, md)
                    assert ds[0, 1] == min(md, tmax)


def test_cpu_gpu_route_consistency(tmp_path, rng):
    # src/hamming_t.cc:346-451 on uniform ACGT- data: both routes, every cell
    c = KATS["consistency"]
    for n_chars in [1, 5, 17, 88, 381, 1023]:
        for n_samples in [2, 3, 4, 7, 11, 127, 128, 257, 703]:
            seqs = random_strings(rng, n_samples, n_chars, "ACGT-")
            md = c["max_dists"][(n_chars + n_samples) % len(c["max_dists"])]
            for use_gpu in (False, True):
                ds = hd.from_stringlist(seqs, False, use_gpu, md)
                np.testing.assert_array_equal(
                    ds.lt_array, expected_cells(seqs, False, use_gpu, 1, md))
            p = str(tmp_path / "f.fa")
            write_fasta(p, seqs)
            ds16 = hd.from_fasta_large(p, False, True, 0, True, md)
            fa_data, _ = read_fasta(p, True, 0)
            np.testing.assert_array_equal(ds16.lt_array,
                                          expected_cells(fa_data, False, True, 2, md))
            out = str(tmp_path / "lt.txt")
            hd.from_fasta_to_lower_triangular(p, out, True, 0, True, md)
            back = hd.from_lower_triangular_large(out) if len(fa_data) > 1 else None
            if back is not None:
                np.testing.assert_array_equal(back.lt_array, ds16.lt_array)


def test_from_fasta_cpu_gpu_grid(tmp_path):
    # src/hamming_t.cc:375-410: from_fasta<uint8_t> and <uint16_t> on uniform
    # ACGT- sequences (src/tests.cc:7-20), with and without duplicate
    # removal, over its whole grid of lengths, caps and sample counts; both
    # routes, every cell, against the oracle.
    rng = np.random.default_rng(12345)
    p = str(tmp_path / "grid.fa")
    calls = 0
    for remove_duplicates in (False, True):
        for n_chars in (17, 88, 381, 1023):
            for md in (0, 1, 2, 3, 89, 497, 9999999):
                for n_samples in (2, 3, 4, 7, 11, 127, 128, 255, 256, 257, 703):
                    seqs = random_strings(rng, n_samples, n_chars, "ACGT-")
                    if remove_duplicates and n_samples > 3:
                        # a few repeats, so that removal has work to do
                        seqs[n_samples - 1] = seqs[0]
                        seqs[n_samples // 2] = seqs[1]
                    write_fasta(p, seqs)
                    data, _ = read_fasta(p, remove_duplicates, 0)
                    for fn, ob in ((hd.from_fasta, 1), (hd.from_fasta_large, 2)):
                        for use_gpu in (False, True):
                            ds = fn(p, False, remove_duplicates, 0, use_gpu, md)
                            want = expected_cells(data, False, use_gpu, ob, md)
                            assert ds.lt_array.dtype == (np.uint8 if ob == 1 else np.uint16)
                            np.testing.assert_array_equal(
                                ds.lt_array, want,
                                err_msg=f"rd={remove_duplicates} L={n_chars} md={md} "
                                        f"n={n_samples} bytes={ob} gpu={use_gpu}")
                            calls += 1
                    if n_chars != 1023 and md != 497:
                        # src/hamming_t.cc:411-451: the streamed lower
                        # triangle, read back as uint16, equals from_fasta<uint16_t>
                        out = str(tmp_path / "grid_lt.txt")
                        hd.from_fasta_to_lower_triangular(p, out, remove_duplicates, 0, True, md)
                        back = hd.from_lower_triangular_large(out)
                        np.testing.assert_array_equal(
                            back.lt_array, expected_cells(data, False, False, 2, md),
                            err_msg=f"streamed: rd={remove_duplicates} L={n_chars} md={md} "
                                    f"n={n_samples}")
                        calls += 1
    assert calls == 2 * 4 * 7 * 11 * 4 + 2 * 3 * 6 * 11


def test_low_diversity_takes_sparse_semantics():
    # 0.2% of sites differ from the consensus, with 'N's: the CPU route
    # (use_gpu=False) is the sparse semantics, the GPU route is dense.
    rng = np.random.default_rng(3)
    base = np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, 2000)]
    seqs = []
    for k in range(60):
        s = base.copy()
        pos = rng.integers(0, 2000, 4)
        s[pos] = np.frombuffer(b"NNC-", np.uint8)
        seqs.append(bytes(s).decode())
    for use_gpu in (False, True):
        want, route = O.distances(O.as_chars(seqs), False, use_gpu, 1, 255)
        assert route == (0 if use_gpu else 1)
        ds = hd.from_stringlist(seqs, False, use_gpu, 255)
        np.testing.assert_array_equal(ds.lt_array, want)


def test_plain_workload_full_size(tmp_path):
    # n=2,000, L=29,903 plain generator, remove_duplicates=True,
    # max_distance=255, uint8, dump_sequence_indices
    from hammingdist_amd import workloads

    spec = workloads.plain()
    p = str(tmp_path / "plain.fa")
    workloads.write_fasta(spec, p)
    ds = hd.from_fasta(p, remove_duplicates=True, max_distance=255)
    data, idx = read_fasta(p, True, 0)
    assert ds.nsamples == len(data) and len(data) < 2000
    want, route = O.distances(O.as_chars(data), False, False, 1, 255)
    np.testing.assert_array_equal(ds.lt_array, want)
    out = str(tmp_path / "idx.txt")
    ds.dump_sequence_indices(out)
    assert list(map(int, open(out).read().split())) == idx
    assert hd.fasta_sequence_indices(p).tolist() == fasta_indices(p)
