"""Host-side logic of the package that runs without a GPU: FASTA ingest and
dedup, loaders, text writers, error messages and argument checks.  The
oracle (oracle/fasta_oracle.py, oracle/hd_oracle.c writers) is the checker."""
import numpy as np
import pytest

import hammingdist_amd as hd
import oracle_lib as O
from conftest import random_strings
from oracle.fasta_oracle import fasta_indices, read_fasta


def write_fasta(path, seqs):
    with open(path, "w") as f:
        for i, s in enumerate(seqs):
            f.write(f">seq{i}\n{s}\n")


@pytest.mark.parametrize("content", [
    ">a\nAC\n>b\nAC\n>c\nGT\n",
    ">a\nAC\n>b\n",
    ">a\nAC\n>b",
    ">a\nAC\n\nGT\n>b\nACGT\n>c\nTT\n",
    "",
    "no header\nAC\n>x\nAC",
    ">a\r\nAC\r\n>b\r\nAC\r\n",
    ">only\n",
])
@pytest.mark.parametrize("n", [0, 1, 2, 5])
def test_sequence_indices_match_oracle(tmp_path, content, n):
    p = str(tmp_path / "f.fa")
    open(p, "w").write(content)
    got = hd.fasta_sequence_indices(p, n).tolist()
    assert got == fasta_indices(p, n)
    assert hd.fasta_sequence_indices(p, n).dtype == np.uint64


FASTA_EDGE_CASES = [
    ">a\nAC\n>b\nAC\n>c\nGT\n",
    ">a\nAC\n>b\n",
    ">a\nAC\n>b",
    ">a\nAC\n\nGT\n>b\nACGT\n>c\nTT\n",
    "",
    "\n",
    "no newline at all",
    "no header\nAC\n>x\nAC",
    ">a\r\nAC\r\n>b\r\nAC\r\n",
    ">only\n",
    "\n>a\nAC\n",
    ">a\n>b\nAC\n>c\n\n",
    ">a\nAC>b\nGT\n",
    ">a\nAC\nGT",
    ">a\n\n\n",
]


@pytest.mark.parametrize("content", FASTA_EDGE_CASES)
@pytest.mark.parametrize("n", [0, 1, 2, 3])
@pytest.mark.parametrize("dedup", [False, True])
def test_read_fasta_records_match_oracle(tmp_path, content, n, dedup):
    # src/hamming_impl.cc:200-242: records, their text and the dedup rows
    from hammingdist_amd import _native

    p = str(tmp_path / "f.fa")
    open(p, "w", newline="").write(content)
    data, idx = _native._read_fasta(p, dedup, n)
    want_data, want_idx = read_fasta(p, dedup, n)
    assert [d.decode("latin-1") for d in data] == want_data
    assert list(idx) == want_idx


def test_read_fasta_large_random_structure_matches_oracle(tmp_path, rng):
    # Several MiB, so the record splitter works on many chunks: random line
    # lengths, header lines, '>' inside lines, empty lines, CR, duplicates.
    lines = []
    pool = ["".join(rng.choice(list("ACGT-N"), size=int(rng.integers(0, 90))))
            for _ in range(40)]
    for _ in range(140_000):
        r = rng.random()
        if r < 0.08:
            lines.append(">h" + str(int(rng.integers(0, 1000))))
        elif r < 0.10:
            lines.append("")
        elif r < 0.12:
            lines.append("AC>GT\r")
        else:
            lines.append(pool[int(rng.integers(0, len(pool)))])
    content = "\n".join(lines) + ("\n" if rng.random() < 0.5 else "")
    p = str(tmp_path / "big.fa")
    open(p, "w", newline="").write(content)
    assert len(content) > 4 << 20
    from hammingdist_amd import _native

    for dedup, n in [(False, 0), (True, 0), (True, 5000)]:
        data, idx = _native._read_fasta(p, dedup, n)
        want_data, want_idx = read_fasta(p, dedup, n)
        assert len(data) == len(want_data)
        assert [d.decode("latin-1") for d in data] == want_data
        assert list(idx) == want_idx


def test_sequence_indices_random_dedup(tmp_path, rng):
    base = random_strings(rng, 30, 40, "ACGT-N")
    seqs = [base[k] for k in rng.integers(0, 30, size=200)]
    p = str(tmp_path / "f.fa")
    write_fasta(p, seqs)
    assert hd.fasta_sequence_indices(p).tolist() == fasta_indices(p)
    assert hd.fasta_sequence_indices(p, 77).tolist() == fasta_indices(p, 77)


def test_missing_file_is_an_error(tmp_path):
    missing = str(tmp_path / "nope.fa")
    with pytest.raises(RuntimeError, match="Failed to open file"):
        hd.fasta_sequence_indices(missing)
    with pytest.raises(RuntimeError, match="Failed to open file"):
        hd.fasta_reference_distances("ACGT", missing)
    with pytest.raises(RuntimeError, match="Failed to open file"):
        hd.from_fasta(missing)


def test_reference_error_messages():
    for data in ([], ["", ""]):
        with pytest.raises(RuntimeError, match="^Error: Empty sequence$"):
            hd.from_stringlist(data, False, False, 255)
    for data in (["ACGT", "A"], ["AC", "ACGTCG"], ["ACGT"] * 5 + ["A"], ["ACGT", "A-----"],
                 ["ACGT", ""]):
        with pytest.raises(RuntimeError,
                           match="^Error: Sequences do not all have the same length$"):
            hd.from_stringlist(data, False, False, 255)
        with pytest.raises(RuntimeError,
                           match="^Error: Sequences do not all have the same length$"):
            hd.distance(data[0], data[-1])
    with pytest.raises(RuntimeError, match="use_gpu=True cannot be used if include_x=True"):
        hd.from_stringlist(["ACGT", "ACGA"], True, True, 255)


def test_stream_requires_gpu_flag(tmp_path):
    p = str(tmp_path / "f.fa")
    write_fasta(p, ["ACGT", "ACGA"])
    with pytest.raises(RuntimeError,
                       match="from_fasta_to_lower_triangular is currently only available on GPU"):
        hd.from_fasta_to_lower_triangular(p, str(tmp_path / "o.txt"), use_gpu=False)


def test_positional_signatures():
    # from_stringlist has no defaults (python/hammingdist.cc:69-70)
    with pytest.raises(TypeError):
        hd.from_stringlist(["AC", "AG"])


def lt_file(path, rows):
    with open(path, "w") as f:
        for r in rows:
            f.write(",".join(str(v) for v in r) + "\n")


@pytest.mark.parametrize("n", [2, 3, 5, 11, 54, 120])
def test_lower_triangular_roundtrip_and_writer(tmp_path, n, rng):
    rows = [rng.integers(0, 300, size=i).tolist() for i in range(1, n)]
    src = str(tmp_path / "lt.txt")
    lt_file(src, rows)
    for loader, dt, cap in [(hd.from_lower_triangular, np.uint8, 255),
                            (hd.from_lower_triangular_large, np.uint16, 65535)]:
        ds = loader(src)
        cells = np.array([min(v, cap) for r in rows for v in r], dt)
        assert ds.lt_array.dtype == dt
        np.testing.assert_array_equal(ds.lt_array, cells)
        assert ds._distances == cells.tolist()
        out = str(tmp_path / "out.txt")
        ds.dump_lower_triangular(out)
        ref_out = str(tmp_path / "ref.txt")
        O.lib().oracle_write_lower_triangular(ref_out.encode(), cells.ctypes.data, n,
                                              cells.itemsize)
        assert open(out).read() == open(ref_out).read()
        full = str(tmp_path / "full.txt")
        ds.dump(full)
        O.lib().oracle_write_full(ref_out.encode(), cells.ctypes.data, n, cells.itemsize)
        assert open(full).read() == open(ref_out).read()
        for i in range(n):
            for j in range(n):
                a, b = max(i, j), min(i, j)
                want = 0 if i == j else int(cells[a * (a - 1) // 2 + b])
                assert ds[i, j] == want
        with pytest.raises(IndexError):
            ds[n, 0]


@pytest.mark.parametrize("top", [2, 10, 300])
def test_writer_many_chunks(tmp_path, top, rng):
    # 700 rows = 244,650 cells: several parallel chunks; top <= 10 takes the
    # one-digit formatting path.
    n = 700
    cells = rng.integers(0, top, size=n * (n - 1) // 2).astype(np.uint16)
    src = str(tmp_path / "lt.txt")
    O.lib().oracle_write_lower_triangular(src.encode(), cells.ctypes.data, n, cells.itemsize)
    ds = hd.from_lower_triangular_large(src)
    np.testing.assert_array_equal(ds.lt_array, cells)
    out = str(tmp_path / "out.txt")
    ds.dump_lower_triangular(out)
    assert open(out, "rb").read() == open(src, "rb").read()
    # a second dump over an existing, longer file truncates it
    open(out, "ab").write(b"stale tail\n" * 1000)
    ds.dump_lower_triangular(out)
    assert open(out, "rb").read() == open(src, "rb").read()


def test_dump_sequence_indices_without_dedup(tmp_path):
    src = str(tmp_path / "lt.txt")
    lt_file(src, [[1], [2, 3]])
    ds = hd.from_lower_triangular(src)
    out = str(tmp_path / "idx.txt")
    ds.dump_sequence_indices(out)
    assert open(out).read() == "0\n1\n2\n"


def test_from_csv_quirk_and_roundtrip(tmp_path, rng):
    # DataSet(filename) sizes the result n(n+1)/2 (include/hamming/hamming.hh:40)
    n = 9
    rows = [rng.integers(0, 200, size=i).tolist() for i in range(1, n)]
    src = str(tmp_path / "lt.txt")
    lt_file(src, rows)
    ds = hd.from_lower_triangular(src)
    csv = str(tmp_path / "m.csv")
    ds.dump(csv)
    back = hd.from_csv(csv)
    assert back.lt_array.shape == (n * (n + 1) // 2,)
    for i in range(n):
        for j in range(n):
            assert back[i, j] == ds[i, j]


def test_empty_lower_triangle_writes_no_file(tmp_path):
    src = str(tmp_path / "lt.txt")
    open(src, "w").write("")
    ds = hd.from_lower_triangular(src)
    out = tmp_path / "out.txt"
    ds.dump_lower_triangular(str(out))
    assert not out.exists()


def test_generator_is_deterministic():
    from hammingdist_amd import workloads

    s = workloads.plain(n=50, L=300, seed=77)
    a, b = workloads.generate_chars(s), workloads.generate_chars(s)
    np.testing.assert_array_equal(a, b)
    p = workloads.generate_packed(s)
    np.testing.assert_array_equal(p, O.pack(a))
    txt = {bytes(r) for r in a}
    assert len(txt) < 50  # duplicates present
    assert set(np.unique(a).tobytes()) <= set(b"ACGTN-")


@pytest.mark.parametrize("map_bytes", [None, "1"])
@pytest.mark.parametrize("top", [3, 10, 70000])
def test_row_writer_bands_match_oracle(tmp_path, top, map_bytes, rng, monkeypatch):
    # The streaming writer's use: one file written band after band, at file
    # offsets that are not page aligned.  Bands are written with pwrite, or
    # (HD_WRITER_MAP_BYTES=1: the path of writes of 1 GiB and more) through
    # a mapping of the file: one-digit text formatted straight into it, other
    # text copied into it.
    from hammingdist_amd import _native

    if map_bytes:
        monkeypatch.setenv("HD_WRITER_MAP_BYTES", map_bytes)

    n = 900
    dt = np.uint8 if top < 256 else np.uint16
    cells = rng.integers(0, top, size=n * (n - 1) // 2).astype(dt)
    want = str(tmp_path / "want.txt")
    O.lib().oracle_write_lower_triangular(want.encode(), cells.ctypes.data, n, cells.itemsize)
    bounds = sorted({0, n} | {int(x) for x in rng.integers(1, n, size=7)})
    out = tmp_path / "out.txt"
    out.write_bytes(b"stale" * 400000)  # longer than the result: truncated on open
    _native._write_row_bands(str(out), cells, bounds, int(cells.max()))
    assert out.read_bytes() == open(want, "rb").read()
    # not a regular file: the pwrite path
    _native._write_row_bands("/dev/null", cells, bounds, int(cells.max()))
