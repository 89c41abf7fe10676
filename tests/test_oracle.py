"""Pins the oracle (oracle/hd_oracle.c, oracle/fasta_oracle.py) before any
GPU result is compared with it.

Each oracle function is checked against (a) the known-answer vectors of the
reference's own tests (tests/golden/reference_kats.json), (b) the reference's
own SIMD pair kernels compiled from its sources (oracle/_ref, where built),
and (c) what the mathematics fixes: closed forms, exhaustive brute force over
small alphabets, and invariants — chosen so that a dropped term, a wrong
sign/index or swapped operands fails one of them.
"""
import itertools
import json
import os

import numpy as np
import pytest

import oracle_lib as O
from conftest import random_strings

HERE = os.path.dirname(os.path.abspath(__file__))
KATS = json.load(open(os.path.join(HERE, "golden", "reference_kats.json")))
VALID, INVALID = KATS["valid_chars"], KATS["invalid_chars"]


def cpu_route(seqs, include_x=False, out_bytes=4, max_distance=2**31 - 1):
    cells, route = O.distances(O.as_chars(seqs), include_x, False, out_bytes, max_distance)
    return cells, route


def gpu_route(seqs, out_bytes=4, max_distance=2**31 - 1):
    return O.distances(O.as_chars(seqs), False, True, out_bytes, max_distance)[0]


# ---- lookup / packing -------------------------------------------------------

def test_pack_layout_and_padding():
    # src/hamming_impl.cc:244-267: low nibble first, odd tail padded with '-',
    # 0xFF up to a multiple of 8 bytes
    p = O.pack(O.as_chars(["ACGTX-N"]))[0]
    assert p.size == 8
    assert p.tolist() == [0x21, 0x84, 0xF0, 0xF0, 0xFF, 0xFF, 0xFF, 0xFF]
    px = O.pack(O.as_chars(["ACGTX-N"]), include_x=True)[0]
    assert px[2] == 0xF3
    for L, bps in [(1, 8), (16, 8), (17, 16), (29903, 14952), (10000, 5000), (1000, 504)]:
        assert O.bytes_per_seq(L) == bps


# ---- per-character rules (src/hamming_t.cc:28-91) ---------------------------

def _char_pairs():
    r = KATS["char_rules"]
    for c in VALID:
        yield c, c, r["equal_valid"]
        yield c, "-", r["valid_gap"]
        yield "-", c, r["valid_gap"]
        for c2 in VALID:
            if c2 != c:
                yield c, c2, r["different_valid"]
        for c2 in INVALID:
            yield c, c2, r["valid_invalid"]
    for c in INVALID:
        yield c, "-", r["invalid_gap"]
        yield "-", c, r["invalid_gap"]
        for c2 in INVALID:
            yield c, c2, r["invalid_invalid"]


def test_char_rules_distance_and_stringlist():
    for a, b, want in _char_pairs():
        assert O.distance_char(a, b) == want, (a, b)
        cells, _ = cpu_route([a, b])
        assert int(cells[0]) == want, (a, b)


def test_distance_two_pairs():
    for a, b in KATS["distance_two_pairs"]:
        assert O.distance_char(a, b) == 2 and O.distance_char(b, a) == 2
        assert O.distance_char(a, a) == 0
        assert int(cpu_route([a, b])[0][0]) == 2
        assert int(gpu_route([a, b])[0]) == 2


def test_distance_kats():
    for a, b, x, want in KATS["distance_kats"]:
        assert O.distance_char(a, b, x) == want


# ---- dense kernel -----------------------------------------------------------

def test_dense_closed_forms():
    # src/hamming_impl_t.cc:6-47: identical -> 0; n x 'A' vs n x 'G' -> min(md, n)
    rng = np.random.default_rng(5)
    for L in KATS["closed_form_lengths"]:
        a = O.pack(O.as_chars(["A" * L, "G" * L]))
        s = random_strings(rng, 1, L, "ACGT-")[0] if L < 70000 else "ACGT-" * (L // 5) + "A" * (L % 5)
        g = O.pack(O.as_chars([s]))[0]
        for md in KATS["closed_form_max_dists"]:
            assert O.distance_dense(a[0], a[1], md) == min(md, L)
            assert O.distance_dense(g, g, md) == 0


def test_dense_equals_reference_simd_kernels():
    ref = O.ref()
    if ref is None:
        pytest.skip("oracle/_ref (the reference's SIMD kernels) not built here")
    rng = np.random.default_rng(123451)
    for L in [1, 5, 17, 64, 65, 127, 128, 129, 255, 256, 257, 1023, 1024, 1025, 4097, 29903]:
        seqs = random_strings(rng, 6, L, "ACGT-NX")
        p = O.pack(O.as_chars(seqs))
        for i in range(0, 6, 2):
            for md in [0, 1, 2, 11, 254, 255, 256, 999, 9876544]:
                want = O.distance_dense(p[i], p[i + 1], md)
                for which in (1, 2, 3):
                    got = ref.ref_distance(which, p[i].ctypes.data, p[i + 1].ctypes.data,
                                           p.shape[1], md)
                    assert got == want, (L, md, which)


SIMD_LENGTHS = [1, 63, 64, 65, 127, 128, 129, 8191, 29903]
SIMD_CAPS = [0, 1, 2, 16, 254, 255, 256, 65535]


@pytest.mark.parametrize("L", SIMD_LENGTHS)
def test_simd_restatement_equals_scalar_and_reference(L):
    # The oracle's AVX-512 restatement (distance_avx512's u8 lane counters and
    # n_inner early-exit cadence under distances<T>'s row loop) against the
    # scalar distance_cpp restatement on every cell, as the reference's own
    # src/distance_avx512_t.cc:6-73 pins its SIMD kernel against distance_cpp;
    # and against the reference's compiled AVX-512 kernel where oracle/_ref is
    # built.  Rows hold long runs of disagreement (distances up to L, past the
    # u8 lane range of one inner block) and identical pairs (distance 0).
    rng = np.random.default_rng(7000 + L)
    n = 23
    base = rng.choice(list("ACGT"), size=L)
    seqs = []
    for r in range(n):
        s = base.copy()
        k = int(rng.integers(0, L + 1)) if r % 4 else (0 if r % 8 == 0 else L)
        pos = rng.choice(L, size=k, replace=False)
        s[pos] = rng.choice(list("ACGT-NX"), size=k)
        seqs.append("".join(s))
    seqs.append(seqs[5])  # an identical pair
    p = O.pack(O.as_chars(seqs))
    n = len(seqs)
    ref = O.ref()
    for md in SIMD_CAPS:
        for ob in (1, 2, 4):
            tmax = (1 << (8 * ob)) - 1
            if md > tmax:
                continue
            want = O.all_pairs_dense(p, ob, md)
            got = O.all_pairs_dense_simd(p, 0, n, ob, md, threads=3)
            np.testing.assert_array_equal(got, want, err_msg=f"L={L} md={md} ob={ob}")
            # bands starting past row 0: cells relative to the band's first cell
            for r0, r1 in [(1, 2), (2, 9), (9, n), (n - 1, n)]:
                first = r0 * (r0 - 1) // 2
                band = O.all_pairs_dense_simd(p, r0, r1, ob, md, threads=2)
                np.testing.assert_array_equal(band, want[first:r1 * (r1 - 1) // 2],
                                              err_msg=f"L={L} md={md} band {r0}:{r1}")
            if ref is not None and ob == 4:
                out = np.zeros_like(want)
                ref.ref_all_pairs(3, p.ctypes.data, n, p.shape[1], 0, n, 4, md,
                                  out.ctypes.data, 2, None)
                np.testing.assert_array_equal(out, want, err_msg=f"ref L={L} md={md}")


@pytest.mark.parametrize("alphabet", ["ACGT-NX", "AAAAAACGT", "ACGTX", "A-"])
def test_sparse_pair_equals_sparse_all_pairs(alphabet, rng):
    # The one-cell sparse oracle (lists against the whole dataset's consensus)
    # gives every cell of the all-pairs sparse route, under every cap.
    for n, L in [(2, 1), (9, 33), (40, 257)]:
        chars = O.as_chars(random_strings(rng, n, L, alphabet))
        for include_x in (False, True):
            cons = O.consensus(chars, include_x)
            I, J = O.tri_index(n)
            for md in (0, 1, 3, 65535):
                want = O.all_pairs_sparse(chars, include_x, 2, md)
                got = [O.distance_sparse_pair(chars[i], chars[j], cons, include_x, md)
                       for i, j in zip(I, J)]
                np.testing.assert_array_equal(np.array(got, np.uint16), want)


def test_simd_restatement_runs_avx512_here():
    # the restatement above is only pinned where it runs its AVX-512 path
    if not O.lib().oracle_have_avx512():
        pytest.skip("this CPU lacks AVX-512BW: the restatement falls back to the scalar kernel")
    assert O.lib().oracle_have_avx512() == 1


def test_dense_brute_force_nibble_rule():
    # every ordered pair of chars over a mixed alphabet: dense counts a site
    # exactly when the lookup codes share no bit
    alphabet = "ACGT-NX*a"
    lut = {"A": 1, "C": 2, "G": 4, "T": 8, "-": 15}
    for a, b in itertools.product(alphabet, repeat=2):
        pa, pb = O.pack(O.as_chars([a])), O.pack(O.as_chars([b]))
        want = int((lut.get(a, 0) & lut.get(b, 0)) == 0)
        assert O.distance_dense(pa[0], pb[0]) == want, (a, b)


def test_dense_invalid_site_lower_bound(rng):
    # The early-exit kernels settle a pair before computing it when
    # max(Z_i, Z_j) reaches the cap, Z the count of invalid (non ACGT-)
    # characters of a row: under the dense rule an invalid site counts against
    # every character, '-' and other invalid ones included.  The oracle's
    # dense distances respect the bound on data with invalid runs at shared
    # positions, and reach it for a row against itself (its letters and '-'
    # add nothing).
    n, L = 60, 400
    root = rng.choice(list("ACGT-"), size=L)
    rows = []
    for k in range(n):
        s = root.copy()
        s[rng.integers(0, L, size=k % 4)] = rng.choice(list("ACGT"))
        a = int(rng.integers(0, L))
        s[a:a + int(rng.integers(0, 40))] = rng.choice(list("N?X*"))
        s[rng.integers(0, L, size=int(rng.integers(0, 12)))] = "N"
        rows.append("".join(s))
    chars = O.as_chars(rows)
    packed = O.pack(chars)
    z = np.array([sum(c not in "ACGT-" for c in r) for r in rows])
    tri = O.all_pairs_dense(packed, 4, 2**31 - 1)
    i, j = O.tri_index(n)
    assert (tri >= np.maximum(z[i], z[j])).all()
    # invalid sites at the same position count once, so the sum of the two
    # counts is no bound: the kernels must not use it
    assert (tri < z[i] + z[j]).any()
    for r in rows[:10]:
        p = O.pack(O.as_chars([r, r]))
        assert O.distance_dense(p[0], p[1]) == sum(c not in "ACGT-" for c in r)


def test_all_pairs_dense_layout(rng):
    seqs = random_strings(rng, 40, 33, "ACGT-N")
    p = O.pack(O.as_chars(seqs))
    cells = O.all_pairs_dense(p, 4, 2**32 - 1)
    I, J = O.tri_index(40)
    for x in range(0, cells.size, 37):
        assert cells[x] == O.distance_dense(p[I[x]], p[J[x]])
        assert cells[x] == O.distance_char(seqs[I[x]], seqs[J[x]])


# ---- sparse route -------------------------------------------------------------

SPARSE_CODES = {"A": 0x11, "C": 0x22, "G": 0x44, "T": 0x88, "-": 0xFF}


def sparse_closed_form(a: str, b: str, include_x: bool) -> int:
    """Derived from distance_sparse + to_sparse_data: a site counts exactly
    when the lookup codes differ and neither is '-' (the consensus letter
    cancels out)."""
    codes = dict(SPARSE_CODES)
    if include_x:
        codes["X"] = 0x33
    d = 0
    for ca, cb in zip(a, b):
        va, vb = codes.get(ca, 0), codes.get(cb, 0)
        d += int(va != vb and va != 0xFF and vb != 0xFF)
    return d


def test_sparse_exhaustive_small_datasets():
    # all 3-sequence datasets of length 2 over a 6-char alphabet
    alphabet = "AC-NXG"
    words = ["".join(w) for w in itertools.product(alphabet, repeat=2)]
    rng = np.random.default_rng(9)
    for _ in range(400):
        seqs = [words[k] for k in rng.integers(0, len(words), size=3)]
        chars = O.as_chars(seqs)
        for include_x in (False, True):
            cells = O.all_pairs_sparse(chars, include_x, 4, 2**31 - 1)
            I, J = O.tri_index(3)
            want = [sparse_closed_form(seqs[i], seqs[j], include_x) for i, j in zip(I, J)]
            assert cells.tolist() == want, (seqs, include_x)


def test_sparse_reference_closed_forms():
    # src/hamming_impl_t.cc:49-104: identical -> 0; A/X vs G/T -> min(md, n)
    rng = np.random.default_rng(12345)
    for L in [1, 2, 7, 31, 32, 33, 255, 256, 1025, 8193]:
        for md in KATS["closed_form_max_dists"]:
            for include_x in (False, True):
                s = random_strings(rng, 1, L, "ACGT-X" if include_x else "ACGT-")[0]
                assert O.all_pairs_sparse(O.as_chars([s, s]), include_x, 4, md)[0] == 0
                for c1 in "AX":
                    for c2 in "GT":
                        got = O.all_pairs_sparse(O.as_chars([c1 * L, c2 * L]), include_x, 4, md)
                        assert got[0] == min(md, L)


def test_sparse_equals_dense_without_x_and_x_substitution(rng):
    # src/hamming_impl_t.cc:106-180
    for L in [1, 5, 16, 17, 63, 64, 65, 513, 2049]:
        for md in KATS["closed_form_max_dists"]:
            s1, s2 = random_strings(rng, 2, L, "ACGT-")
            p = O.pack(O.as_chars([s1, s2]))
            dense = O.distance_dense(p[0], p[1], md)
            for include_x in (False, True):
                sp = O.all_pairs_sparse(O.as_chars([s1, s2]), include_x, 4, md)[0]
                assert sp == dense
            for rc in "ACGT":
                x1, x2 = s1.replace(rc, "X"), s2.replace(rc, "X")
                assert O.all_pairs_sparse(O.as_chars([x1, x2]), True, 4, md)[0] == dense


# ---- heuristic --------------------------------------------------------------

def test_consensus_first_maximum():
    # ties resolve to the first of A, C, G, T (std::max_element); columns
    # without a letter read 'A'
    seqs = ["ACN-", "CAN-", "GTNT", "TGNT"]
    assert O.consensus(O.as_chars(seqs), False) == b"AAAT"
    seqs = ["XX", "XA", "CA"]
    assert O.consensus(O.as_chars(seqs), False) == b"CA"
    assert O.consensus(O.as_chars(seqs), True) == b"XA"


def test_heuristic_threshold_boundary():
    # frac_diff = n_diff / (n L) < 0.005 selects the sparse route.  n=10,
    # L=1000: 49 differing entries -> sparse, 50 -> dense.
    base = "A" * 1000
    for diffs, want in [(49, True), (50, False), (0, True), (1000, False)]:
        seqs = [base] * 10
        s = list(seqs[1])
        for k in range(diffs):
            s[k] = "C"
        seqs[1] = "".join(s)
        assert O.use_sparse(O.as_chars(seqs), False) == want, diffs
    assert O.use_sparse(O.as_chars([base, base]), True)


def test_routes_diverge_only_on_invalid_sites():
    # N vs N: dense (GPU route) counts 1, sparse counts 0
    seqs = ["A" * 400 + "N", "A" * 400 + "N", "A" * 401]
    cells, route = cpu_route(seqs)
    assert route == 1 and cells.tolist() == [0, 1, 1]
    assert gpu_route(seqs).tolist() == [1, 1, 1]


# ---- clamping (src/hamming_t.cc:299-318) ------------------------------------

def test_saturation_grid():
    sat = KATS["saturation"]
    for ob, tmax in [(1, 255), (2, 65535)]:
        for off in sat["n_offsets"]:
            L = tmax + off
            seqs = ["A" * L, "T" * L]
            for md in sat["max_dists"]:
                cells, _ = cpu_route(seqs, False, ob, md)
                assert int(cells[0]) == min(md, tmax), (ob, L, md)


def test_clamp_max_distance_casts():
    assert O.clamp_max_distance(255, 1) == 255
    assert O.clamp_max_distance(9999999, 1) == 255
    assert O.clamp_max_distance(9999999, 2) == 65535
    assert O.clamp_max_distance(-1, 1) == 255
    assert O.clamp_max_distance(-256, 1) == 0
    assert O.clamp_max_distance(0, 2) == 0


# ---- FASTA / dedup ------------------------------------------------------------

def write_records(path, records):
    with open(path, "w") as f:
        for i, rec in enumerate(records):
            f.write(f">seq{i}\n")
            for line in rec:
                f.write(line + "\n")


def test_dedup_kat(tmp_path):
    from oracle.fasta_oracle import read_fasta

    d = KATS["dedup"]
    path = str(tmp_path / "f.fa")
    write_records(path, [[s] for s in d["sequences"]])
    for n in d["ns"]:
        data, idx = read_fasta(path, True, n)
        assert idx == d["sequence_indices"]
        assert len(data) == 3
        for md in d["max_dists"]:
            cells, _ = cpu_route(data, False, 1, md)
            for key, want in d["unique_distances"].items():
                i, j = map(int, key.split(","))
                a, b = max(i, j), min(i, j)
                assert int(cells[a * (a - 1) // 2 + b]) == min(md, want, 255)


def test_fasta_reader_edge_cases(tmp_path):
    from oracle.fasta_oracle import read_fasta

    p = str(tmp_path / "f.fa")
    ml = KATS["multi_line"]
    write_records(p, ml["records"])
    data, _ = read_fasta(p, False, ml["n"])
    assert data == ["".join(r) for r in ml["records"]]
    # trailing header with newline -> one extra empty record; without -> none
    open(p, "w").write(">a\nAC\n>b\n")
    assert read_fasta(p)[0] == ["AC", ""]
    open(p, "w").write(">a\nAC\n>b")
    assert read_fasta(p)[0] == ["AC"]
    open(p, "w").write(">a\nAC\n\nGT\n>b\nTT")
    assert read_fasta(p)[0] == ["ACGT", "TT"]
    for n, want in [(0, 6), (2, 2), (3, 3), (8, 6)]:
        write_records(p, [[s] for s in KATS["from_fasta"]["sequences"]])
        assert len(read_fasta(p, False, n)[0]) == want


# ---- writers ----------------------------------------------------------------

def test_lower_triangular_writer_kat(tmp_path):
    lt = KATS["lower_triangular"]
    cells, _ = cpu_route(lt["sequences"], False, 2, 65535)
    for md in lt["max_distances"]:
        c = np.minimum(cells.astype(np.int64), md).astype(np.uint16)
        path = str(tmp_path / "lt.txt")
        O.lib().oracle_write_lower_triangular(path.encode(), c.ctypes.data, 3, 2)
        rows = [list(map(int, line.split(","))) for line in open(path).read().splitlines()]
        assert rows == [[min(v, md) for v in r] for r in lt["rows"]]


def test_writer_formats(tmp_path):
    cells = np.array([5, 0, 300], np.uint16)
    p = str(tmp_path / "x.txt")
    O.lib().oracle_write_lower_triangular(p.encode(), cells.ctypes.data, 3, 2)
    assert open(p).read() == "5\n0,300\n"
    O.lib().oracle_write_sparse(p.encode(), cells.ctypes.data, 3, 2, 5)
    assert open(p).read() == "1 0 5\n2 0 0\n"
    O.lib().oracle_write_full(p.encode(), cells.ctypes.data, 3, 2)
    assert open(p).read() == "0, 5, 0\n5, 0, 300\n0, 300, 0\n"
    assert O.lib().oracle_write_lower_triangular(p.encode(), cells.ctypes.data, 1, 2) == 0


# ---- reference distances -----------------------------------------------------

def test_reference_distance_matches_matrix_column(rng):
    # python/tests/test_hammingdist.py:90-136: every column of the matrix
    # equals the reference distances of that sequence (off the diagonal
    # without include_x, where invalid chars are 1 from themselves)
    for chars, include_x in [("A", False), ("ACGT-", False), ("ACGT-", True),
                             ("ACGT-X", False), ("ACGT-X", True)]:
        seqs = random_strings(rng, 30, 25, chars)
        cells, _ = O.distances(O.as_chars(seqs), include_x, False, 4, 2**31 - 1)
        for i in range(30):
            for j in range(30):
                d = O.distance_char(seqs[i], seqs[j], include_x)
                if include_x or i != j:
                    a, b = max(i, j), min(i, j)
                    got = 0 if i == j else int(cells[a * (a - 1) // 2 + b])
                    assert got == d, (chars, include_x, i, j)
