"""The pair functions of the all-pairs kernel, checked exhaustively on the CPU.

The tile kernel (hammingdist_amd/csrc/kernels.h) evaluates each semantics on
bit planes with a fixed chain of v_bitop3 instructions.  Here the chain is
evaluated bit by bit with the same truth tables for every pair of nibble
classes and compared with the reference's per-site rules
(src/hamming_impl.cc:131-140 for the dense rule; src/hamming_impl.cc:93-129
for the sparse rule: codes differ and neither is '-').  The GPU parity tests
check the kernels themselves; this file checks that the algebra they
implement is the reference's rule on every input class.
"""
import itertools

import pytest

NIBBLE = {"inv": 0, "A": 1, "C": 2, "G": 4, "T": 8, "X": 3, "-": 15}


def bitop3(x: int, y: int, z: int, imm: int) -> int:
    """v_bitop3_b32 on single bits: the result is bit (4x + 2y + z) of imm
    (truth tables written over src0 = 0xF0, src1 = 0xCC, src2 = 0xAA)."""
    return (imm >> (4 * x + 2 * y + z)) & 1


def dense_rule(a: str, b: str) -> int:
    return int((NIBBLE[a] & NIBBLE[b]) == 0)


def sparse_rule(a: str, b: str) -> int:
    return int(a != b and a != "-" and b != "-")


def dense3_planes(nib: int):
    """Planes (u, v, w, y) from the nibble bits, as the encode kernel derives
    them: u = {A, C, invalid}, v = {C, T, invalid}, w = {'-', invalid},
    y = {G, T}."""
    n0, n1, n2, n3 = (nib >> 0) & 1, (nib >> 1) & 1, (nib >> 2) & 1, (nib >> 3) & 1
    inv = 1 - (n0 | n1 | n2 | n3)
    u = (n0 ^ n1) | inv
    v = ((n1 | n3) & (1 - n0)) | inv
    w = (n0 & n1) | inv
    y = (n2 | n3) & (1 - n0)
    return u, v, w, y


DENSE3_TABLE = {"A": (1, 0, 0, 0), "C": (1, 1, 0, 0), "G": (0, 0, 0, 1), "T": (0, 1, 0, 1),
                "-": (0, 0, 1, 0), "inv": (1, 1, 1, 0)}


def dense3_site(a: str, b: str) -> int:
    """The kernel's chain on one site: row a's planes (u, v, w) against
    column b's (v, w, y)."""
    u_i, v_i, w_i, _ = DENSE3_TABLE[a]
    _, v_j, w_j, y_j = DENSE3_TABLE[b]
    s = v_i ^ v_j
    s = bitop3(u_i, y_j, s, 0x14)
    return bitop3(w_i, w_j, s, 0x4D)


def test_dense3_encode_matches_table():
    for cls, planes in DENSE3_TABLE.items():
        assert dense3_planes(NIBBLE[cls]) == planes, cls


def test_dense3_counts_from_planes():
    # the encoder counts invalid sites as u & v & w and '-' sites as w & ~u
    for cls, (u, v, w, y) in DENSE3_TABLE.items():
        assert (u & v & w) == (cls == "inv"), cls
        assert (w & (1 - u)) == (cls == "-"), cls


def test_dense3_chain_with_corrections_is_the_dense_rule():
    classes = list(DENSE3_TABLE)
    for a, b in itertools.product(classes, classes):
        site = dense3_site(a, b) + (a == "inv") - (b == "-")
        assert site == dense_rule(a, b), (a, b)


def test_dense3_partial_sums_never_exceed_the_distance_bound():
    # Per site the term is s3 + [row invalid] - [column '-'] = the dense
    # rule (0 or 1), so every prefix of sites sums to the exact partial
    # distance; the early-exit vote relies on this.
    classes = list(DENSE3_TABLE)
    for a, b in itertools.product(classes, classes):
        assert 0 <= dense3_site(a, b) + (a == "inv") - (b == "-") <= 1


def test_dense3_chain_catches_a_wrong_table():
    # a plausible slip (a neighbouring truth table, a wrong plane index, the
    # chain's operands in another order) breaks the rule on some pair, so the
    # checks above pin the tables and the operand order
    classes = list(DENSE3_TABLE)

    def rule_holds(t1, t2, slip=None):
        for a, b in itertools.product(classes, classes):
            u_i, v_i, w_i, _ = DENSE3_TABLE[a]
            _, v_j, w_j, y_j = DENSE3_TABLE[b]
            if slip == "column v for y":
                y_j = v_j
            if slip == "row v for u":
                u_i = v_i
            s = v_i ^ v_j
            if slip == "w before u":
                s = bitop3(u_i, y_j, bitop3(w_i, w_j, s, t2), t1)
            else:
                s = bitop3(w_i, w_j, bitop3(u_i, y_j, s, t1), t2)
            if slip == "w operands swapped":
                s = bitop3(w_j, w_i, bitop3(u_i, y_j, v_i ^ v_j, t1), t2)
            if s + (a == "inv") - (b == "-") != dense_rule(a, b):
                return False
        return True

    assert rule_holds(0x14, 0x4D)
    for slip in ("column v for y", "row v for u", "w before u", "w operands swapped"):
        assert not rule_holds(0x14, 0x4D, slip), slip
    for bit in range(8):
        assert not rule_holds(0x14 ^ (1 << bit), 0x4D)
        assert not rule_holds(0x14, 0x4D ^ (1 << bit))


def class_code_table(present, bits):
    """Class code of each class present (codes 0, 1, ... in the engine's
    order: invalid, A, C, G, T, X) and '-' as the plane after the code bits."""
    order = [c for c in ("inv", "A", "C", "G", "T", "X") if c in present]
    table = {c: [(k >> b) & 1 for b in range(bits)] + [0] for k, c in enumerate(order)}
    table["-"] = [0] * bits + [1]
    return table


@pytest.mark.parametrize("present", [("A", "C", "G", "T"),
                                     ("inv", "A", "C", "G", "T"),
                                     ("A", "C", "G", "T", "X"),
                                     ("inv", "A", "C", "G", "T", "X"),
                                     ("inv", "X")])
def test_class_code_chain_is_the_sparse_rule(present):
    # kClass3, the code of rows of 2^28 sites or more: always 3 code bits
    bits = 3
    table = class_code_table(present, bits)
    classes = list(present) + ["-"]
    for a, b in itertools.product(classes, classes):
        pa, pb = table[a], table[b]
        s = pa[0] ^ pb[0]
        for p in range(1, bits):
            s = bitop3(pa[p], pb[p], s, 0xBE)
        s = bitop3(pa[bits], pb[bits], s, 0x02)
        assert s == sparse_rule(a, b), (present, a, b)


def test_one_hot_chain_is_the_dense_rule_on_nibbles():
    # The 4-plane nibble code: s = a0 & b0, (ap & bp) | s, ~((a3 & b3) | s).
    for x, y in itertools.product(range(16), range(16)):
        a = [(x >> p) & 1 for p in range(4)]
        b = [(y >> p) & 1 for p in range(4)]
        s = a[0] & b[0]
        for p in (1, 2):
            s = bitop3(a[p], b[p], s, 0xEA)
        s = bitop3(a[3], b[3], s, 0x15)
        assert s == int((x & y) == 0)


def gap_code_table(present):
    """kGap2 / kGap3: codes 0, 1, ... for the classes present (invalid, A, C,
    G, T, X) and then '-'; 2 code bits for up to 4 codes, else 3; the plane
    after the code bits marks '-'."""
    order = [c for c in ("inv", "A", "C", "G", "T", "X") if c in present] + ["-"]
    bits = 2 if len(order) <= 4 else 3
    table = {c: [(k >> b) & 1 for b in range(bits)] + [int(c == "-")]
             for k, c in enumerate(order)}
    return table, bits


@pytest.mark.parametrize("present", [("A", "C", "G"),
                                     ("A", "C", "G", "T"),
                                     ("inv", "A", "C", "G", "T"),
                                     ("A", "C", "G", "T", "X"),
                                     ("inv", "A", "C", "G", "T", "X"),
                                     ("inv", "X")])
def test_gap_code_chain_with_corrections_is_the_sparse_rule(present):
    # Per site: f = codes differ (XOR, then (x ^ y) | s per further bit);
    # f - [row '-'] - [column '-'] + 2 [both '-'] is the sparse rule, and
    # without the last term it never exceeds it (the early-exit votes use
    # that lower bound).
    table, bits = gap_code_table(present)
    assert len(set(tuple(v[:bits]) for v in table.values())) == len(table)
    classes = list(present) + ["-"]
    for a, b in itertools.product(classes, classes):
        pa, pb = table[a], table[b]
        s = pa[0] ^ pb[0]
        for p in range(1, bits):
            s = bitop3(pa[p], pb[p], s, 0xBE)
        g_a, g_b = pa[bits], pb[bits]
        assert s - g_a - g_b + 2 * (g_a & g_b) == sparse_rule(a, b), (present, a, b)
        assert s - g_a - g_b <= sparse_rule(a, b)
