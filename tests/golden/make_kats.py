"""Writes tests/golden/reference_kats.json: the known-answer vectors that the
reference's own tests hold for the all-pairs path, as data (inputs and
expected outputs).  Each entry names the reference test it comes from.
Nothing here runs the reference; the values are the literal inputs and
expected results of those tests.

    python tests/golden/make_kats.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))

KATS = {
    # src/hamming_t.cc:9-11
    "valid_chars": ["A", "C", "G", "T"],
    "invalid_chars": [" ", "N", "*", "?", "a", "."],
    # src/hamming_t.cc:28-91: per-char distances (both from_stringlist and
    # distance()): equal valid 0, valid vs '-' 0, different valid 1,
    # valid vs invalid 1, invalid vs invalid 1, invalid vs '-' 1
    "char_rules": {"equal_valid": 0, "valid_gap": 0, "different_valid": 1,
                   "valid_invalid": 1, "invalid_invalid": 1, "invalid_gap": 1},
    # src/hamming_t.cc:93-127: every pair has distance exactly 2
    "distance_two_pairs": [
        ["AC", "CA"], ["AC", "TG"], ["ACG", "AGT"], ["ACG", "-TT"], ["ACG", "T-T"],
        ["ACG", "TA-"], ["ACG", "CCC"], ["ACGT", "AGGG"],
        ["ACGTGTCGTGTCGACGTGTCG", "ACGTGTCGTTTCGACGAGTCG"],
        ["ACGTGTCGTGTCGACGTGTCGT", "ACGTGTCGTTTCGACGAGTCGT"],
        ["ACGTGTCGTGTCGACGTGTCGT-", "ACGTGTCGTTTCGACGAGTCGTA"],
        ["ACGTGTCGTGTCGACGTGTCG----------", "ACGTGTCGTTTCGACGAGTCGGGG-------"],
        ["ACGTGTCGTGTCGACGTGTCG----------A", "ACGTGTCGTTTCGACGAGTCGGGG-------A"],
        ["ACGTGTCGTGTCGACGTGTCG----------AG", "ACGTGTCGTTTCGACGAGTCGGGG-------AG"],
        ["ACGTGTCGTGTCGACGTGTCG---ACGTGTCGTGTCGACGTGTCG---ACGTGTCGTGTCGACGTGTCG---",
         "ACGTGTCGTTTCGACGAGTCGGGGACGTGTCGTGTCGACGTGTCG---ACGTGTCGTGTCGACGTGTCG---"],
    ],
    # src/hamming_t.cc:165-215: dedup FASTA, indices, and the clamped 3x3
    # matrix d[0,1]=2, d[0,2]=1, d[1,2]=2 for max_dist {0,1,2,3,999999999}
    # and n in {0, 6, 22}
    "dedup": {
        "sequences": ["ACGTGTCGTGTCGACGTGTCG", "ACGTGTCGTTTCGACGAGTCG",
                      "ACGTGTCGTTTCGACGAGTCG", "ACGTGTCGTTTCGACGAGTCG",
                      "ACGTGTCGTGTCGACGTGTCG", "ACGTGTCGTATCGACGTGTCG"],
        "sequence_indices": [0, 1, 1, 1, 0, 2],
        "unique_distances": {"0,1": 2, "0,2": 1, "1,2": 2},
        "max_dists": [0, 1, 2, 3, 999999999],
        "ns": [0, 6, 22],
    },
    # src/hamming_t.cc:129-163, 217-252: single- and multi-line records
    "single_line": {"records": [["ACGTGTCGTGTCGACGTGTCG"], ["ACGTGTCGTTTCGACGAGTCG"]],
                    "distance": 2, "ns": [0, 2, 3, 8]},
    "multi_line": {"records": [["ACGTGTCGTGTCGACGTGTCG", "ACGTGTCGTGTCGACGTGTCG", "ACGTGTCGTGTCG"],
                               ["ACGTGTCGTTTCGACGAGTCG", "ACGTGTCGTGTCGACGTGTCG", "ACGTGTCGTGTCG"]],
                   "distance": 2, "n": 2},
    # src/hamming_t.cc:254-276
    "errors": {
        "empty": [[], ["", ""]],
        "empty_message": "Error: Empty sequence",
        "ragged": [["ACGT", "A"], ["AC", "ACGTCG"], ["ACGT", "ACGT", "ACGT", "ACGT", "ACGT", "A"],
                   ["ACGT", "A-----"], ["ACGT", ""]],
        "ragged_message": "Error: Sequences do not all have the same length",
    },
    # src/hamming_t.cc:299-318: n x 'A' vs n x 'T' saturates at
    # min(max_dist, type max) for n in {tmax, tmax+1, tmax+99}
    "saturation": {"max_dists": [0, 1, 2, 3, 7, 21, 886, 2784904, 99999999],
                   "type_max": {"1": 255, "2": 65535}, "n_offsets": [0, 1, 99]},
    # python/tests/test_hammingdist.py:139-146
    "distance_kats": [["ACGT", "ACCT", False, 1], ["ACGTX", "ACCTX", False, 2],
                      ["ACGTX", "ACCTX", True, 1]],
    # python/tests/test_hammingdist.py:149-174: clamped lower triangle
    "lower_triangular": {
        "sequences": ["ACGTGTCGTGTCGACGTGTCGCAGGTGTCGACGTGTCGCAGGTGTCGACGTGTCGCAG",
                      "CCGTGTCGTGTCGACGTGTCGC-GGTGTCGACGTGTCGCAGGTGTCGACGTGTCGCAG",
                      "CAGTGT-GTGTCGACGTGTCGCAGGTGTCGACGTGTCGCAGGTGTCGACGTG--GCAG"],
        "rows": [[1], [2, 1]],
        "max_distances": [0, 1, 2, 3, 89, 497, 9999999],
    },
    # python/tests/test_hammingdist.py:39-87: sizes for the 6-record file
    "from_fasta": {
        "sequences": ["ACGTGTCGTGTCGACGTGTCG", "ACGTGTCGTTTCGACGAGTCG", "ACGTGTCGTTTCGACGAGTCG",
                      "ACGTGTCGTTTCGACGAGTCG", "ACGTGACGTGTCGACGTGTCG", "ACGTGXCGTGTCGACGTGTCG"],
        # (n, remove_duplicates) -> (inputs, rows)
        "cases": [[0, False, 6, 6], [5, False, 5, 5], [0, True, 6, 4], [2, False, 2, 2],
                  [3, True, 3, 2], [5, True, 5, 3]],
    },
    # src/hamming_impl_t.cc:9-18 (lengths used by the closed-form kernel tests)
    "closed_form_lengths": [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19,
                            20, 31, 32, 33, 63, 64, 65, 127, 128, 129, 254, 255, 256, 511, 512,
                            513, 1023, 1024, 1025, 2047, 2048, 2049, 4095, 4096, 4097, 8191,
                            8192, 8193, 32767, 32768, 32769, 65535, 65536, 65537, 131071,
                            131072, 131073, 262143, 262144, 262145, 524287, 524288, 524289,
                            1048575, 1048576, 1048577],
    "closed_form_max_dists": [0, 1, 2, 11, 999, 9876544],
    # src/hamming_t.cc:346-451 (shapes of the CPU==GPU consistency tests)
    "consistency": {"lengths": [1, 5, 13, 17, 32, 88, 89, 185, 381, 497, 1023, 1092],
                    "samples": [2, 3, 4, 7, 11, 32, 33, 127, 128, 255, 256, 257, 689, 703],
                    "max_dists": [0, 1, 2, 3, 89, 497, 9999999]},
}

if __name__ == "__main__":
    with open(os.path.join(HERE, "reference_kats.json"), "w") as f:
        json.dump(KATS, f, indent=1)
        f.write("\n")
