"""The text loaders: from_csv (include/hamming/hamming.hh:34-55) and
from_lower_triangular(_large) (include/hamming/hamming.hh:152-166).

The loaders' oracle (oracle/loaders_oracle.cc, the reference's stream loop
restated) is pinned first: round trips through the oracle writers (what a
writer writes, the loader reads back), the reference's own round-trip and
KAT (src/hamming_t.cc:278-297, python/tests/test_hammingdist.py:149-174),
and cases whose answer the standard library's stream and std::stoi rules fix
by hand.  Then the engine's parallel loaders are compared with it on those
cases and on random files, including the error each one raises.  CPU only:
the loaders compute nothing on the GPU."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import hammingdist
import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
KATS = json.load(open(os.path.join(HERE, "golden", "reference_kats.json")))
DT = {1: np.uint8, 2: np.uint16}


def oracle_lib():
    lib = O.lib()
    lib.oracle_from_csv.restype = C.c_int
    lib.oracle_from_csv.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_void_p),
                                    C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    lib.oracle_from_lower_triangular.restype = C.c_int
    lib.oracle_from_lower_triangular.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_void_p),
                                                 C.POINTER(C.c_uint64)]
    lib.oracle_free.argtypes = [C.c_void_p]
    return lib


ERRORS = {1: ValueError, 2: IndexError}  # std::invalid_argument, std::out_of_range


def oracle_csv(path):
    lib = oracle_lib()
    p, nc, ns = C.c_void_p(), C.c_uint64(), C.c_uint64()
    err = lib.oracle_from_csv(str(path).encode(), 1, C.byref(p), C.byref(nc), C.byref(ns))
    if err:
        return err, None, None
    out = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), (max(nc.value, 1),))
    cells = out[: nc.value].copy()
    lib.oracle_free(p)
    return 0, cells, ns.value


def oracle_lt(path, out_bytes):
    lib = oracle_lib()
    p, nc = C.c_void_p(), C.c_uint64()
    err = lib.oracle_from_lower_triangular(str(path).encode(), out_bytes, C.byref(p),
                                           C.byref(nc))
    if err:
        return err, None
    ct = C.c_uint8 if out_bytes == 1 else C.c_uint16
    out = np.ctypeslib.as_array(C.cast(p, C.POINTER(ct)), (max(nc.value, 1),))
    cells = out[: nc.value].copy()
    lib.oracle_free(p)
    return 0, cells


def write(tmp_path, text, name="f.txt"):
    p = tmp_path / name
    p.write_bytes(text.encode() if isinstance(text, str) else text)
    return p


# ---- pinning the loaders' oracle ---------------------------------------------

@pytest.mark.parametrize("n", [2, 3, 17, 200])
@pytest.mark.parametrize("ob", [1, 2])
def test_oracle_round_trips_through_the_writers(tmp_path, n, ob, rng):
    # what the lower-triangle writer writes the loader reads back, and the
    # full-matrix dump reads back as the lower triangle plus the reference's
    # n trailing zero cells (result sized n(n+1)/2)
    cells = rng.integers(0, 256 if ob == 1 else 65536, size=n * (n - 1) // 2).astype(DT[ob])
    lt = tmp_path / "lt.txt"
    O.lib().oracle_write_lower_triangular(str(lt).encode(), cells.ctypes.data, n, ob)
    err, back = oracle_lt(lt, ob)
    assert err == 0 and np.array_equal(back, cells)
    if ob == 1:
        full = tmp_path / "full.txt"
        O.lib().oracle_write_full(str(full).encode(), cells.ctypes.data, n, 1)
        err, back, ns = oracle_csv(full)
        assert err == 0 and ns == n
        assert np.array_equal(back, np.concatenate([cells, np.zeros(n, np.uint8)]))


def test_oracle_reference_kat(tmp_path):
    # python/tests/test_hammingdist.py:149-174: rows [1] and [2, 1]
    lt = KATS["lower_triangular"]
    text = "".join(",".join(map(str, r)) + "\n" for r in lt["rows"])
    err, back = oracle_lt(write(tmp_path, text), 1)
    assert err == 0 and back.tolist() == [v for r in lt["rows"] for v in r]


# (text, from_lower_triangular uint8 cells or error, from_csv cells or error),
# each answer fixed by the stream rules: a field after the line's end repeats
# the last one (std::getline leaves d unchanged once eof is set), a field
# after a trailing ',' is empty, std::stoi skips leading space, ignores
# trailing text, needs a digit, and fails outside int.
HAND_CASES = [
    ("1\n2,3\n", [1, 2, 3], None),
    ("1\n2,3", [1, 2, 3], None),            # no final '\n': still a line
    ("7\n5\n9\n", [7, 5, 9], [5, 9, 9, 0, 0, 0]),   # csv: row 2 repeats its only field
    ("x\n4,5\n", 1, [4, 0, 0]),             # csv skips line 0 entirely
    ("1,2,\n", 1, None),                    # empty field after a trailing ','
    ("1\n\n2\n", 1, 1),                     # an empty line is one empty field
    (" 12abc, -1,\t+7\r\n", [12, 255, 7], None),  # space, sign, trailing text, wrap
    ("300,70000\n", [255, 255], None),      # safe_int_cast caps at 255
    ("2147483647\n", [255], None),
    ("2147483648\n", 2, None),              # past int: out_of_range
    ("-2147483648\n", [0], None),           # INT_MIN wraps to 0 in uint8
    ("-2147483649\n", 2, None),
    ("+\n", 1, None),
    ("0x1F\n", [0], None),                  # base 10: reads the 0
    ("", [], []),
]


@pytest.mark.parametrize("text,lt_want,_", HAND_CASES)
def test_oracle_hand_cases_lower_triangular(tmp_path, text, lt_want, _):
    err, back = oracle_lt(write(tmp_path, text), 1)
    if isinstance(lt_want, int):
        assert err == lt_want
    else:
        assert err == 0 and back.tolist() == lt_want


@pytest.mark.parametrize("text,_,csv_want", [c for c in HAND_CASES if c[2] is not None])
def test_oracle_hand_cases_csv(tmp_path, text, _, csv_want):
    err, back, ns = oracle_csv(write(tmp_path, text))
    if isinstance(csv_want, int):
        assert err == csv_want
    else:
        assert err == 0 and back.tolist() == csv_want and ns == text.count("\n")


# ---- the engine's loaders against the oracle ---------------------------------

def engine_lt(path, ob):
    fn = hammingdist.from_lower_triangular if ob == 1 else hammingdist.from_lower_triangular_large
    return fn(str(path))


def check_same(tmp_path, text, name="f.txt"):
    p = write(tmp_path, text, name)
    for ob in (1, 2):
        err, want = oracle_lt(p, ob)
        if err:
            with pytest.raises(ERRORS[err]):
                engine_lt(p, ob)
        else:
            ds = engine_lt(p, ob)
            assert ds.lt_array.dtype == DT[ob]
            np.testing.assert_array_equal(ds.lt_array, want)
            n = ds.nsamples
            assert n == (int(round((8 * want.size + 1) ** 0.5)) + 1) // 2
    err, want, ns = oracle_csv(p)
    if err:
        with pytest.raises(ERRORS[err]):
            hammingdist.from_csv(str(p))
    else:
        ds = hammingdist.from_csv(str(p))
        assert ds.nsamples == ns
        np.testing.assert_array_equal(ds.lt_array, want)


@pytest.mark.parametrize("case", range(len(HAND_CASES)))
def test_engine_hand_cases(tmp_path, case):
    check_same(tmp_path, HAND_CASES[case][0])


def test_engine_missing_file_reads_empty(tmp_path):
    missing = str(tmp_path / "absent.txt")
    assert hammingdist.from_lower_triangular(missing).lt_array.size == 0
    assert hammingdist.from_csv(missing).lt_array.size == 0


@pytest.mark.parametrize("n", [2, 5, 64, 700])
def test_engine_round_trips(tmp_path, n, rng):
    for ob in (1, 2):
        cells = rng.integers(0, 256 if ob == 1 else 65536,
                             size=n * (n - 1) // 2).astype(DT[ob])
        lt = tmp_path / "lt.txt"
        O.lib().oracle_write_lower_triangular(str(lt).encode(), cells.ctypes.data, n, ob)
        ds = engine_lt(lt, ob)
        assert ds.nsamples == n
        np.testing.assert_array_equal(ds.lt_array, cells)
    cells = rng.integers(0, 256, size=n * (n - 1) // 2).astype(np.uint8)
    full = tmp_path / "full.txt"
    O.lib().oracle_write_full(str(full).encode(), cells.ctypes.data, n, 1)
    check_same(tmp_path, full.read_bytes(), "full2.txt")


def test_engine_random_text(tmp_path, rng):
    # random lines of random fields, some malformed, spread over many
    # threads' ranges (several MiB): values, errors and the first error agree
    pieces = ["0", "1", "255", "256", "65535", "70000", "-1", " 3", "4x", "", "+", "12\r",
              "2147483648"]
    for trial in range(12):
        lines = []
        bad = trial % 3 == 0
        for r in range(int(rng.integers(1, 3000))):
            k = int(rng.integers(1, 40))
            choice = rng.integers(0, len(pieces) if bad else 8, size=k)
            lines.append(",".join(pieces[c] for c in choice))
        text = "\n".join(lines) + ("\n" if trial % 2 else "")
        check_same(tmp_path, text, f"r{trial}.txt")


def test_engine_large_file_many_threads(tmp_path, rng):
    # a 20 MB lower triangle: the byte-balanced split across threads
    n = 2600
    cells = rng.integers(0, 65536, size=n * (n - 1) // 2).astype(np.uint16)
    lt = tmp_path / "big.txt"
    O.lib().oracle_write_lower_triangular(str(lt).encode(), cells.ctypes.data, n, 2)
    ds = hammingdist.from_lower_triangular_large(str(lt))
    np.testing.assert_array_equal(ds.lt_array, cells)
    ds8 = hammingdist.from_lower_triangular(str(lt))
    np.testing.assert_array_equal(ds8.lt_array, np.minimum(cells, 255).astype(np.uint8))
