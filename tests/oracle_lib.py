"""Loader of the oracle (oracle/liboracle.so) and of the reference's own SIMD
kernels (oracle/_ref/libref.so) for the tests.  Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libref.so")

vp, u64, i64 = C.c_void_p, C.c_uint64, C.c_int64

_SIGS = {
    "oracle_bytes_per_seq": (u64, [u64]),
    "oracle_pack": (None, [vp, u64, C.c_int, vp, u64]),
    "oracle_distance_dense": (C.c_int, [vp, vp, u64, C.c_int]),
    "oracle_distance_char": (C.c_uint32, [vp, vp, u64, C.c_int]),
    "oracle_reference_distance": (C.c_uint32, [vp, vp, u64, C.c_int]),
    "oracle_consensus": (None, [vp, u64, u64, C.c_int, vp]),
    "oracle_clamp_max_distance": (C.c_uint32, [i64, C.c_int]),
    "oracle_use_sparse": (C.c_int, [vp, u64, u64, C.c_int]),
    "oracle_all_pairs_sparse": (None, [vp, u64, u64, C.c_int, C.c_int, C.c_uint32, vp]),
    "oracle_all_pairs_dense": (None, [vp, u64, u64, C.c_int, C.c_uint32, vp]),
    "oracle_distance_sparse_pair": (C.c_int, [vp, vp, vp, u64, C.c_int, C.c_uint32]),
    "oracle_distances": (C.c_int, [vp, u64, u64, C.c_int, C.c_int, C.c_int, i64, vp]),
    "oracle_have_avx512": (C.c_int, []),
    "oracle_all_pairs_dense_simd": (C.c_int, [vp, u64, u64, u64, u64, C.c_int, C.c_uint32,
                                              vp, C.c_int]),
    "oracle_write_lower_triangular": (i64, [C.c_char_p, vp, u64, C.c_int]),
    "oracle_write_sparse": (i64, [C.c_char_p, vp, u64, C.c_int, i64]),
    "oracle_write_full": (i64, [C.c_char_p, vp, u64, C.c_int]),
}

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        _lib = C.CDLL(ORACLE)
        for name, (res, args) in _SIGS.items():
            f = getattr(_lib, name)
            f.restype = res
            f.argtypes = args
    return _lib


def ref():
    """The reference's SIMD kernels, or None where oracle/_ref was not built."""
    global _ref
    if _ref is None and os.path.exists(REF):
        _ref = C.CDLL(REF)
        _ref.ref_distance.restype = C.c_int
        _ref.ref_distance.argtypes = [C.c_int, vp, vp, u64, C.c_int]
        _ref.ref_all_pairs.restype = C.c_int
        _ref.ref_all_pairs.argtypes = [C.c_int, vp, u64, u64, u64, u64, C.c_int, C.c_int, vp,
                                       C.c_int, C.POINTER(C.c_double)]
    return _ref


DTYPE = {1: np.uint8, 2: np.uint16, 4: np.uint32}


def as_chars(seqs) -> np.ndarray:
    """list of equal-length str -> (n, L) uint8"""
    if len(seqs) == 0:
        return np.zeros((0, 0), np.uint8)
    L = len(seqs[0])
    arr = np.frombuffer("".join(seqs).encode("latin-1"), dtype=np.uint8)
    return arr.reshape(len(seqs), L).copy()


def bytes_per_seq(L: int) -> int:
    return int(lib().oracle_bytes_per_seq(L))


def pack(chars: np.ndarray, include_x: bool = False) -> np.ndarray:
    n, L = chars.shape
    bps = bytes_per_seq(L)
    out = np.empty((n, bps), np.uint8)
    c = np.ascontiguousarray(chars)
    for i in range(n):
        lib().oracle_pack(c[i].ctypes.data, L, int(include_x), out[i].ctypes.data, bps)
    return out


def distances(chars: np.ndarray, include_x: bool, gpu_route: bool, out_bytes: int,
              max_distance: int):
    """distances<T> of the reference; returns (cells, route: 0 dense / 1 sparse)."""
    n, L = chars.shape
    c = np.ascontiguousarray(chars)
    out = np.zeros(n * (n - 1) // 2, DTYPE[out_bytes])
    route = lib().oracle_distances(c.ctypes.data, n, L, int(include_x), int(gpu_route),
                                   out_bytes, max_distance, out.ctypes.data)
    return out, route


def all_pairs_dense(packed: np.ndarray, out_bytes: int, max_dist: int) -> np.ndarray:
    n, bps = packed.shape
    out = np.zeros(n * (n - 1) // 2, DTYPE[out_bytes])
    p = np.ascontiguousarray(packed)
    lib().oracle_all_pairs_dense(p.ctypes.data, n, bps, out_bytes, max_dist, out.ctypes.data)
    return out


def all_pairs_sparse(chars: np.ndarray, include_x: bool, out_bytes: int, max_dist: int):
    n, L = chars.shape
    out = np.zeros(n * (n - 1) // 2, DTYPE[out_bytes])
    c = np.ascontiguousarray(chars)
    lib().oracle_all_pairs_sparse(c.ctypes.data, n, L, int(include_x), out_bytes, max_dist,
                                  out.ctypes.data)
    return out


def distance_char(a: str, b: str, include_x: bool = False) -> int:
    ba, bb = a.encode("latin-1"), b.encode("latin-1")
    assert len(ba) == len(bb)
    return int(lib().oracle_distance_char(ba, bb, len(ba), int(include_x)))


def distance_dense(pa: np.ndarray, pb: np.ndarray, max_dist: int = 2**31 - 1) -> int:
    return int(lib().oracle_distance_dense(pa.ctypes.data, pb.ctypes.data, pa.size, max_dist))


def consensus(chars: np.ndarray, include_x: bool) -> bytes:
    n, L = chars.shape
    out = C.create_string_buffer(max(L, 1))
    c = np.ascontiguousarray(chars)
    lib().oracle_consensus(c.ctypes.data, n, L, int(include_x), out)
    return out.raw[:L]


def use_sparse(chars: np.ndarray, include_x: bool) -> bool:
    n, L = chars.shape
    c = np.ascontiguousarray(chars)
    return bool(lib().oracle_use_sparse(c.ctypes.data, n, L, int(include_x)))


def clamp_max_distance(max_distance: int, out_bytes: int) -> int:
    return int(lib().oracle_clamp_max_distance(max_distance, out_bytes))


def tri_index(n: int):
    """(i, j) arrays of the linear lower triangle, i > j."""
    i, j = np.tril_indices(n, -1)
    order = np.lexsort((j, i))
    return i[order], j[order]


def distance_sparse_pair(a: np.ndarray, b: np.ndarray, cons: bytes, include_x: bool,
                         max_dist: int = 2**32 - 1) -> int:
    """One cell of the sparse route: rows a, b (uint8 chars) against the
    dataset consensus `cons` (from consensus() over the whole dataset)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return int(lib().oracle_distance_sparse_pair(a.ctypes.data, b.ctypes.data, cons, a.size,
                                                 int(include_x), max_dist))


def all_pairs_dense_simd(packed: np.ndarray, row0: int, row1: int, out_bytes: int,
                         max_dist: int, threads: int = 0) -> np.ndarray:
    """Rows [row0, row1) of the triangle through the oracle's restatement of
    the reference's AVX-512 pair kernel (scalar kernel where the CPU lacks
    AVX-512BW)."""
    n, bps = packed.shape
    first = row0 * (row0 - 1) // 2 if row0 else 0
    out = np.zeros(row1 * (row1 - 1) // 2 - first if row1 else 0, DTYPE[out_bytes])
    p = np.ascontiguousarray(packed)
    lib().oracle_all_pairs_dense_simd(p.ctypes.data, n, bps, row0, row1, out_bytes, max_dist,
                                      out.ctypes.data, threads)
    return out


def sample_cells(rng: np.random.Generator, n: int, k: int) -> np.ndarray:
    """k linear triangle indices spread over the whole triangle of n rows:
    uniform draws plus the first cells, the whole last row's ends, and the
    cells around 2^31 and 2^32 where the triangle reaches them."""
    cells = n * (n - 1) // 2
    picks = [rng.integers(0, cells, size=k, dtype=np.uint64)]
    fixed = [0, 1, 2, cells - 1, cells - 2, cells - (n - 1), cells - (n - 1) + 1]
    for edge in (2**31, 2**32):
        if cells > edge + 2:
            fixed += [edge - 2, edge - 1, edge, edge + 1]
    picks.append(np.array([x for x in fixed if 0 <= x < cells], np.uint64))
    return np.unique(np.concatenate(picks))


def cell_rows(x: np.ndarray):
    """(i, j) of linear triangle indices x (i > j, cell at i(i-1)/2 + j)."""
    x = np.asarray(x, np.uint64)
    i = np.floor(np.sqrt(2.0 * x.astype(np.float64) + 0.25) + 0.5).astype(np.uint64)
    # exact integer correction of the floating-point row
    for _ in range(2):
        i = np.where(i * (i - 1) // 2 > x, i - 1, i)
        i = np.where((i + 1) * i // 2 <= x, i + 1, i)
    j = x - i * (i - 1) // 2
    return i.astype(np.int64), j.astype(np.int64)
