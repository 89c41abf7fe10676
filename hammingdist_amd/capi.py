"""ctypes binding of the engine's C ABI (include/hammingdist_amd.h).

This is the same boundary a C, Go or Rust caller binds; tests and the
benchmark use it to drive the engine directly (device pointers come from
torch tensors there).  Every call raises ``EngineError`` with the engine's
message on a non-zero status.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "libhd_engine.so")

HD_OK = 0
HD_ERR_INVALID = 1
HD_ERR_NO_DEVICE = 2
HD_ERR_HIP = 3
HD_ERR_CALLBACK = 4
HD_ERR_CAPACITY = 5

HD_SEM_DENSE = 0
HD_SEM_SPARSE = 1
HD_SEM_SPARSE_X = 2
HD_SEM_CHAR_X = 3

HD_REQ_PENDING = 0
HD_REQ_ADMITTED = 1
HD_REQ_RUNNING = 2
HD_REQ_DONE = 3
HD_REQ_FAILED = 4

u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
i32p = C.POINTER(C.c_int32)
intp = C.POINTER(C.c_int)
vp = C.c_void_p

BAND_CB = C.CFUNCTYPE(C.c_int, vp, vp, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64)

# name -> (restype, argtypes); the declarations of include/hammingdist_amd.h
SIGNATURES = {
    "hd_last_error": (C.c_char_p, []),
    "hd_device_count": (C.c_int, []),
    "hd_device_arch": (C.c_int, [C.c_char_p, C.c_size_t]),
    "hd_distance_pair": (C.c_int, [vp, vp, C.c_uint64, C.c_int, C.c_int32, i32p]),
    "hd_all_pairs": (C.c_int, [vp, C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.c_uint32, vp]),
    "hd_all_pairs_stream": (C.c_int, [vp, C.c_uint64, C.c_uint64, C.c_int, C.c_int,
                                      C.c_uint32, C.c_uint64, BAND_CB, vp]),
    "hd_reference_distances": (C.c_int, [vp, vp, C.c_uint64, C.c_uint64, vp, C.c_uint64,
                                         C.c_int, vp]),
    "hd_consensus_diff": (C.c_int, [vp, C.c_uint64, C.c_uint64, C.c_uint64, u64p, u32p]),
    "hd_sparse_emit": (C.c_int, [vp, C.c_uint64, C.c_int, C.c_int64, vp, C.c_uint64, u64p]),
    "hd_sparse_emit_dev": (C.c_int, [vp, C.c_uint64, C.c_int, C.c_int64, vp, C.c_uint64, u64p,
                                     vp]),
    "hd_sparse_emit_dev_noalloc": (C.c_int, [vp, C.c_uint64, C.c_int, C.c_int64, vp, C.c_uint64,
                                             u64p, vp]),
    "hd_planes_layout": (C.c_int, [C.c_uint64, C.c_uint64, C.c_int, C.c_uint32, u64p,
                                   u32p, u32p, u64p]),
    "hd_nibble_mask_dev": (C.c_int, [vp, C.c_uint64, C.c_uint64, u32p, vp]),
    "hd_encode_dev": (C.c_int, [vp, C.c_uint64, C.c_uint64, C.c_int, C.c_uint32, vp, vp]),
    "hd_all_pairs_dev": (C.c_int, [vp, C.c_uint64, C.c_uint64, C.c_int, C.c_uint32, C.c_int,
                                   C.c_uint32, C.c_int, C.c_int, vp, vp]),
    "hd_all_pairs_dev_rows": (C.c_int, [vp, C.c_uint64, C.c_uint64, C.c_int, C.c_uint32,
                                        C.c_int, C.c_uint32, C.c_int, C.c_int, C.c_uint64,
                                        C.c_uint64, vp, vp]),
    "hd_reference_distances_dev": (C.c_int, [vp, vp, C.c_uint64, C.c_uint64, vp, C.c_uint64,
                                             C.c_int, vp, vp]),
    "hd_batch_create": (C.c_int, [C.POINTER(vp)]),
    "hd_batch_destroy": (C.c_int, [vp]),
    "hd_batch_add": (C.c_int, [vp, vp, C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.c_uint32,
                               u64p]),
    "hd_batch_prepare": (C.c_int, [vp, C.c_int, C.c_int]),
    "hd_batch_run": (C.c_int, [vp, vp]),
    "hd_batch_run_sequential": (C.c_int, [vp, vp]),
    "hd_batch_output": (C.c_int, [vp, C.c_uint64, vp]),
    "hd_batch_output_dev": (C.c_int, [vp, C.c_uint64, C.POINTER(vp), u64p]),
    "hd_batch_info": (C.c_int, [vp, u64p, u64p, u32p, intp]),
    "hd_batch_variant_count": (C.c_int, []),
    "hd_batch_variant_info": (C.c_int, [C.c_int, intp, intp, intp]),
    "hd_server_create": (C.c_int, [vp, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int,
                                   C.c_uint64, C.POINTER(vp)]),
    "hd_server_destroy": (C.c_int, [vp]),
    "hd_server_submit_query": (C.c_int, [vp, vp, vp, C.c_uint64, C.c_int, u64p]),
    "hd_server_submit_all_pairs": (C.c_int, [vp, vp, C.c_uint64, C.c_int, C.c_int, C.c_uint32,
                                             u64p]),
    "hd_server_wait": (C.c_int, [vp, C.c_uint64, C.c_double, intp]),
    "hd_server_result": (C.c_int, [vp, C.c_uint64, vp, C.c_uint64, C.POINTER(C.c_double)]),
    "hd_server_release": (C.c_int, [vp, C.c_uint64]),
    "hd_server_stats": (C.c_int, [vp, u64p, u64p, u64p, u64p, u64p]),
    "hd_server_shared_stats": (C.c_int, [vp, u64p, u64p]),
    "hd_server_rect_stats": (C.c_int, [vp, intp, u64p, u64p]),
    "hd_variant_count": (C.c_int, []),
    "hd_variant_info": (C.c_int, [C.c_int, intp, intp, intp, intp]),
    "hd_default_variant": (C.c_int, [C.c_uint64, C.c_uint64, C.c_int, C.c_uint32, C.c_int,
                                     C.c_uint32, intp]),
}


class EngineError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[{status}] {message}")
        self.status = status


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY):
            raise ImportError(f"{LIBRARY} is not built; run hammingdist_amd/build.py")
        from ._runtime import ensure_single_runtime

        ensure_single_runtime()
        handle = C.CDLL(LIBRARY)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status: int) -> None:
    if status != HD_OK:
        raise EngineError(status, lib().hd_last_error().decode())


def ptr(a) -> int:
    """Address of a numpy array or a torch tensor's storage."""
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data


def device_count() -> int:
    return lib().hd_device_count()


def device_arch() -> str:
    buf = C.create_string_buffer(64)
    check(lib().hd_device_arch(buf, 64))
    return buf.value.decode()


def all_pairs(packed, n: int, bps: int, semantics: int, out_bytes: int, max_dist: int, out):
    check(lib().hd_all_pairs(ptr(packed), n, bps, semantics, out_bytes, max_dist, ptr(out)))


def all_pairs_stream(packed, n, bps, semantics, out_bytes, max_dist, band_cells, callback):
    """callback(cells_address, first_cell, n_cells, row_begin, row_end) -> None."""
    errors = []

    def cb(_user, cells, first, count, r0, r1):
        try:
            callback(cells, first, count, r0, r1)
            return 0
        except Exception as exc:  # reported after the call
            errors.append(exc)
            return 1

    fn = BAND_CB(cb)
    st = lib().hd_all_pairs_stream(ptr(packed), n, bps, semantics, out_bytes, max_dist,
                                   band_cells, fn, None)
    if errors:
        raise errors[0]
    check(st)


def distance_pair(a, b, nbytes: int, semantics: int, max_dist: int) -> int:
    out = C.c_int32(0)
    check(lib().hd_distance_pair(ptr(a), ptr(b), nbytes, semantics, max_dist, C.byref(out)))
    return out.value


def reference_distances(query, db, n_db: int, bps: int, rows, m: int, semantics: int, out):
    check(lib().hd_reference_distances(ptr(query), ptr(db), n_db, bps,
                                       None if rows is None else ptr(rows), m, semantics,
                                       ptr(out)))


def consensus_diff(packed, n: int, bps: int, L: int) -> tuple[int, int]:
    nd, mask = C.c_uint64(0), C.c_uint32(0)
    check(lib().hd_consensus_diff(ptr(packed), n, bps, L, C.byref(nd), C.byref(mask)))
    return nd.value, mask.value


def sparse_emit(cells, n: int, out_bytes: int, threshold: int, triplets, capacity: int) -> int:
    count = C.c_uint64(0)
    st = lib().hd_sparse_emit(ptr(cells), n, out_bytes, threshold,
                              None if triplets is None else ptr(triplets), capacity,
                              C.byref(count))
    if st not in (HD_OK, HD_ERR_CAPACITY) or (st == HD_ERR_CAPACITY and capacity):
        check(st)
    return count.value


def sparse_emit_dev(d_cells, n: int, out_bytes: int, threshold: int, d_triplets,
                    capacity: int, stream: int = 0) -> int:
    """Device triangle -> device (i, j, d) triplets; returns the total found
    (raises when it exceeds a non-zero capacity)."""
    count = C.c_uint64(0)
    st = lib().hd_sparse_emit_dev(ptr(d_cells), n, out_bytes, threshold,
                                  None if d_triplets is None else ptr(d_triplets), capacity,
                                  C.byref(count), stream)
    if st not in (HD_OK, HD_ERR_CAPACITY) or (st == HD_ERR_CAPACITY and capacity):
        check(st)
    return count.value


HD_EMIT_NOALLOC_MAX_CELLS = 1 << 33


def sparse_emit_dev_noalloc(d_cells, n: int, out_bytes: int, threshold: int, d_triplets,
                            capacity: int, stream: int = 0) -> int:
    """sparse_emit_dev through the entry point that allocates no device
    memory (its scratch is the library's static device array)."""
    count = C.c_uint64(0)
    st = lib().hd_sparse_emit_dev_noalloc(ptr(d_cells), n, out_bytes, threshold,
                                          None if d_triplets is None else ptr(d_triplets),
                                          capacity, C.byref(count), stream)
    if st not in (HD_OK, HD_ERR_CAPACITY) or (st == HD_ERR_CAPACITY and capacity):
        check(st)
    return count.value


def planes_layout(n: int, bps: int, semantics: int, mask: int):
    n_pad, W, P, total = C.c_uint64(), C.c_uint32(), C.c_uint32(), C.c_uint64()
    check(lib().hd_planes_layout(n, bps, semantics, mask, C.byref(n_pad), C.byref(W),
                                 C.byref(P), C.byref(total)))
    return n_pad.value, W.value, P.value, total.value


def default_variant(n: int, bps: int, semantics: int, mask: int, early_exit: bool,
                    max_dist: int) -> int:
    """The tile shape variant -1 selects for these inputs."""
    v = C.c_int()
    check(lib().hd_default_variant(n, bps, semantics, mask, 1 if early_exit else 0, max_dist,
                                   C.byref(v)))
    return v.value


def nibble_mask_dev(d_packed, n: int, bps: int, stream: int = 0) -> int:
    mask = C.c_uint32()
    check(lib().hd_nibble_mask_dev(ptr(d_packed), n, bps, C.byref(mask), stream))
    return mask.value


def encode_dev(d_packed, n: int, bps: int, semantics: int, mask: int, d_planes, stream: int = 0):
    check(lib().hd_encode_dev(ptr(d_packed), n, bps, semantics, mask, ptr(d_planes), stream))


def all_pairs_dev(d_planes, n, bps, semantics, mask, out_bytes, max_dist, early_exit, variant,
                  d_out, stream: int = 0):
    check(lib().hd_all_pairs_dev(ptr(d_planes), n, bps, semantics, mask, out_bytes, max_dist,
                                 1 if early_exit else 0, variant, ptr(d_out), stream))


def all_pairs_dev_rows(d_planes, n, bps, semantics, mask, out_bytes, max_dist, early_exit,
                       variant, row_begin, row_end, d_out, stream: int = 0):
    """Rows [row_begin, row_end) of the triangle into d_out (band-relative
    cells)."""
    check(lib().hd_all_pairs_dev_rows(ptr(d_planes), n, bps, semantics, mask, out_bytes,
                                      max_dist, int(early_exit), variant, row_begin, row_end,
                                      ptr(d_out), stream))


def reference_distances_dev(d_query, d_db, n_db, bps, d_rows, m, semantics, d_out,
                            stream: int = 0):
    check(lib().hd_reference_distances_dev(ptr(d_query), ptr(d_db), n_db, bps,
                                           None if d_rows is None else ptr(d_rows), m,
                                           semantics, ptr(d_out), stream))


def variants() -> list[dict]:
    out = []
    for v in range(lib().hd_variant_count()):
        r, c, w, p = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(lib().hd_variant_info(v, C.byref(r), C.byref(c), C.byref(w), C.byref(p)))
        out.append({"variant": v, "rows_per_lane": r.value, "cols_per_wave": c.value,
                    "waves": w.value})
    return out


class Batch:
    """Many independent all-pairs datasets run as one graph-captured
    persistent launch (hd_batch_* of the C ABI)."""

    def __init__(self):
        self._h = C.c_void_p()
        check(lib().hd_batch_create(C.byref(self._h)))
        self.jobs = []

    def close(self):
        if self._h:
            lib().hd_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, packed, semantics: int, out_bytes: int, max_dist: int) -> int:
        import numpy as np

        p = np.ascontiguousarray(packed)
        n, bps = p.shape
        job = C.c_uint64()
        check(lib().hd_batch_add(self._h, p.ctypes.data, n, bps, semantics, out_bytes,
                                 max_dist, C.byref(job)))
        self.jobs.append((n, out_bytes))
        return job.value

    def prepare(self, variant: int = -1, early_exit: bool = True):
        check(lib().hd_batch_prepare(self._h, variant, 1 if early_exit else 0))

    def run(self, stream: int = 0):
        check(lib().hd_batch_run(self._h, stream))

    def run_sequential(self, stream: int = 0):
        check(lib().hd_batch_run_sequential(self._h, stream))

    def output(self, job: int):
        import numpy as np

        n, ob = self.jobs[job]
        out = np.zeros(n * (n - 1) // 2, np.uint8 if ob == 1 else np.uint16)
        check(lib().hd_batch_output(self._h, job, out.ctypes.data))
        return out

    def info(self):
        jobs, tiles, blocks, variant = C.c_uint64(), C.c_uint64(), C.c_uint32(), C.c_int()
        check(lib().hd_batch_info(self._h, C.byref(jobs), C.byref(tiles), C.byref(blocks),
                                  C.byref(variant)))
        return {"jobs": jobs.value, "tiles": tiles.value, "blocks": blocks.value,
                "variant": variant.value}


def batch_variants() -> list[dict]:
    out = []
    for v in range(lib().hd_batch_variant_count()):
        r, c, w = C.c_int(), C.c_int(), C.c_int()
        check(lib().hd_batch_variant_info(v, C.byref(r), C.byref(c), C.byref(w)))
        out.append({"variant": v, "rows_per_lane": r.value, "cols_per_wave": c.value,
                    "waves": w.value})
    return out


class Server:
    """hd_server_*: a device-resident packed database serving reference
    distance queries and small all-pairs jobs with continuous batching."""

    def __init__(self, db, device_bytes: int = 8 << 30, host_bytes: int = 2 << 30,
                 streams: int = 2, max_batch: int = 64):
        import numpy as np

        d = np.ascontiguousarray(db)
        self.n_db, self.bps = d.shape
        self._h = C.c_void_p()
        check(lib().hd_server_create(d.ctypes.data, self.n_db, self.bps, device_bytes,
                                     host_bytes, streams, max_batch, C.byref(self._h)))
        self._meta = {}

    def close(self):
        if self._h:
            lib().hd_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit_query(self, query, rows, semantics: int = HD_SEM_DENSE) -> int:
        import numpy as np

        q = np.ascontiguousarray(query, dtype=np.uint8)
        r = np.ascontiguousarray(rows, dtype=np.uint64)
        rid = C.c_uint64()
        check(lib().hd_server_submit_query(self._h, q.ctypes.data, r.ctypes.data, r.size,
                                           semantics, C.byref(rid)))
        self._meta[rid.value] = ("query", r.size, 4)
        return rid.value

    def submit_all_pairs(self, rows, semantics: int = HD_SEM_DENSE, out_bytes: int = 1,
                         max_dist: int = 255) -> int:
        import numpy as np

        r = np.ascontiguousarray(rows, dtype=np.uint64)
        rid = C.c_uint64()
        check(lib().hd_server_submit_all_pairs(self._h, r.ctypes.data, r.size, semantics,
                                               out_bytes, max_dist, C.byref(rid)))
        self._meta[rid.value] = ("all_pairs", r.size, out_bytes)
        return rid.value

    def wait(self, rid: int, timeout: float = -1.0) -> int:
        st = C.c_int()
        check(lib().hd_server_wait(self._h, rid, timeout, C.byref(st)))
        return st.value

    def result(self, rid: int, release: bool = True):
        """(result array, latency seconds)."""
        import numpy as np

        kind, m, ob = self._meta[rid]
        if kind == "query":
            out = np.empty(m, np.uint32)
        else:
            out = np.empty(m * (m - 1) // 2, {1: np.uint8, 2: np.uint16, 4: np.uint32}[ob])
        lat = C.c_double()
        check(lib().hd_server_result(self._h, rid, out.ctypes.data if out.size else None,
                                     out.nbytes, C.byref(lat)))
        if release:
            check(lib().hd_server_release(self._h, rid))
            del self._meta[rid]
        return out, lat.value

    def stats(self) -> dict:
        v = [C.c_uint64() for _ in range(5)]
        check(lib().hd_server_stats(self._h, *[C.byref(x) for x in v]))
        out = dict(zip(["done", "batches", "overlapped_batches", "max_running",
                        "device_used"], [x.value for x in v]))
        g, q = C.c_uint64(), C.c_uint64()
        check(lib().hd_server_shared_stats(self._h, C.byref(g), C.byref(q)))
        out["shared_groups"], out["shared_queries"] = g.value, q.value
        en = C.c_int()
        check(lib().hd_server_rect_stats(self._h, C.byref(en), C.byref(g), C.byref(q)))
        out["rect_enabled"], out["rect_groups"], out["rect_queries"] = bool(en.value), g.value, \
            q.value
        return out
