"""hd_sparse_emit_dev_noalloc (the sparse emit whose scratch is the library's
static device array) against hd_sparse_emit_dev (scratch allocated per call):
the same count and the same triplets, on two calls in a row, the second with
a smaller triangle (the array keeps the first call's contents), for every
cell width, under a capacity below the total, and a triangle past the stated
maximum rejected before any device work."""
import numpy as np
import pytest

import oracle_lib as O
from conftest import random_strings
from hammingdist_amd import capi

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def device_triangle(rng, n, L, out_bytes, max_dist):
    seqs = random_strings(rng, n, L, "ACGT-N")
    packed = O.pack(O.as_chars(seqs))
    n, bps = packed.shape
    sh = torch.cuda.current_stream().cuda_stream
    d_packed = torch.from_numpy(packed.reshape(-1)).cuda()
    mask = capi.nibble_mask_dev(d_packed, n, bps, sh)
    _, _, _, total = capi.planes_layout(n, bps, capi.HD_SEM_DENSE, mask)
    d_planes = torch.empty(total, dtype=torch.int32, device="cuda")
    capi.encode_dev(d_packed, n, bps, capi.HD_SEM_DENSE, mask, d_planes, sh)
    d_out = torch.empty(n * (n - 1) // 2 * out_bytes, dtype=torch.uint8, device="cuda")
    capi.all_pairs_dev(d_planes, n, bps, capi.HD_SEM_DENSE, mask, out_bytes, max_dist, False, -1,
                       d_out, sh)
    torch.cuda.synchronize()
    want = O.all_pairs_dense(packed, out_bytes, max_dist)
    assert np.array_equal(d_out.cpu().numpy().view(want.dtype), want)
    return d_out


def both(d_cells, n, out_bytes, thr, capacity):
    sh = torch.cuda.current_stream().cuda_stream
    cap = max(capacity, 1)
    a = torch.full((cap * 3,), -1, dtype=torch.int32, device="cuda")
    b = torch.full((cap * 3,), -1, dtype=torch.int32, device="cuda")
    ca = capi.sparse_emit_dev(d_cells, n, out_bytes, thr, a, capacity, sh)
    cb = capi.sparse_emit_dev_noalloc(d_cells, n, out_bytes, thr, b, capacity, sh)
    torch.cuda.synchronize()
    return ca, cb, a, b


@pytest.mark.parametrize("out_bytes,max_dist", [(1, 255), (2, 700), (4, 2**32 - 1)])
def test_noalloc_emit_equals_allocating_emit(out_bytes, max_dist):
    rng = np.random.default_rng(77 + out_bytes)
    sh = torch.cuda.current_stream().cuda_stream
    # first call: a larger triangle; second: a smaller one with another
    # threshold, over the scratch the first call left behind
    for n, L, thr in [(2300, 300, 160), (700, 90, 50)]:
        print(f"noalloc emit: out_bytes={out_bytes} n={n} L={L} thr={thr}", flush=True)
        d_cells = device_triangle(rng, n, L, out_bytes, max_dist)
        total = capi.sparse_emit_dev(d_cells, n, out_bytes, thr, None, 0, sh)
        assert capi.sparse_emit_dev_noalloc(d_cells, n, out_bytes, thr, None, 0, sh) == total
        assert 0 < total < n * (n - 1) // 2
        ca, cb, a, b = both(d_cells, n, out_bytes, thr, total)
        assert ca == cb == total
        assert torch.equal(a, b)
        # capacity below the total: HD_ERR_CAPACITY, the first triplets written
        cap = total // 3
        sh = torch.cuda.current_stream().cuda_stream
        for fn in (capi.sparse_emit_dev, capi.sparse_emit_dev_noalloc):
            d_cap = torch.full((cap * 3 + 3,), -1, dtype=torch.int32, device="cuda")
            with pytest.raises(capi.EngineError, match="capacity"):
                fn(d_cells, n, out_bytes, thr, d_cap, cap, sh)
            torch.cuda.synchronize()
            assert torch.equal(d_cap[: 3 * cap], a[: 3 * cap])
            assert int((d_cap[3 * cap:] == -1).sum()) == 3


def test_noalloc_emit_rejects_past_maximum():
    # the size check runs on the host before any device work: a tiny valid
    # buffer stands in for the (absent) triangle of 131,073 rows
    n = 131_073
    assert n * (n - 1) // 2 > capi.HD_EMIT_NOALLOC_MAX_CELLS
    tiny = torch.zeros(16, dtype=torch.uint8, device="cuda")
    sh = torch.cuda.current_stream().cuda_stream
    with pytest.raises(capi.EngineError, match="maximum"):
        capi.sparse_emit_dev_noalloc(tiny, n, 1, 3, None, 0, sh)
    # the largest triangle within the maximum passes the check (n = 131,072)
    assert 131_072 * 131_071 // 2 <= capi.HD_EMIT_NOALLOC_MAX_CELLS
