"""Device-resident entry points (torch device memory, torch's stream): every
tile variant, early exit on and off, and every output width give the same
triangle as the oracle."""
import numpy as np
import pytest

import oracle_lib as O
from conftest import random_strings
from hammingdist_amd import capi

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

DT = {1: np.uint8, 2: np.uint16, 4: np.uint32}


def run_dev(packed, sem, out_bytes, max_dist, early, variant):
    n, bps = packed.shape
    stream = torch.cuda.current_stream().cuda_stream
    d_packed = torch.from_numpy(np.ascontiguousarray(packed)).cuda()
    mask = capi.nibble_mask_dev(d_packed, n, bps, stream)
    _, _, _, total = capi.planes_layout(n, bps, sem, mask)
    d_planes = torch.empty(total, dtype=torch.int32, device="cuda")
    capi.encode_dev(d_packed, n, bps, sem, mask, d_planes, stream)
    d_out = torch.zeros(n * (n - 1) // 2 * out_bytes, dtype=torch.uint8, device="cuda")
    capi.all_pairs_dev(d_planes, n, bps, sem, mask, out_bytes, max_dist, early, variant, d_out,
                       stream)
    torch.cuda.synchronize()
    return d_out.cpu().numpy().view(DT[out_bytes])


@pytest.mark.parametrize("variant", range(5))
@pytest.mark.parametrize("alphabet,include_x", [("ACGT-N", False), ("ACGTX-N", True)])
def test_every_variant_matches_oracle(variant, alphabet, include_x, rng):
    # ACGT-N: the dense 3-op code; X packed as code 3: the 4-plane code.
    for n, L in [(2, 3), (65, 64), (600, 300), (1100, 97)]:
        seqs = random_strings(rng, n, L, alphabet)
        packed = O.pack(O.as_chars(seqs), include_x)
        for out_bytes, md in [(1, 255), (2, 7), (4, 2**32 - 1)]:
            want = O.all_pairs_dense(packed, out_bytes, md)
            for early in (False, True):
                got = run_dev(packed, capi.HD_SEM_DENSE, out_bytes, md, early, variant)
                np.testing.assert_array_equal(got, want, err_msg=f"v={variant} n={n} L={L}")


@pytest.mark.parametrize("variant", range(5))
def test_every_variant_sparse_x(variant, rng):
    n, L = 700, 150
    seqs = random_strings(rng, n, L, "ACGTX-N")
    chars = O.as_chars(seqs)
    packed = O.pack(chars, include_x=True)
    want = O.all_pairs_sparse(chars, True, 2, 65535)
    got = run_dev(packed, capi.HD_SEM_SPARSE_X, 2, 65535, True, variant)
    np.testing.assert_array_equal(got, want)


def test_early_exit_identical_on_saturating_data():
    # Every pair far above max_dist: the early exit fires in the first chunk.
    rng = np.random.default_rng(7)
    n, L = 900, 5000
    seqs = random_strings(rng, n, L, "ACGT")
    packed = O.pack(O.as_chars(seqs))
    want = O.all_pairs_dense(packed, 1, 2)
    assert (want == 2).all()
    for v in range(5):
        got = run_dev(packed, capi.HD_SEM_DENSE, 1, 2, True, v)
        np.testing.assert_array_equal(got, want)


def test_reference_distances_dev_matches_host(rng):
    m, L = 1000, 3000
    seqs = random_strings(rng, m + 1, L, "ACGT-N")
    db = O.pack(O.as_chars(seqs[1:]))
    q = O.pack(O.as_chars(seqs[:1]))[0]
    d_db = torch.from_numpy(db).cuda()
    d_q = torch.from_numpy(q).cuda()
    d_out = torch.zeros(m, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    capi.reference_distances_dev(d_q, d_db, m, db.shape[1], None, m, capi.HD_SEM_DENSE, d_out,
                                 stream)
    torch.cuda.synchronize()
    want = [O.distance_char(seqs[0], s) for s in seqs[1:]]
    assert d_out.cpu().numpy().astype(np.uint32).tolist() == want


@pytest.mark.parametrize("out_bytes", [1, 2, 4])
def test_sparse_emit_dev_on_resident_triangle(out_bytes, rng):
    # all-pairs into HBM, then the (i, j, d) emit of cells <= threshold from
    # the resident triangle, row-major (include/hamming/hamming.hh:79-114)
    n, L = 1300, 400
    seqs = random_strings(rng, n, L, "ACGT-N")
    base = seqs[0]
    seqs = [base if k % 5 == 0 else s for k, s in enumerate(seqs)]  # some exact duplicates
    packed = O.pack(O.as_chars(seqs))
    md = {1: 255, 2: 65535, 4: 2**32 - 1}[out_bytes]
    want_cells = O.all_pairs_dense(packed, out_bytes, md)
    n_, bps = packed.shape
    stream = torch.cuda.current_stream().cuda_stream
    d_packed = torch.from_numpy(np.ascontiguousarray(packed)).cuda()
    mask = capi.nibble_mask_dev(d_packed, n, bps, stream)
    _, _, _, total = capi.planes_layout(n, bps, capi.HD_SEM_DENSE, mask)
    d_planes = torch.empty(total, dtype=torch.int32, device="cuda")
    capi.encode_dev(d_packed, n, bps, capi.HD_SEM_DENSE, mask, d_planes, stream)
    d_cells = torch.zeros(want_cells.size * out_bytes, dtype=torch.uint8, device="cuda")
    capi.all_pairs_dev(d_planes, n, bps, capi.HD_SEM_DENSE, mask, out_bytes, md, True, -1,
                       d_cells, stream)
    I, J = O.tri_index(n)
    for thr in [0, 3, 150]:
        sel = want_cells <= thr
        count = capi.sparse_emit_dev(d_cells, n, out_bytes, thr, None, 0, stream)
        assert count == int(sel.sum())
        d_trip = torch.zeros(max(count, 1) * 3, dtype=torch.int32, device="cuda")
        assert capi.sparse_emit_dev(d_cells, n, out_bytes, thr, d_trip, count, stream) == count
        torch.cuda.synchronize()
        trip = d_trip[:count * 3].cpu().numpy().view(np.uint32).reshape(-1, 3)
        np.testing.assert_array_equal(trip[:, 0], I[sel])
        np.testing.assert_array_equal(trip[:, 1], J[sel])
        np.testing.assert_array_equal(trip[:, 2], want_cells[sel])
        if count > 1:
            with pytest.raises(capi.EngineError, match="capacity"):
                capi.sparse_emit_dev(d_cells, n, out_bytes, thr, d_trip, count - 1, stream)
