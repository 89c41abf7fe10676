"""Batched all-pairs (one persistent, graph-captured launch over many uneven
datasets): every job's triangle equals the oracle's, for every batch tile
shape, on repeated graph replays with eager work in between, and equals the
sequential plain launches."""
import numpy as np
import pytest

import oracle_lib as O
from conftest import random_strings
from hammingdist_amd import capi

pytestmark = pytest.mark.gpu


def make_jobs(rng, count):
    jobs = []
    for k in range(count):
        n = int(rng.choice([1, 2, 3, 17, 64, 65, 130, 257, 600, 1100]))
        L = int(rng.choice([1, 15, 33, 100, 511, 1500, 3001]))
        alphabet = ["ACGT-", "ACGT-N", "AAAAAAAAAAAAAAACGTN-", "ACGTX-N"][k % 4]
        seqs = random_strings(rng, n, L, alphabet)
        chars = O.as_chars(seqs)
        kind = k % 3
        if kind == 0:  # dense (reference GPU route)
            packed, sem = O.pack(chars), capi.HD_SEM_DENSE
        elif kind == 1:  # sparse without X
            packed, sem = O.pack(chars), capi.HD_SEM_SPARSE
        else:  # sparse with X
            packed, sem = O.pack(chars, True), capi.HD_SEM_SPARSE_X
        out_bytes = 1 if k % 2 == 0 else 2
        md = [255, 65535, 3, 40][k % 4] if out_bytes == 2 else [255, 3, 40, 0][k % 4]
        if n >= 2:
            if sem == capi.HD_SEM_DENSE:
                want = O.all_pairs_dense(packed, out_bytes, md)
            else:
                want = O.all_pairs_sparse(chars, sem == capi.HD_SEM_SPARSE_X, out_bytes, md)
        else:
            want = np.zeros(0, np.uint8 if out_bytes == 1 else np.uint16)
        jobs.append((packed, sem, out_bytes, md, want))
    return jobs


@pytest.fixture(scope="module")
def jobs():
    return make_jobs(np.random.default_rng(31), 40)


@pytest.mark.parametrize("variant", range(5))
def test_batch_matches_oracle_on_replays(jobs, variant):
    import torch

    b = capi.Batch()
    ids = [b.add(p, sem, ob, md) for p, sem, ob, md, _ in jobs]
    b.prepare(variant, early_exit=True)
    stream = torch.cuda.current_stream().cuda_stream
    for replay in range(3):
        b.run(stream)
        torch.cuda.synchronize()
        for k, (_, _, _, _, want) in zip(ids, jobs):
            np.testing.assert_array_equal(b.output(k), want, err_msg=f"v={variant} job={k}")
        # eager work between replays
        x = torch.ones(1 << 20, device="cuda")
        assert float(x.sum()) == float(1 << 20)
    b.close()


def test_sequential_equals_batched(jobs):
    import torch

    b = capi.Batch()
    ids = [b.add(p, sem, ob, md) for p, sem, ob, md, _ in jobs]
    b.prepare(-1, early_exit=False)
    stream = torch.cuda.current_stream().cuda_stream
    b.run_sequential(stream)
    torch.cuda.synchronize()
    seq = [b.output(k) for k in ids]
    b.run(stream)
    torch.cuda.synchronize()
    for k, (_, _, _, _, want) in zip(ids, jobs):
        np.testing.assert_array_equal(seq[k], want)
        np.testing.assert_array_equal(b.output(k), want)
    info = b.info()
    assert info["jobs"] == len(jobs) and info["tiles"] > 0 and info["blocks"] > 0
    b.close()


def test_batch_of_generated_workloads():
    # a few datasets of the batched workload's generator (plain generator,
    # uneven n and L, alternating uint8 max 255 / uint16 max 65535)
    from hammingdist_amd import workloads

    specs = workloads.batched(count=12, n_max=700, L_max=6000)
    b = capi.Batch()
    wants = []
    for spec, ob in specs:
        packed = workloads.generate_packed(spec)
        md = 255 if ob == 1 else 65535
        b.add(packed, capi.HD_SEM_DENSE, ob, md)
        wants.append(O.all_pairs_dense(packed, ob, md))
    b.prepare(-1)
    b.run(0)
    import torch

    torch.cuda.synchronize()
    for k, want in enumerate(wants):
        np.testing.assert_array_equal(b.output(k), want)
    b.close()
