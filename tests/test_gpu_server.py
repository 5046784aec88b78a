"""The request server (continuous batching) on the GPU: every result in a
concurrent mix equals the same request made alone through the C ABI."""
import random
import threading

import pytest

import vkzg_amd
from conftest import random_blob

pytestmark = pytest.mark.gpu

CELLS = 128 * 2048


def test_mix_equals_alone(gpu_prover, oracle_mod):
    rng = random.Random(41)
    pool = [random_blob(rng) for _ in range(6)]
    cells_pool = []
    for b in pool[:3]:
        rc, c = oracle_mod.blob_to_cells(b)
        assert rc == 0
        cells_pool.append(c)
    reqs = []
    for i in range(40):
        if rng.random() < 0.7:
            k = rng.randint(1, 20)
            blobs = b"".join(pool[rng.randrange(6)] for _ in range(k))
            reqs.append(("p", k, blobs, None))
        else:
            k = rng.randint(1, 12)
            cells, masks = [], []
            for _ in range(k):
                src = cells_pool[rng.randrange(3)]
                keep = set(rng.sample(range(128), rng.randint(64, 127)))
                c = bytearray(rng.randbytes(CELLS))
                for q in keep:
                    c[2048 * q:2048 * (q + 1)] = src[2048 * q:2048 * (q + 1)]
                cells.append(bytes(c))
                masks.append(bytes(1 if q in keep else 0 for q in range(128)))
            reqs.append(("r", k, b"".join(cells), b"".join(masks)))
    # alone: one synchronous call per request
    alone = []
    for kind, k, data, mask in reqs:
        if kind == "p":
            alone.append(gpu_prover.compute_cell_kzg_proofs_batch(data, k))
        else:
            alone.append(gpu_prover.recover_cells_batch(data, mask, k))
    # in the mix: all submitted from 4 threads at once to a server with small batches
    with vkzg_amd.Server(gpu_prover, batch_blobs=32, slots=3) as srv:
        results = [None] * len(reqs)
        batches = [0] * len(reqs)

        def client(idx):
            tickets = []
            for i in idx:
                kind, k, data, mask = reqs[i]
                t = srv.submit_proofs(data, k) if kind == "p" else srv.submit_recovery(data, mask, k)
                tickets.append((i, t))
            for i, t in tickets:
                results[i], meta = srv.wait(t)
                batches[i] = meta["batches"]

        threads = [threading.Thread(target=client, args=(list(range(j, len(reqs), 4)),))
                   for j in range(4)]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        st = srv.stats()
    assert st["requests"] == len(reqs)
    assert st["batches"] < len(reqs)  # requests were packed together
    for i, (a, b) in enumerate(zip(alone, results)):
        assert a == b, (i, reqs[i][0], reqs[i][1])


def test_invalid_blob_fails_only_its_request(gpu_prover, vectors):
    good = vectors.blob("compute_cells_and_kzg_proofs_case_valid_2")
    want = vectors.proofs("compute_cells_and_kzg_proofs_case_valid_2")
    bad = vectors.blob(vectors.invalid_compute_131072()[1])
    with vkzg_amd.Server(gpu_prover, batch_blobs=16, slots=1) as srv:
        t1 = srv.submit_proofs(good * 3, 3)
        t2 = srv.submit_proofs(good + bad, 2)
        t3 = srv.submit_proofs(good * 2, 2)
        assert srv.wait(t1)[0] == want * 3
        with pytest.raises(vkzg_amd.VkzgError) as e:
            srv.wait(t2)
        assert e.value.code == vkzg_amd.ERR_INVALID_BLOB
        assert srv.wait(t3)[0] == want * 2
        # more than 64 missing cells is rejected at submission
        with pytest.raises(vkzg_amd.VkzgError) as e:
            srv.submit_recovery(bytes(CELLS), bytes([1] * 63 + [0] * 65), 1)
        assert e.value.code == vkzg_amd.ERR_BADARGS


def test_large_request_spans_batches_and_budget(gpu_prover, vectors):
    blob = vectors.blob("compute_cells_and_kzg_proofs_case_valid_3")
    want = vectors.proofs("compute_cells_and_kzg_proofs_case_valid_3")
    with vkzg_amd.Server(gpu_prover, batch_blobs=1, slots=1) as one:
        per_blob = one.stats()["slot_bytes"]
    # a budget of 10.5 blobs of working set: one slot of 10 blobs
    with vkzg_amd.Server(gpu_prover, batch_blobs=64, slots=4,
                         hbm_budget_bytes=per_blob * 21 // 2) as srv:
        st = srv.stats()
        assert st["slots"] == 1 and st["batch_blobs"] == 10
        t = srv.submit_proofs(blob * 23, 23)
        out, meta = srv.wait(t)
        assert out == want * 23
        assert meta["batches"] >= 3
        t0 = srv.submit_proofs(b"", 0)
        assert srv.wait(t0)[0] == b""
