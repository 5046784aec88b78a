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
    """Small batches and 3 slots: each result of a concurrent mix equals the
    same request made alone through the synchronous C ABI."""
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


def test_serving_mix_at_full_size(gpu_prover, vectors, cells_of):
    """The serving mix at its stated sizes on the default server (256-blob
    batches, 2 slots): 70% proof requests of k ~ U[1,48] blobs drawn from the
    7 consensus vectors, 30% recovery requests of k ~ U[1,64] blobs with
    64-127 of 128 cells present and garbage in the others.  Every proof is
    checked against the consensus proofs and every recovered blob against its
    original cells.  Meanwhile another thread makes synchronous proof and
    recovery calls on the same prover and a third changes the prover's
    tunings; neither may change any result."""
    rng = random.Random(4242)
    names = vectors.valid_compute()
    reqs = []
    for i in range(48):
        if rng.random() < 0.7:
            k = rng.randint(1, 48)
            pick = [names[rng.randrange(len(names))] for _ in range(k)]
            reqs.append(("p", k, b"".join(vectors.blob(n) for n in pick),
                         None, b"".join(vectors.proofs(n) for n in pick)))
        else:
            k = rng.randint(1, 64)
            garbage = rng.randbytes(CELLS)
            cells, masks, want = [], [], []
            for _ in range(k):
                src = cells_of(names[rng.randrange(len(names))])
                keep = set(rng.sample(range(128), rng.randint(64, 127)))
                c = bytearray(garbage)
                for q in keep:
                    c[2048 * q:2048 * (q + 1)] = src[2048 * q:2048 * (q + 1)]
                cells.append(bytes(c))
                masks.append(bytes(1 if q in keep else 0 for q in range(128)))
                want.append(src)
            reqs.append(("r", k, b"".join(cells), b"".join(masks), b"".join(want)))
    assert max(r[1] for r in reqs if r[0] == "p") > 32
    assert max(r[1] for r in reqs if r[0] == "r") > 32

    stop = threading.Event()
    side_errors = []

    def sync_caller():
        local = random.Random(7)
        while not stop.is_set():
            n = names[local.randrange(len(names))]
            k = local.randint(1, 8)
            if gpu_prover.compute_cell_kzg_proofs_batch(vectors.blob(n) * k, k) != vectors.proofs(n) * k:
                side_errors.append(("sync proofs", n, k))
            src = cells_of(n)
            keep = set(local.sample(range(128), 64))
            c = bytearray(CELLS)
            for q in keep:
                c[2048 * q:2048 * (q + 1)] = src[2048 * q:2048 * (q + 1)]
            pm = bytes(1 if q in keep else 0 for q in range(128))
            if gpu_prover.recover_cells_batch(bytes(c), pm, 1) != src:
                side_errors.append(("sync recovery", n))

    saved = {nb: gpu_prover.get_tuning(nb) for nb in (1, 4, 16, 64, 256)}

    def tuner():
        local = random.Random(8)
        while not stop.is_set():
            nb = local.choice([1, 4, 16, 64, 256])
            gpu_prover.set_tuning(nb, phase_a_groups=local.choice([1, 2, 4]),
                                  reduce_lanes=local.choice([8, 16]),
                                  recovery_waves=local.choice([1, 2, 4]),
                                  recovery_radix=local.choice([4, 8]))
            stop.wait(0.005)

    side = [threading.Thread(target=sync_caller), threading.Thread(target=tuner)]
    try:
        with vkzg_amd.Server(gpu_prover) as srv:
            st = srv.stats()
            assert st["batch_blobs"] == gpu_prover.info()["max_batch_size"] and st["slots"] == 2
            for th in side:
                th.start()
            results = [None] * len(reqs)

            def client(idx):
                tickets = []
                for i in idx:
                    kind, k, data, mask, _ = reqs[i]
                    t = srv.submit_proofs(data, k) if kind == "p" else srv.submit_recovery(data, mask, k)
                    tickets.append((i, t))
                for i, t in tickets:
                    results[i], _meta = srv.wait(t)

            clients = [threading.Thread(target=client, args=(list(range(j, len(reqs), 4)),))
                       for j in range(4)]
            for th in clients:
                th.start()
            for th in clients:
                th.join()
    finally:
        stop.set()
        for th in side:
            if th.is_alive() or th.ident:
                th.join()
        for nb, t in saved.items():
            gpu_prover.set_tuning(nb, **t)
    assert not side_errors, side_errors[:3]
    for i, (kind, k, _data, _mask, want) in enumerate(reqs):
        assert results[i] == want, (i, kind, k)


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
    # a budget of the resident tables plus 10.5 blobs of working set: one slot
    # of 10 blobs; the tables alone do not leave room for any slot
    tables = gpu_prover.info()["resident_table_bytes"]
    with pytest.raises(vkzg_amd.VkzgError) as e:
        vkzg_amd.Server(gpu_prover, hbm_budget_bytes=tables)
    assert e.value.code == vkzg_amd.ERR_MALLOC
    with vkzg_amd.Server(gpu_prover, batch_blobs=64, slots=4,
                         hbm_budget_bytes=tables + per_blob * 21 // 2) as srv:
        st = srv.stats()
        assert st["slots"] == 1 and st["batch_blobs"] == 10
        t = srv.submit_proofs(blob * 23, 23)
        out, meta = srv.wait(t)
        assert out == want * 23
        assert meta["batches"] >= 3
        t0 = srv.submit_proofs(b"", 0)
        assert srv.wait(t0)[0] == b""
