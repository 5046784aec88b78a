"""FK20 cell proofs on the GPU through the C ABI, bit-exact against the
consensus vectors, the oracle and the known-tau closed form (reference
tests/test.cpp:55-202)."""
import ctypes
import random
import threading

import pytest

import vkzg_amd
from conftest import SEEDED_TAU, random_blob

pytestmark = pytest.mark.gpu

PROOFS = 6144


def test_vectors_bit_exact(gpu_prover, vectors):
    for name in vectors.valid_compute():
        assert gpu_prover.compute_cell_kzg_proofs_batch(vectors.blob(name), 1) == \
            vectors.proofs(name), name


def test_invalid_blobs_rejected(gpu_prover, vectors):
    for name in vectors.invalid_compute_131072():
        with pytest.raises(vkzg_amd.VkzgError) as e:
            gpu_prover.compute_cell_kzg_proofs_batch(vectors.blob(name), 1)
        assert e.value.code == vkzg_amd.ERR_INVALID_BLOB


def test_batched_equals_single(gpu_prover, vectors):
    names = vectors.valid_compute()
    out = gpu_prover.compute_cell_kzg_proofs_batch(b"".join(vectors.blob(n) for n in names),
                                                   len(names))
    assert out == b"".join(vectors.proofs(n) for n in names)


def test_api_contract(gpu_prover, vectors):
    L = vkzg_amd._lib.load()
    blob = vectors.blob("compute_cells_and_kzg_proofs_case_valid_3")
    buf = ctypes.create_string_buffer(PROOFS)
    assert L.vkzg_compute_proofs_batch(gpu_prover.handle, buf, None, 1) == vkzg_amd.ERR_BADARGS
    assert L.vkzg_compute_proofs_batch(gpu_prover.handle, None, blob, 1) == vkzg_amd.ERR_BADARGS
    assert L.vkzg_compute_proofs_batch(gpu_prover.handle, buf, blob, 0) == vkzg_amd.OK
    assert gpu_prover.gpu_core_count() > 0
    assert gpu_prover.device_name()


def test_chunking_with_small_batch(gpu_prover_small_batch, vectors):
    blob = vectors.blob("compute_cells_and_kzg_proofs_case_valid_4")
    want = vectors.proofs("compute_cells_and_kzg_proofs_case_valid_4")
    n = 4 * 2 + 1
    assert gpu_prover_small_batch.compute_cell_kzg_proofs_batch(blob * n, n) == want * n


def test_invalid_blob_in_second_chunk(gpu_prover_small_batch, vectors):
    good = vectors.blob("compute_cells_and_kzg_proofs_case_valid_2")
    bad = vectors.blob(vectors.invalid_compute_131072()[1])
    with pytest.raises(vkzg_amd.VkzgError) as e:
        gpu_prover_small_batch.compute_cell_kzg_proofs_batch(good * 5 + bad + good, 7)
    assert e.value.code == vkzg_amd.ERR_INVALID_BLOB


def test_concurrent_calls(gpu_prover, vectors):
    names = vectors.valid_compute()
    bad = []

    def worker(t):
        for r in range(3):
            n = names[(t + r) % len(names)]
            if gpu_prover.compute_cell_kzg_proofs_batch(vectors.blob(n), 1) != vectors.proofs(n):
                bad.append(n)

    threads = [threading.Thread(target=worker, args=(t,)) for t in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not bad


def test_seeded_setup_closed_form(gpu_seeded_prover, oracle_mod, seeded_verify_setup):
    """A seeded-tau setup: every GPU proof equals [q_c(tau)]G1 and passes
    the pairing check of batch cell-proof verification."""
    rng = random.Random(31)
    blobs = [random_blob(rng) for _ in range(3)]
    out = gpu_seeded_prover.compute_cell_kzg_proofs_batch(b"".join(blobs), 3)
    for i, blob in enumerate(blobs):
        rc, want = oracle_mod.proofs_from_tau(blob, SEEDED_TAU)
        assert rc == 0
        assert out[PROOFS * i:PROOFS * (i + 1)] == want
    assert seeded_verify_setup.verify_blobs(b"".join(blobs), out, 3)


def test_seeded_setup_256_blobs(gpu_seeded_prover, oracle_mod, seeded_verify_setup):
    """The 256-blob batch: spot-check blobs against the closed form, the
    whole batch for linearity (proofs(a)+proofs(b) = proofs(a+b)), and every
    one of the 32768 proofs by the pairing check."""
    R = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
    rng = random.Random(32)
    base = [random_blob(rng) for _ in range(3)]
    blobs = [base[i % 3] for i in range(256)]
    a = [int.from_bytes(base[0][32 * k:32 * k + 32], "big") for k in range(4096)]
    b = [int.from_bytes(base[1][32 * k:32 * k + 32], "big") for k in range(4096)]
    blobs[255] = b"".join(((x + y) % R).to_bytes(32, "big") for x, y in zip(a, b))
    out = gpu_seeded_prover.compute_cell_kzg_proofs_batch(b"".join(blobs), 256)
    for i in (0, 1, 2, 130):
        rc, want = oracle_mod.proofs_from_tau(blobs[i], SEEDED_TAU)
        assert out[PROOFS * i:PROOFS * (i + 1)] == want, i
    for c in range(0, 128, 9):
        s0 = out[PROOFS * 0 + 48 * c:PROOFS * 0 + 48 * c + 48]
        s1 = out[PROOFS * 1 + 48 * c:PROOFS * 1 + 48 * c + 48]
        s255 = out[PROOFS * 255 + 48 * c:PROOFS * 255 + 48 * c + 48]
        assert oracle_mod.g1_add(s0, s1) == s255
    assert seeded_verify_setup.verify_blobs(b"".join(blobs), out, 256, seed=256)


def test_device_entry_point_matches_host(gpu_prover, vectors):
    from vkzg_amd import hip
    names = vectors.valid_compute()
    d_blobs = hip.DeviceBuffer.from_bytes(b"".join(vectors.blob(n) for n in names))
    d_proofs = hip.DeviceBuffer(PROOFS * len(names))
    stream = hip.Stream()
    gpu_prover.compute_proofs_device(d_proofs.ptr, d_blobs.ptr, len(names), stream.handle)
    assert d_proofs.to_bytes() == b"".join(vectors.proofs(n) for n in names)
    # invalid blob through the device entry
    bad = hip.DeviceBuffer.from_bytes(vectors.blob(vectors.invalid_compute_131072()[0]))
    with pytest.raises(vkzg_amd.VkzgError) as e:
        gpu_prover.compute_proofs_device(d_proofs.ptr, bad.ptr, 1, None)
    assert e.value.code == vkzg_amd.ERR_INVALID_BLOB
    stream.destroy()


def test_prover_info(gpu_prover, gpu_prover_small_batch):
    info = gpu_prover.info()
    print("prover info:", info, gpu_prover.device_name())
    # window plan (c, W, k): W windows of c bits, the top k one bit wider,
    # covering the 255 bits of a signed scalar's magnitude and sign
    c, W = info["window_bits"], info["windows"]
    entries = info["table_bytes"] // (8192 * 128)
    k = entries // (1 << (c - 1)) - W
    assert entries == (W + k) << (c - 1) and 0 <= k <= W
    assert W * c + k >= 255 and (W - 1) * c + max(k - 1, 0) < 255
    assert info["max_batch_size"] == 256
    assert gpu_prover_small_batch.info()["max_batch_size"] == 4
