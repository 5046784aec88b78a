"""GPU tests of the consensus-spec pair around the two entry points:
compute_cells_and_kzg_proofs (cells + proofs of a blob) and
recover_cells_and_kzg_proofs (cells + proofs from any half of the cells),
against the consensus vectors and the oracle."""
import hashlib
import random

import pytest

import vkzg_amd
from conftest import random_blob

pytestmark = pytest.mark.gpu

CELLS = 128 * 2048
PROOFS = 128 * 48


def test_cells_and_proofs_match_vectors(gpu_prover, vectors):
    names = vectors.valid_compute()
    blobs = b"".join(vectors.blob(n) for n in names)
    cells, proofs = gpu_prover.compute_cells_and_kzg_proofs_batch(blobs, len(names))
    for i, n in enumerate(names):
        assert hashlib.sha256(cells[CELLS * i:CELLS * (i + 1)]).hexdigest() == vectors.cells_sha256(n), n
        assert proofs[PROOFS * i:PROOFS * (i + 1)] == vectors.proofs(n), n


def test_cells_and_proofs_random_blobs_and_chunks(gpu_prover_small_batch, oracle_mod):
    rng = random.Random(41)
    blobs = [random_blob(rng) for _ in range(6)]  # chunks of 4 + 2
    cells, proofs = gpu_prover_small_batch.compute_cells_and_kzg_proofs_batch(b"".join(blobs), 6)
    for i, b in enumerate(blobs):
        rc, want = oracle_mod.blob_to_cells(b)
        assert rc == 0 and cells[CELLS * i:CELLS * (i + 1)] == want, i
    assert proofs == gpu_prover_small_batch.compute_cell_kzg_proofs_batch(b"".join(blobs), 6)


def test_cells_and_proofs_invalid_blob(gpu_prover, vectors):
    bad = vectors.invalid_compute_131072()[0]
    with pytest.raises(vkzg_amd.VkzgError) as e:
        gpu_prover.compute_cells_and_kzg_proofs_batch(vectors.blob(bad), 1)
    assert e.value.code == vkzg_amd.ERR_INVALID_BLOB


def test_recover_cells_and_proofs_vectors(gpu_prover, vectors, cells_of):
    for case in vectors.recover_cases():
        src_cells = cells_of(case["source"]) if case["source"] else bytes(CELLS)
        present = bytearray(128)
        for c in case["cell_indices"]:
            present[c] = 1
        if not case["valid"]:
            with pytest.raises(vkzg_amd.VkzgError) as e:
                gpu_prover.recover_cells_and_kzg_proofs_batch(src_cells, bytes(present), 1)
            assert e.value.code == vkzg_amd.ERR_BADARGS
            continue
        out, proofs = gpu_prover.recover_cells_and_kzg_proofs_batch(src_cells, bytes(present), 1)
        assert hashlib.sha256(out).hexdigest() == case["cells_sha256"], case["name"]
        assert case["proofs_equal_source"]
        assert proofs == vectors.proofs(case["source"]), case["name"]


def test_recover_cells_and_proofs_random_patterns(gpu_seeded_prover, oracle_mod, seeded_verify_setup):
    """Seeded setup: 5 blobs with different missing sets in one call; cells
    equal the originals and every proof passes the pairing check."""
    rng = random.Random(43)
    blobs = [random_blob(rng) for _ in range(5)]
    full = b"".join(oracle_mod.blob_to_cells(b)[1] for b in blobs)
    present = bytearray(128 * 5)
    cells = bytearray(full)
    for b in range(5):
        keep = set(rng.sample(range(128), 64 + 15 * b))
        for c in range(128):
            if c in keep:
                present[128 * b + c] = 1
            else:
                cells[CELLS * b + 2048 * c:CELLS * b + 2048 * (c + 1)] = rng.randbytes(2048)
    out, proofs = gpu_seeded_prover.recover_cells_and_kzg_proofs_batch(bytes(cells), bytes(present), 5)
    assert out == full
    assert proofs == gpu_seeded_prover.compute_cell_kzg_proofs_batch(b"".join(blobs), 5)
    assert seeded_verify_setup.verify_blobs(b"".join(blobs), proofs, 5, cells=out, seed=7)
