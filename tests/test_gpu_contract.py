"""The reference's call contract on the GPU, through raw ctypes calls:
per-chunk partial output on error (src/vkzg.cpp:1034,1191-1197,1209-1215),
the 64-blob / exactly-64-present recovery shape, the checked recompute of
the proof pipeline, and the Java binding's call sequence (Vkzg.java:53-184)
replayed by tests/c/java_like.c in full."""
import ctypes
import os
import random
import subprocess

import pytest

import vkzg_amd
from conftest import REPO, random_blob

pytestmark = pytest.mark.gpu

PROOFS = 128 * 48
CELL = 2048
CELLS = 128 * CELL
SENTINEL = 0xA5


def masked(src, keep, rng):
    cells = bytearray(rng.randbytes(CELLS))
    present = bytearray(128)
    for k in keep:
        cells[CELL * k:CELL * (k + 1)] = src[CELL * k:CELL * (k + 1)]
        present[k] = 1
    return bytes(cells), bytes(present)


def test_invalid_blob_in_chunk_two_keeps_chunk_one(gpu_prover_small_batch, vectors):
    """max_batch 4, 10 blobs, a non-canonical blob at index 5: the call
    returns INVALID_BLOB with chunk 1 (blobs 0-3) written and chunks 2 and 3
    untouched, as the reference's per-chunk loop leaves them."""
    p = gpu_prover_small_batch
    assert p.info()["max_batch_size"] == 4
    names = vectors.valid_compute()
    bad = vectors.blob(vectors.invalid_compute_131072()[0])
    blobs = [vectors.blob(names[i % len(names)]) for i in range(10)]
    blobs[5] = bad
    out = ctypes.create_string_buffer(bytes([SENTINEL]) * (10 * PROOFS), 10 * PROOFS)
    rc = vkzg_amd._lib.load().vkzg_compute_proofs_batch(p.handle, out, b"".join(blobs), 10)
    assert rc == vkzg_amd.ERR_INVALID_BLOB
    raw = out.raw
    for i in range(4):
        assert raw[i * PROOFS:(i + 1) * PROOFS] == vectors.proofs(names[i % len(names)]), i
    assert raw[4 * PROOFS:] == bytes([SENTINEL]) * (6 * PROOFS)


def test_too_many_missing_in_chunk_two_keeps_chunk_one(gpu_prover_small_batch, vectors, cells_of):
    """Recovery with max_batch 4: blob 6 has 65 missing cells, so the call
    returns BADARGS after chunk 1; chunk 1 holds the original cells and the
    rest of the output is untouched."""
    p = gpu_prover_small_batch
    rng = random.Random(77)
    names = vectors.valid_compute()
    ins, pms, srcs = [], [], []
    for b in range(9):
        src = cells_of(names[b % len(names)])
        keep = rng.sample(range(128), 63 if b == 6 else rng.randint(64, 128))
        c, pm = masked(src, keep, rng)
        ins.append(c)
        pms.append(pm)
        srcs.append(src)
    out = ctypes.create_string_buffer(bytes([SENTINEL]) * (9 * CELLS), 9 * CELLS)
    rc = vkzg_amd._lib.load().vkzg_recover_cells_batch(p.handle, out, b"".join(ins), b"".join(pms), 9)
    assert rc == vkzg_amd.ERR_BADARGS
    raw = out.raw
    assert raw[:4 * CELLS] == b"".join(srcs[:4])
    assert raw[4 * CELLS:] == bytes([SENTINEL]) * (5 * CELLS)


def test_recovery_64_blobs_exactly_half_present(gpu_prover, oracle_mod):
    """64 random blobs, each with exactly 64 of its 128 cells kept (a seeded
    choice per blob), garbage in the others: every blob returns its original
    cells, and the first four agree with the oracle's recovery."""
    rng = random.Random(0xC0F1)
    ins, pms, srcs = [], [], []
    for b in range(64):
        rc, src = oracle_mod.blob_to_cells(random_blob(rng))
        assert rc == 0
        c, pm = masked(src, rng.sample(range(128), 64), rng)
        assert sum(pm) == 64
        ins.append(c)
        pms.append(pm)
        srcs.append(src)
    out = gpu_prover.recover_cells_batch(b"".join(ins), b"".join(pms), 64)
    assert out == b"".join(srcs)
    rc, want = oracle_mod.recover_cells(b"".join(ins[:4]), b"".join(pms[:4]), 4)
    assert rc == 0 and out[:4 * CELLS] == want


@pytest.mark.parametrize("which", ["default", "small_batch"])
def test_checked_recompute_path_is_bit_exact(which, gpu_prover, gpu_prover_small_batch, vectors):
    """Every tile of phase A, every bucket of phase B and every output of the
    bucket reduction forced through the checked recompute (the path a
    doubling or cancellation inside the fast additions takes): the proofs of
    all 7 consensus vectors are unchanged."""
    p = gpu_prover if which == "default" else gpu_prover_small_batch
    names = vectors.valid_compute()
    blobs = b"".join(vectors.blob(n) for n in names)
    want = b"".join(vectors.proofs(n) for n in names)
    p.debug_force_checked(True)
    try:
        got = p.compute_cell_kzg_proofs_batch(blobs, len(names))
    finally:
        p.debug_force_checked(False)
    assert got == want
    assert p.compute_cell_kzg_proofs_batch(blobs, len(names)) == want


def _build_java_like(tmp_path):
    exe = tmp_path / "java_like"
    subprocess.run(["gcc", "-O1", "-o", str(exe), os.path.join(REPO, "tests", "c", "java_like.c"), "-ldl"],
                   check=True)
    return exe


def test_java_binding_call_sequence_end_to_end(tmp_path, gpu_prover, vectors, cells_of):
    """tests/c/java_like.c in full: dlopen/dlsym of the 8 symbols, a 4-byte
    options struct, VKZG_TABLES_PATH set before vkzg_prover_new (once to a
    missing file, once to a version-1 tables file written by this library),
    a 7-blob proof call memcmp'd against the consensus proofs and a 5-blob
    recover call memcmp'd against the original cells, in a fresh process."""
    exe = _build_java_like(tmp_path)
    names = vectors.valid_compute()
    (tmp_path / "blobs.bin").write_bytes(b"".join(vectors.blob(n) for n in names))
    (tmp_path / "proofs.bin").write_bytes(b"".join(vectors.proofs(n) for n in names))
    rng = random.Random(91)
    ins, pms, srcs = [], [], []
    for b in range(5):
        src = cells_of(names[b])
        c, pm = masked(src, rng.sample(range(128), rng.randint(64, 127)), rng)
        ins.append(c)
        pms.append(pm)
        srcs.append(src)
    (tmp_path / "cells.bin").write_bytes(b"".join(ins))
    (tmp_path / "present.bin").write_bytes(b"".join(pms))
    (tmp_path / "want.bin").write_bytes(b"".join(srcs))
    tables = tmp_path / "precomputed_tables.bin"
    gpu_prover.write_v1_tables(tables)
    assert tables.stat().st_size > 25_000_000
    for tables_path in [str(tmp_path / "missing" / "precomputed_tables.bin"), str(tables)]:
        r = subprocess.run([str(exe), vkzg_amd._lib.LIB_PATH, tables_path, str(tmp_path / "blobs.bin"),
                            str(tmp_path / "proofs.bin"), str(len(names)), str(tmp_path / "cells.bin"),
                            str(tmp_path / "present.bin"), str(tmp_path / "want.bin"), "5"],
                           capture_output=True, text=True, timeout=100)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "proofs ok" in r.stdout and "recover ok" in r.stdout, r.stdout
