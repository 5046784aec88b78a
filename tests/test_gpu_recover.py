"""Cell recovery on the GPU through the C ABI, bit-exact against the
consensus vectors and the oracle (reference tests/test.cpp:204-276)."""
import hashlib
import random

import pytest

import vkzg_amd

pytestmark = pytest.mark.gpu

CELL = 2048
CELLS = 128 * CELL


def mask_cells(src, keep, rng=None):
    cells = bytearray(rng.randbytes(CELLS)) if rng else bytearray(CELLS)
    present = bytearray(128)
    for k in keep:
        cells[CELL * k:CELL * (k + 1)] = src[CELL * k:CELL * (k + 1)]
        present[k] = 1
    return bytes(cells), bytes(present)


def test_recover_vectors(gpu_prover, vectors, cells_of):
    valid = []
    for case in vectors.recover_cases():
        cells, present = mask_cells(cells_of(case["source"]), case["cell_indices"])
        if case["valid"]:
            out = gpu_prover.recover_cells_batch(cells, present, 1)
            assert hashlib.sha256(out).hexdigest() == case["cells_sha256"], case["name"]
            valid.append((cells, present, out))
        else:
            with pytest.raises(vkzg_amd.VkzgError) as e:
                gpu_prover.recover_cells_batch(cells, present, 1)
            assert e.value.code == vkzg_amd.ERR_BADARGS
    # all valid vectors in one batched call
    n = len(valid)
    out = gpu_prover.recover_cells_batch(b"".join(v[0] for v in valid),
                                         b"".join(v[1] for v in valid), n)
    assert out == b"".join(v[2] for v in valid)


def test_recover_random_patterns_match_originals(gpu_prover, vectors, cells_of):
    rng = random.Random(21)
    names = vectors.valid_compute()
    ins, pms, want = [], [], []
    for i in range(24):
        src = cells_of(names[i % len(names)])
        keep = rng.sample(range(128), rng.randint(64, 128))
        c, p = mask_cells(src, keep, rng)  # random garbage in the missing cells
        ins.append(c)
        pms.append(p)
        want.append(src)
    out = gpu_prover.recover_cells_batch(b"".join(ins), b"".join(pms), len(ins))
    assert out == b"".join(want)


def test_recover_inconsistent_cells_match_oracle(gpu_prover, oracle_mod):
    """Present cells that are not a codeword (random bytes, some >= r): the
    output is the reference's linear map of them, bit for bit."""
    rng = random.Random(22)
    n = 3
    cells = rng.randbytes(CELLS * n)
    masks = []
    for b in range(n):
        keep = set(rng.sample(range(128), 64 + 20 * b))
        masks.append(bytes(1 if k in keep else 0 for k in range(128)))
    present = b"".join(masks)
    rc, want = oracle_mod.recover_cells(cells, present, n)
    assert rc == 0
    assert gpu_prover.recover_cells_batch(cells, present, n) == want


@pytest.mark.parametrize("radix", [2, 8])
def test_recover_modes_match_oracle_on_inconsistent_cells(gpu_prover, oracle_mod, radix):
    """The three-kernel (radix 2) and radix-8 recoveries on present cells that
    are not a codeword, uneven present counts (64..127, all present, all but
    one pattern per blob): the reference's linear map, bit for bit."""
    rng = random.Random(40 + radix)
    n = 7
    cells = rng.randbytes(CELLS * n)
    masks = []
    for b in range(n):
        keep = set(rng.sample(range(128), [64, 65, 90, 127, 128, 64, 100][b]))
        masks.append(bytes(1 if k in keep else 0 for k in range(128)))
    present = b"".join(masks)
    rc, want = oracle_mod.recover_cells(cells, present, n)
    assert rc == 0
    saved = gpu_prover.get_tuning(n)
    try:
        gpu_prover.set_tuning(n, recovery_radix=radix)
        assert gpu_prover.recover_cells_batch(cells, present, n) == want
    finally:
        gpu_prover.set_tuning(n, **saved)


def test_recover_chunking_and_bad_chunk(gpu_prover_small_batch, vectors, cells_of):
    rng = random.Random(23)
    src = cells_of("compute_cells_and_kzg_proofs_case_valid_2")
    ins, pms = [], []
    for _ in range(9):
        c, p = mask_cells(src, rng.sample(range(128), 70), rng)
        ins.append(c)
        pms.append(p)
    out = gpu_prover_small_batch.recover_cells_batch(b"".join(ins), b"".join(pms), 9)
    assert out == src * 9
    # a blob with 65 missing cells in the second chunk -> BADARGS
    bad_c, bad_p = mask_cells(src, range(63))
    pms[5] = bad_p
    with pytest.raises(vkzg_amd.VkzgError) as e:
        gpu_prover_small_batch.recover_cells_batch(b"".join(ins), b"".join(pms), 9)
    assert e.value.code == vkzg_amd.ERR_BADARGS


def test_recover_zero_blobs_and_nulls(gpu_prover):
    assert gpu_prover.recover_cells_batch(b"", b"", 0) == b""
    L = vkzg_amd._lib.load()
    import ctypes
    buf = ctypes.create_string_buffer(CELLS)
    for args in [(None, buf, buf), (buf, None, buf), (buf, buf, None)]:
        assert L.vkzg_recover_cells_batch(gpu_prover.handle, *args, 1) == vkzg_amd.ERR_BADARGS


def test_uneven_batch_128_blobs(gpu_prover, oracle_mod, rng):
    """Many blobs with different missing sets (distinct vanishing polynomials)
    in one call; every blob must return its original cells."""
    from conftest import random_blob
    local = random.Random(24)
    srcs = []
    for i in range(4):
        rc, cells = oracle_mod.blob_to_cells(random_blob(local))
        assert rc == 0
        srcs.append(cells)
    ins, pms, want = [], [], []
    for b in range(128):
        src = srcs[b % 4]
        c, p = mask_cells(src, local.sample(range(128), local.randint(64, 127)), local)
        ins.append(c)
        pms.append(p)
        want.append(src)
    out = gpu_prover.recover_cells_batch(b"".join(ins), b"".join(pms), 128)
    assert out == b"".join(want)


def test_device_entry_point_matches_host(gpu_prover, vectors, cells_of):
    from vkzg_amd import hip
    rng = random.Random(25)
    src = cells_of("compute_cells_and_kzg_proofs_case_valid_5")
    ins, pms = [], []
    for _ in range(5):
        c, p = mask_cells(src, rng.sample(range(128), 80), rng)
        ins.append(c)
        pms.append(p)
    d_cells = hip.DeviceBuffer.from_bytes(b"".join(ins))
    d_present = hip.DeviceBuffer.from_bytes(b"".join(pms))
    d_out = hip.DeviceBuffer(CELLS * 5)
    gpu_prover.recover_cells_device(d_out.ptr, d_cells.ptr, d_present.ptr, b"".join(pms), 5)
    assert d_out.to_bytes() == src * 5
