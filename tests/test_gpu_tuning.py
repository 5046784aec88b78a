"""GPU tests of the launch-shape knobs: every tuning (phase-A work-queue tiles,
persistent grid, normalisation chunk, sub-streams) and the autotuner's choice
give bit-identical proofs, and captured HIP graphs replay to the same bytes as
kernel-by-kernel launches."""
import random

import pytest

import vkzg_amd

pytestmark = pytest.mark.gpu

PROOFS = 128 * 48
NAMES = ["compute_cells_and_kzg_proofs_case_valid_%d" % i for i in (2, 3, 4)]


@pytest.fixture
def restore_tuning(gpu_prover):
    saved = {n: gpu_prover.get_tuning(n) for n in (1, 3, 10, 40, 200)}
    yield
    for n, t in saved.items():
        gpu_prover.set_tuning(n, **t)


@pytest.mark.parametrize("groups", [1, 2, 4, 5, 10, 20])
def test_phase_a_groups_bit_exact(gpu_prover, vectors, restore_tuning, groups):
    blobs = b"".join(vectors.blob(n) for n in NAMES)
    want = b"".join(vectors.proofs(n) for n in NAMES)
    for bpc, chunk, splits, lanes in ((1, 1, 1, 8), (2, 64, 2, 16), (2, 8, 4, 32), (2, 0, 1, 4), (1, 0, 4, 64)):
        gpu_prover.set_tuning(3, phase_a_groups=groups, phase_a_blocks_per_cu=bpc, normalize_chunk=chunk,
                              ladder_splits=splits, reduce_lanes=lanes)
        assert gpu_prover.get_tuning(3)["phase_a_groups"] == groups
        assert gpu_prover.compute_cell_kzg_proofs_batch(blobs, 3) == want, (bpc, chunk, splits)


@pytest.mark.parametrize("lanes", [4, 8, 16, 32, 64])
def test_reduce_lanes_bit_exact(gpu_prover, vectors, restore_tuning, lanes):
    """Every bucket-reduction shape (64: 16 bucket ranges, four lanes sharing
    each addition), on the fast path and with every output forced through the
    checked recompute, gives the consensus proofs."""
    names = vectors.valid_compute()
    blobs = b"".join(vectors.blob(n) for n in names)
    want = b"".join(vectors.proofs(n) for n in names)
    gpu_prover.set_tuning(len(names), reduce_lanes=lanes)
    assert gpu_prover.get_tuning(len(names))["reduce_lanes"] == lanes
    assert gpu_prover.compute_cell_kzg_proofs_batch(blobs, len(names)) == want
    gpu_prover.debug_force_checked(True)
    try:
        got = gpu_prover.compute_cell_kzg_proofs_batch(blobs, len(names))
    finally:
        gpu_prover.debug_force_checked(False)
    assert got == want


def test_splits_and_odd_sizes_bit_exact(gpu_prover, vectors, restore_tuning):
    names = [NAMES[i % 3] for i in range(37)]
    blobs = b"".join(vectors.blob(n) for n in names)
    want = b"".join(vectors.proofs(n) for n in names)
    gpu_prover.set_tuning(37, phase_a_groups=5, splits=3, normalize_chunk=16, ladder_splits=2)
    assert gpu_prover.compute_cell_kzg_proofs_batch(blobs, 37) == want


@pytest.mark.parametrize("radix,waves", [(2, 1), (2, 2), (2, 4), (4, 1), (4, 2), (4, 4), (8, 1), (8, 2), (8, 4)])
def test_recovery_radix_and_waves_bit_exact(gpu_prover, vectors, cells_of, restore_tuning, radix, waves):
    rng = random.Random(radix * waves)
    names = [NAMES[i % 3] for i in range(3)]
    cells = b"".join(cells_of(n) for n in names)
    present = bytearray(128 * 3)
    for b in range(3):
        for c in rng.sample(range(128), 64 + 20 * b):
            present[128 * b + c] = 1
    gpu_prover.set_tuning(3, recovery_waves=waves, recovery_radix=radix)
    assert gpu_prover.recover_cells_batch(cells, bytes(present), 3) == cells
    out, proofs = gpu_prover.compute_cells_and_kzg_proofs_batch(b"".join(vectors.blob(n) for n in names), 3)
    assert out == cells and proofs == b"".join(vectors.proofs(n) for n in names)


def test_smaller_window_table_bit_exact(gpu_prover_small_batch, gpu_seeded_prover, vectors):
    """A second, smaller phase-A table beside the main one: tunings that name
    it give the same proofs; windows outside 10-13, at or above the main
    table's, or not resident are rejected."""
    p = gpu_seeded_prover
    main = p.info()["window_bits"]
    assert main > 10
    saved = p.get_tuning(3)
    with pytest.raises(vkzg_amd.VkzgError) as e:
        p.set_tuning(3, phase_a_window_bits=10)  # not resident yet
    assert e.value.code == vkzg_amd.ERR_BADARGS
    for bad in (9, 14, main):
        with pytest.raises(vkzg_amd.VkzgError):
            p.load_window_table(bad)
    p.load_window_table(10)
    info = p.info()
    assert info["extra_table_windows"] & (1 << 10)
    assert info["resident_table_bytes"] > info["table_bytes"]
    blobs = b"".join(vectors.blob(n) for n in NAMES)
    want = p.compute_cell_kzg_proofs_batch(blobs, 3)
    try:
        p.set_tuning(3, phase_a_window_bits=10)
        assert p.get_tuning(3)["phase_a_window_bits"] == 10
        assert p.compute_cell_kzg_proofs_batch(blobs, 3) == want
    finally:
        p.set_tuning(3, **saved)
    # the main 10-bit table of the small-batch prover gives the same proofs
    # of the mainnet vectors as its default
    assert gpu_prover_small_batch.compute_cell_kzg_proofs_batch(blobs, 3) == \
        b"".join(vectors.proofs(n) for n in NAMES)


def test_tuning_rejects_out_of_range(gpu_prover, restore_tuning):
    for bad in (dict(phase_a_groups=0), dict(phase_a_blocks_per_cu=3),
                dict(normalize_chunk=257), dict(splits=5),
                dict(recovery_waves=3), dict(recovery_waves=8), dict(recovery_radix=16), dict(ladder_splits=3),
                dict(phase_a_groups=21), dict(reduce_lanes=2), dict(reduce_lanes=12), dict(reduce_lanes=128)):
        with pytest.raises(vkzg_amd.VkzgError) as e:
            gpu_prover.set_tuning(3, **bad)
        assert e.value.code == vkzg_amd.ERR_BADARGS


def test_autotune_keeps_results(gpu_prover_small_batch, vectors):
    p = gpu_prover_small_batch
    ms = p.autotune()
    info = p.info()
    assert info["autotuned"] == 1 and ms > 0
    for n in (1, 4):
        t = p.get_tuning(n)
        assert 1 <= t["phase_a_groups"] <= 20 and 1 <= t["splits"] <= 4 and t["ladder_splits"] in (1, 2, 4)
        assert t["reduce_lanes"] in (4, 8, 16, 32, 64)
    names = [NAMES[i % 3] for i in range(6)]  # chunks of 4 + 2
    blobs = b"".join(vectors.blob(n) for n in names)
    assert p.compute_cell_kzg_proofs_batch(blobs, 6) == b"".join(vectors.proofs(n) for n in names)


def test_graph_replays_match_eager(gpu_prover, vectors, cells_of):
    info0 = gpu_prover.info()
    if not info0["graphs_enabled"]:
        pytest.skip("graphs disabled by VKZG_AMD_GRAPHS=0")
    rng = random.Random(3)
    names = [NAMES[rng.randrange(3)] for _ in range(5)]
    blobs = b"".join(vectors.blob(n) for n in names)
    want = b"".join(vectors.proofs(n) for n in names)
    cells = b"".join(cells_of(n) for n in names)
    present = bytearray(128 * 5)
    for b in range(5):
        for c in rng.sample(range(128), 64 + b * 10):
            present[128 * b + c] = 1
    # eager reference (stage timing forces kernel-by-kernel launches)
    gpu_prover.set_stage_timing(True)
    eager_p = gpu_prover.compute_cell_kzg_proofs_batch(blobs, 5)
    eager_r = gpu_prover.recover_cells_batch(cells, bytes(present), 5)
    gpu_prover.set_stage_timing(False)
    assert eager_p == want and eager_r == cells
    # graph capture, then several replays with other work in between
    for rep in range(4):
        assert gpu_prover.compute_cell_kzg_proofs_batch(blobs, 5) == want, rep
        assert gpu_prover.recover_cells_batch(cells, bytes(present), 5) == cells, rep
        assert gpu_prover.compute_cell_kzg_proofs_batch(blobs[:131072], 1) == want[:PROOFS], rep
    info1 = gpu_prover.info()
    assert info1["graph_captures"] + info1["graph_replays"] >= info0["graph_captures"] + info0["graph_replays"] + 12
    assert info1["graph_replays"] >= info0["graph_replays"] + 9
