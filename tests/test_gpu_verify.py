"""GPU tests of batch cell-proof verification, commitments, the device pairing
and the device MSM against the CPU oracle (seeded-tau setup: the oracle
supplies [tau^i]G1 and [tau^i]Q, so every verdict is decided by the
mathematics, not by stored answers).

  * pairing: e(P, Q)^3 from the device equals the oracle's e(3P, Q) (the
    device's final exponentiation computes the cube, see pairing.hip);
  * MSM: bit-exact against the oracle's Pippenger on random, repeated,
    identity and extreme scalars;
  * commitments: bit-exact against the oracle's;
  * verification: the verdict on valid and tampered batches, the two G1
    inputs of the pairing check bit-exact against the oracle's spec-weighted
    verifier with the same Fiat-Shamir challenge, and the error cases of the
    specs' assertions.
"""
import random

import pytest

pytestmark = pytest.mark.gpu

R = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
P = 0x1A0111EA397FE69A4B1BA7B6434BACD764774B84F38512BF6730D2A0F6B0F6241EABFFFEB153FFFFB9FEFFFFFFFFAAAB
LAMBDA = 0xD201000000010000 ** 2 - 1
G1_INF = bytes([0xC0]) + bytes(47)


@pytest.fixture(scope="module")
def g2_setup(oracle_mod):
    from tests.conftest import SEEDED_TAU
    raw = oracle_mod.g2_powers_from_tau(SEEDED_TAU, 65)
    return b"".join(oracle_mod.g2_compress(raw[192 * i:192 * (i + 1)]) for i in range(65))


@pytest.fixture()
def verifier(gpu_seeded_prover, g2_setup):
    gpu_seeded_prover.load_g2_setup(g2_setup)
    return gpu_seeded_prover


def _blob_data(prover, blobs, n):
    cells, proofs = prover.compute_cells_and_kzg_proofs_batch(blobs, n)
    coms = prover.blob_to_kzg_commitments(blobs, n)
    return cells, proofs, coms


def _select(cells, proofs, coms, picks):
    """picks: (blob, cell index) pairs -> (commitments, indices, cells, proofs)."""
    c = b"".join(coms[48 * b:48 * (b + 1)] for b, _ in picks)
    idx = [i for _, i in picks]
    ce = b"".join(cells[(128 * b + i) * 2048:(128 * b + i + 1) * 2048] for b, i in picks)
    pr = b"".join(proofs[(128 * b + i) * 48:(128 * b + i + 1) * 48] for b, i in picks)
    return c, idx, ce, pr


def _non_group_point():
    """A point of E(Fp) outside G1, compressed (x = 1, 2, ... with a square x^3 + 4)."""
    for x in range(1, 1000):
        rhs = (x ** 3 + 4) % P
        y = pow(rhs, (P + 1) // 4, P)
        if y * y % P == rhs:
            b = bytearray(x.to_bytes(48, "big"))
            b[0] |= 0x80 | (0x20 if y > (P - 1) // 2 else 0)
            return bytes(b)
    raise AssertionError


def test_pairing_matches_oracle(gpu_seeded_prover, oracle_mod):
    rng = random.Random(21)
    q0 = oracle_mod.g2_generator()
    g1s, g2s, want = [], [], []
    for _ in range(3):
        a, b = rng.randrange(1, R), rng.randrange(1, R)
        p = oracle_mod.g1_mul_gen(a)
        q = oracle_mod.g2_mul(q0, b)
        g1s.append(p)
        g2s.append(oracle_mod.g2_compress(q))
        want.append(oracle_mod.pairing(oracle_mod.g1_mul(p, 3), q))
    got = gpu_seeded_prover.debug_pairing(b"".join(g1s), b"".join(g2s))
    for i in range(3):
        dev = got[576 * i:576 * (i + 1)]
        ora = want[i]
        for d in range(2):
            for k in range(3):
                for e in range(2):
                    m = 2 * k + d
                    assert dev[48 * (2 * m + e):48 * (2 * m + e + 1)] == ora[48 * (6 * d + 2 * k + e):
                                                                          48 * (6 * d + 2 * k + e + 1)]
    # P at infinity: e = 1
    one = gpu_seeded_prover.debug_pairing(G1_INF, g2s[0])
    assert int.from_bytes(one[:48], "big") == 1 and one[48:] == bytes(528)


@pytest.mark.parametrize("n", [1, 2, 3, 64, 333, 2048])
def test_msm_matches_oracle(gpu_seeded_prover, oracle_mod, n):
    rng = random.Random(1000 + n)
    pts = [oracle_mod.g1_mul_gen(rng.randrange(1, R)) for _ in range(n)]
    ks = [rng.randrange(R) for _ in range(n)]
    if n >= 3:
        pts[1] = G1_INF
        pts[2] = pts[0]            # the same point twice with the same scalar: a doubling in a bucket
        ks[2] = ks[0]
    if n >= 64:
        ks[3], ks[4], ks[5], ks[6], ks[7] = 0, 1, R - 1, LAMBDA, LAMBDA + 1
        pts[8] = pts[9]
    points = b"".join(pts)
    scalars = b"".join(k.to_bytes(32, "big") for k in ks)
    assert gpu_seeded_prover.debug_g1_msm(points, scalars) == oracle_mod.g1_msm(points, scalars)


def test_msm_cancellation(gpu_seeded_prover, oracle_mod):
    p = oracle_mod.g1_mul_gen(12345)
    pts = p + p
    scal = (7).to_bytes(32, "big") + (R - 7).to_bytes(32, "big")
    assert gpu_seeded_prover.debug_g1_msm(pts, scal) == G1_INF


def test_commitments_match_oracle(gpu_seeded_prover, seeded_verify_setup, rng):
    from tests.conftest import random_blob
    blobs = [random_blob(rng) for _ in range(3)]
    blobs.append(bytes(131072))                                  # zero polynomial
    blobs.append(bytes(131071) + b"\x05")                         # one nonzero element
    data = b"".join(blobs)
    got = gpu_seeded_prover.blob_to_kzg_commitments(data, len(blobs))
    assert got == seeded_verify_setup.commit_batch(data, len(blobs))
    assert got[48 * 3:48 * 4] == G1_INF
    import vkzg_amd
    bad = bytearray(data)
    bad[131072 * 2:131072 * 2 + 32] = b"\xff" * 32
    with pytest.raises(vkzg_amd.VkzgError) as e:
        gpu_seeded_prover.blob_to_kzg_commitments(bytes(bad), len(blobs))
    assert e.value.code == vkzg_amd.ERR_INVALID_BLOB


def test_commitments_large_batch(gpu_seeded_prover, seeded_verify_setup, rng):
    """24 blobs in one call: over 2^18 buckets, the MSM's large-launch regime
    (16 buckets per window thread, hot buckets summed by 16-lane teams).  Each
    commitment equals the one-blob call's (the small regime) and, for three of
    them, the oracle's."""
    from tests.conftest import random_blob
    blobs = [random_blob(rng) for _ in range(23)] + [bytes(131072)]
    got = gpu_seeded_prover.blob_to_kzg_commitments(b"".join(blobs), len(blobs))
    for i, b in enumerate(blobs):
        assert got[48 * i:48 * (i + 1)] == gpu_seeded_prover.blob_to_kzg_commitments(b, 1), i
    for i in (0, 11, 22):
        assert got[48 * i:48 * (i + 1)] == seeded_verify_setup.commit_batch(blobs[i], 1), i
    assert got[48 * 23:] == G1_INF


def test_verify_valid_and_terms(verifier, seeded_verify_setup, oracle_mod, rng):
    from tests.conftest import random_blob
    blobs = b"".join(random_blob(rng) for _ in range(3))
    cells, proofs, coms = _blob_data(verifier, blobs, 3)
    assert coms == seeded_verify_setup.commit_batch(blobs, 3)
    picks = [(0, i) for i in range(128)]
    picks += [(1, i) for i in rng.sample(range(128), 40)]
    picks += [(2, 9), (2, 9), (2, 100)]      # a repeated cell
    rng.shuffle(picks)
    c, idx, ce, pr = _select(cells, proofs, coms, picks)
    assert verifier.verify_cell_kzg_proof_batch(c, idx, ce, pr)
    ok, a, b = verifier.debug_verify_terms(c, idx, ce, pr)
    r = oracle_mod.verify_challenge(c, idx, ce, pr)
    import vkzg_amd
    assert vkzg_amd.verify_challenge(c, idx, ce, pr) == r
    ok_o, a_o, b_o = seeded_verify_setup.verify_r(c, idx, ce, pr, r)
    assert ok and ok_o and a == a_o and b == b_o
    # one cell alone, and a single blob
    c1, i1, ce1, pr1 = _select(cells, proofs, coms, [(2, 127)])
    assert verifier.verify_cell_kzg_proof_batch(c1, i1, ce1, pr1)
    assert verifier.verify_cell_kzg_proof_batch(b"", [], b"", b"")


def test_verify_rejects_inconsistent_batches(verifier, rng):
    from tests.conftest import random_blob
    blobs = b"".join(random_blob(rng) for _ in range(2))
    cells, proofs, coms = _blob_data(verifier, blobs, 2)
    picks = [(0, 3), (0, 64), (1, 17), (1, 90)]
    c, idx, ce, pr = _select(cells, proofs, coms, picks)
    assert verifier.verify_cell_kzg_proof_batch(c, idx, ce, pr)
    t = bytearray(ce)
    t[2048 * 2 + 32 * 5 + 31] ^= 0x01            # one element changed (still canonical)
    assert not verifier.verify_cell_kzg_proof_batch(c, idx, bytes(t), pr)
    swapped = pr[48:96] + pr[:48] + pr[96:]      # proofs of cells 0 and 1 exchanged
    assert not verifier.verify_cell_kzg_proof_batch(c, idx, ce, swapped)
    wrong_c = coms[48:96] + c[48:]               # cell 0 claimed for blob 1
    assert not verifier.verify_cell_kzg_proof_batch(wrong_c, idx, ce, pr)
    assert not verifier.verify_cell_kzg_proof_batch(c, [64, 3] + idx[2:], ce, pr)


def test_verify_identity_proofs(verifier):
    blobs = bytes(131072)
    cells, proofs, coms = _blob_data(verifier, blobs, 1)
    assert coms == G1_INF and proofs == G1_INF * 128
    c, idx, ce, pr = _select(cells, proofs, coms, [(0, i) for i in range(0, 128, 3)])
    assert verifier.verify_cell_kzg_proof_batch(c, idx, ce, pr)


def test_verify_errors(verifier, gpu_prover, rng):
    import vkzg_amd
    from tests.conftest import random_blob
    blobs = random_blob(rng)
    cells, proofs, coms = _blob_data(verifier, blobs, 1)
    c, idx, ce, pr = _select(cells, proofs, coms, [(0, 1), (0, 2)])

    def code(*args):
        with pytest.raises(vkzg_amd.VkzgError) as e:
            verifier.verify_cell_kzg_proof_batch(*args)
        return e.value.code

    noncanon = bytearray(ce)
    noncanon[2048 + 32 * 7:2048 + 32 * 8] = R.to_bytes(32, "big")
    assert code(c, idx, bytes(noncanon), pr) == vkzg_amd.ERR_BADARGS
    assert code(c, [1, 128], ce, pr) == vkzg_amd.ERR_BADARGS
    noflag = bytearray(pr)
    noflag[48] &= 0x7F
    assert code(c, idx, ce, bytes(noflag)) == vkzg_amd.ERR_BADARGS
    outside = _non_group_point()
    assert code(c, idx, ce, pr[:48] + outside) == vkzg_amd.ERR_BADARGS
    assert code(outside + c[48:], idx, ce, pr) == vkzg_amd.ERR_BADARGS
    with pytest.raises(vkzg_amd.VkzgError) as e:
        gpu_prover.verify_cell_kzg_proof_batch(c, idx, ce, pr)   # no G2 setup loaded
    assert e.value.code == vkzg_amd.ERR_SETUP
    # malformed G2 setups, then the good one again
    with pytest.raises(vkzg_amd.VkzgError) as e:
        verifier.load_g2_setup(bytes(96 * 64))
    assert e.value.code == vkzg_amd.ERR_BADARGS
    assert verifier.verify_cell_kzg_proof_batch(c, idx, ce, pr)


def test_verify_bad_g2_setup(gpu_seeded_prover, g2_setup):
    import vkzg_amd
    bad = bytearray(g2_setup)
    bad[96 * 5] &= 0x7F                          # no compression flag
    with pytest.raises(vkzg_amd.VkzgError) as e:
        gpu_seeded_prover.load_g2_setup(bytes(bad))
    assert e.value.code == vkzg_amd.ERR_SETUP
    inf = bytearray(g2_setup)
    inf[96 * 64:96 * 65] = bytes([0xC0]) + bytes(95)
    with pytest.raises(vkzg_amd.VkzgError) as e:
        gpu_seeded_prover.load_g2_setup(bytes(inf))
    assert e.value.code == vkzg_amd.ERR_SETUP
    gpu_seeded_prover.load_g2_setup(g2_setup)


def test_verify_64_blobs_all_cells(verifier, rng):
    """8192 cells of 64 blobs in one call (the size of one block's columns)."""
    from tests.conftest import random_blob
    n = 64
    blobs = b"".join(random_blob(rng) for _ in range(n))
    cells, proofs, coms = _blob_data(verifier, blobs, n)
    c = b"".join(coms[48 * b:48 * (b + 1)] * 128 for b in range(n))
    idx = list(range(128)) * n
    assert verifier.verify_cell_kzg_proof_batch(c, idx, cells, proofs)
    t = bytearray(proofs)
    t[48 * 5000:48 * 5001] = proofs[48 * 5001:48 * 5002]
    assert not verifier.verify_cell_kzg_proof_batch(c, idx, cells, bytes(t))
