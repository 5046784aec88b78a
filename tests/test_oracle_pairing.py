"""CPU tests of the oracle's pairing and batch cell-proof verification.

No G2 data or pairing vectors ship with the reference, so these pin the code by
what the mathematics fixes: the constants against their defining formulas,
group orders, bilinearity and non-degeneracy of the pairing, the commitment
against [p(tau)]G1, and the verifier against proofs made by definition
(accepted) and against each kind of wrong input (rejected).
"""
import random

import pytest

from conftest import SEEDED_TAU, random_blob

R = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
P = 0x1A0111EA397FE69A4B1BA7B6434BACD764774B84F38512BF6730D2A0F6B0F6241EABFFFEB153FFFFB9FEFFFFFFFFAAAB
Z = -0xD201000000010000
INF2 = bytes(192)


def test_constants_match_their_formulas(oracle_mod):
    h2 = (Z**8 - 4 * Z**7 + 5 * Z**6 - 4 * Z**4 + 6 * Z**3 - 4 * Z**2 - 4 * Z + 13)
    assert h2 % 9 == 0 and oracle_mod.g2_cofactor() == h2 // 9
    assert (P**6 + 1) % R == 0 and oracle_mod.final_exp_hard() == (P**6 + 1) // R


def test_g2_group_orders(oracle_mod):
    q = oracle_mod.g2_generator()
    assert q != INF2 and oracle_mod.g2_on_curve(q)
    assert oracle_mod.g2_mul(q, R) == INF2
    h2 = oracle_mod.g2_cofactor()
    for seed in (0, 17):
        raw = oracle_mod.g2_raw_point(seed)
        assert oracle_mod.g2_on_curve(raw)
        # #E'(Fp2) = h2 * r: the raw point has order dividing it, but not r alone
        assert oracle_mod.g2_mul(oracle_mod.g2_mul(raw, h2), R) == INF2
        assert oracle_mod.g2_mul(raw, R) != INF2
    # group law: [a]Q + [b]Q = [a+b]Q, [a]Q + [-a]Q = O
    a, b = 0x1234567, 0xFEDCBA987654321
    assert oracle_mod.g2_add(oracle_mod.g2_mul(q, a), oracle_mod.g2_mul(q, b)) == oracle_mod.g2_mul(q, a + b)
    assert oracle_mod.g2_add(oracle_mod.g2_mul(q, a), oracle_mod.g2_mul(q, R - a)) == INF2


def test_pairing_bilinear_and_nondegenerate(oracle_mod):
    rng = random.Random(5)
    q = oracle_mod.g2_generator()
    g = oracle_mod.g1_mul_gen(1)
    inf1 = oracle_mod.g1_mul_gen(0)
    e = oracle_mod.pairing(g, q)
    assert e != oracle_mod.pairing(inf1, q)  # e(P, Q) != 1
    for _ in range(2):
        a, b = rng.randrange(1, R), rng.randrange(1, R)
        lhs = oracle_mod.pairing(oracle_mod.g1_mul_gen(a), oracle_mod.g2_mul(q, b))
        assert lhs == oracle_mod.pairing(oracle_mod.g1_mul_gen(a * b % R), q)
        assert lhs == oracle_mod.pairing(g, oracle_mod.g2_mul(q, a * b % R))
    neg = oracle_mod.g1_mul_gen(R - 1)
    assert oracle_mod.pairing_check([(g, q), (neg, q)])
    assert not oracle_mod.pairing_check([(g, q), (g, q)])
    q2 = oracle_mod.g2_mul(q, 987654321)
    assert oracle_mod.pairing_check([(g, q), (g, q2), (neg, oracle_mod.g2_add(q, q2))])
    # e(P, Q)^r = 1 through bilinearity: e([r-1]P, Q) e(P, Q) = 1
    assert oracle_mod.pairing_check([(neg, q), (g, q)])


def test_commitment_is_p_of_tau(oracle_mod, seeded_verify_setup):
    rng = random.Random(6)
    blob = random_blob(rng)
    rc, coeffs = oracle_mod.blob_to_poly(blob)
    assert rc == 0
    acc = 0
    for k in range(4095, -1, -1):
        acc = (acc * SEEDED_TAU + int.from_bytes(coeffs[32 * k:32 * k + 32], "big")) % R
    assert seeded_verify_setup.commit(blob) == oracle_mod.g1_mul_gen(acc)


def test_verifier_accepts_proofs_by_definition(oracle_mod, seeded_verify_setup, seeded_setup_bytes):
    rng = random.Random(8)
    blobs = [random_blob(rng), bytes(131072)]  # a random blob and the zero blob
    blobs.append(b"".join((7 if k == 100 else 0).to_bytes(32, "big") for k in range(4096)))
    proofs = b"".join(oracle_mod.proofs_from_tau(b, SEEDED_TAU)[1] for b in blobs)
    assert seeded_verify_setup.verify_blobs(b"".join(blobs), proofs, len(blobs))
    # the FK20 oracle's proofs on the same setup verify too
    rc, fk = oracle_mod.Setup(seeded_setup_bytes).compute_proofs(b"".join(blobs), len(blobs))
    assert rc == 0 and seeded_verify_setup.verify_blobs(b"".join(blobs), fk, len(blobs), seed=99)


def test_verifier_rejects_wrong_inputs(oracle_mod, seeded_verify_setup):
    rng = random.Random(9)
    blob = random_blob(rng)
    other = random_blob(rng)
    proofs = oracle_mod.proofs_from_tau(blob, SEEDED_TAU)[1]
    cells = oracle_mod.blob_to_cells(blob)[1]
    com = seeded_verify_setup.commit(blob)
    idx = list(range(8))
    coms = com * 8
    ok = seeded_verify_setup.verify(coms, idx, cells[:8 * 2048], proofs[:8 * 48])
    assert ok
    # proof of the wrong cell
    bad = bytearray(proofs[:8 * 48])
    bad[48 * 2:48 * 3] = proofs[48 * 3:48 * 4]
    assert not seeded_verify_setup.verify(coms, idx, cells[:8 * 2048], bytes(bad))
    # one cell element changed
    badc = bytearray(cells[:8 * 2048])
    badc[2048 * 5 + 31] ^= 1
    assert not seeded_verify_setup.verify(coms, idx, bytes(badc), proofs[:8 * 48])
    # wrong commitment for one item
    badcom = bytearray(coms)
    badcom[48 * 7:48 * 8] = seeded_verify_setup.commit(other)
    assert not seeded_verify_setup.verify(bytes(badcom), idx, cells[:8 * 2048], proofs[:8 * 48])
    # cells listed under the wrong index
    assert not seeded_verify_setup.verify(coms, [1, 0] + idx[2:], cells[:8 * 2048], proofs[:8 * 48])
    # malformed input raises
    with pytest.raises(ValueError):
        seeded_verify_setup.verify(coms[:48], [128], cells[:2048], proofs[:48])


def test_glv_endomorphism_constants(oracle_mod, mainnet_setup_bytes):
    """The device's phase-B split uses phi(x, y) = (beta x, y) = [lambda] P with
    lambda = z^2 - 1; check that pairing of beta and lambda on G1 points."""
    import importlib.util
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    spec = importlib.util.spec_from_file_location(
        "gen_constants", os.path.join(here, "..", "vkzg_amd", "csrc", "gen_constants.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    beta = gen.BETA
    lam = (Z * Z - 1) % R
    assert lam == 0xAC45A4010001A40200000000FFFFFFFF and (lam * lam + lam + 1) % R == 0
    assert pow(beta, 3, P) == 1 and beta != 1
    for pt in (oracle_mod.g1_mul_gen(1), mainnet_setup_bytes[48 * 7:48 * 8], mainnet_setup_bytes[-48:]):
        lp = oracle_mod.g1_mul(pt, lam)
        x = int.from_bytes(bytes([pt[0] & 0x1F]) + pt[1:], "big")
        lx = int.from_bytes(bytes([lp[0] & 0x1F]) + lp[1:], "big")
        assert lx == beta * x % P
        assert (lp[0] & 0x20) == (pt[0] & 0x20)  # same y
