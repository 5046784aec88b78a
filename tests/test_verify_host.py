"""CPU tests of the host side of batch cell-proof verification and of the
oracle functions the GPU verification tests rely on:

  * the library's SHA-256 (SHA-extension and portable paths) against hashlib
    and the FIPS 180-4 example digests;
  * the library's Fiat-Shamir challenge against the oracle's restatement of
    the consensus specs' compute_verify_cell_kzg_proof_batch_challenge
    (hashlib), with duplicated commitments;
  * the GLV split of vkzg_amd/csrc/glv.h (host build) against Python integers;
  * the oracle's G2 compression against the standard BLS12-381 G2 generator
    encoding, its MSM against a plain sum of scalar multiples, and its
    spec-weighted verifier against its seeded one.
"""
import hashlib
import os
import random
import subprocess

import pytest

R = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
LAMBDA = 0xD201000000010000 ** 2 - 1
HERE = os.path.dirname(os.path.abspath(__file__))

# The BLS12-381 G2 generator (x.c1 first, big-endian, compression flag set),
# the standard encoding of the IETF pairing-friendly-curves draft and of
# every BLS12-381 library.
G2_GEN_COMPRESSED = bytes.fromhex(
    "93e02b6052719f607dacd3a088274f65596bd0d09920b61ab5da61bbdc7f5049334cf11213945d57e5ac7d055d042b7e"
    "024aa2b2f08f0a91260805272dc51051c6e47ad4fa403b02b4510b647ae3d1770bac0326a805bbefd48056c8c121bdb8")
G2_GEN_Y = (0x0ce5d527727d6e118cc9cdc6da2e351aadfd9baa8cbdd3a76d429a695160d12c923ac9cc3baca289e193548608b82801,
            0x0606c4a02ea734cc32acd2b02bc28b99cb3e287e85a763af267492ab572e99ab3f370d275cec1da1aaa9075ff05f79be)


def test_sha256_fips_examples_and_hashlib():
    import vkzg_amd
    kat = {
        b"abc": "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad",
        b"": "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855",
        b"abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq":
            "248d6a61d20638b8e5c026930c3e6039a33ce45964ff2167f6ecedd419db06c1",
    }
    paths = [True, False] if vkzg_amd.sha256_hw() else [True]
    for portable in paths:
        for msg, hexd in kat.items():
            assert vkzg_amd.sha256(msg, portable).hex() == hexd
        rng = random.Random(7)
        for n in list(range(0, 200)) + [447, 448, 511, 512, 513, 4095, 65536, 200003]:
            d = rng.randbytes(n)
            assert vkzg_amd.sha256(d, portable) == hashlib.sha256(d).digest(), (n, portable)


def test_challenge_matches_spec_restatement(oracle_mod):
    import vkzg_amd
    rng = random.Random(11)
    for n, ncom in [(1, 1), (3, 2), (16, 3), (130, 5)]:
        coms_u = [rng.randbytes(48) for _ in range(ncom)]
        coms = b"".join(coms_u[rng.randrange(ncom)] for _ in range(n))
        idx = [rng.randrange(128) for _ in range(n)]
        cells = rng.randbytes(2048 * n)
        proofs = rng.randbytes(48 * n)
        want = oracle_mod.verify_challenge(coms, idx, cells, proofs)
        assert int.from_bytes(want, "big") < R
        assert vkzg_amd.verify_challenge(coms, idx, cells, proofs) == want
        assert vkzg_amd.verify_challenge(coms, idx, cells, proofs, portable_sha=True) == want
    # the order of commitments and indices is bound into r
    coms = rng.randbytes(48) + rng.randbytes(48)
    cells, proofs = rng.randbytes(4096), rng.randbytes(96)
    a = vkzg_amd.verify_challenge(coms, [3, 5], cells, proofs)
    assert a != vkzg_amd.verify_challenge(coms, [5, 3], cells, proofs)
    assert a != vkzg_amd.verify_challenge(coms[48:] + coms[:48], [3, 5], cells, proofs)


def test_glv_split_host(tmp_path):
    exe = str(tmp_path / "glv_host")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "c", "glv_host.cpp")], check=True)
    rng = random.Random(99)
    ks = [0, 1, 2, LAMBDA - 1, LAMBDA, LAMBDA + 1, 2 * LAMBDA, LAMBDA * LAMBDA, LAMBDA * LAMBDA + LAMBDA,
          R - 1, R - 2, R - LAMBDA, (1 << 127), (1 << 128) - 1, (1 << 254), R - (1 << 128)]
    ks += [q * LAMBDA + rr for q in (0, 1, LAMBDA - 1, LAMBDA, LAMBDA + 1) for rr in (0, 1, LAMBDA - 1)
           if q * LAMBDA + rr < R]
    ks += [rng.randrange(R) for _ in range(20000)]
    out = subprocess.run([exe], input="\n".join("%x" % k for k in ks) + "\n", capture_output=True, text=True,
                         check=True).stdout.split()
    assert len(out) == 2 * len(ks)
    for i, k in enumerate(ks):
        k1, k2 = int(out[2 * i], 16), int(out[2 * i + 1], 16)
        assert k1 + LAMBDA * k2 == k, hex(k)
        assert k1 < LAMBDA and k2 <= LAMBDA + 1, hex(k)


def test_oracle_g2_encoding_standard_generator(oracle_mod):
    q = oracle_mod.g2_decompress(G2_GEN_COMPRESSED)
    assert oracle_mod.g2_on_curve(q)
    assert oracle_mod.g2_mul(q, R) == bytes(192)  # order r
    assert int.from_bytes(q[96:144], "big") == G2_GEN_Y[0] and int.from_bytes(q[144:], "big") == G2_GEN_Y[1]
    assert oracle_mod.g2_compress(q) == G2_GEN_COMPRESSED
    neg = oracle_mod.g2_mul(q, R - 1)
    c = oracle_mod.g2_compress(neg)
    assert c[0] == G2_GEN_COMPRESSED[0] | 0x20 and c[1:] == G2_GEN_COMPRESSED[1:]
    assert oracle_mod.g2_decompress(c) == neg
    inf = bytes([0xc0]) + bytes(95)
    assert oracle_mod.g2_decompress(inf) == bytes(192)
    with pytest.raises(ValueError):
        oracle_mod.g2_decompress(bytes(96))  # no compression flag


def test_oracle_g1_msm_against_scalar_multiples(oracle_mod):
    rng = random.Random(5)
    pts = [oracle_mod.g1_mul_gen(rng.randrange(1, R)) for _ in range(9)]
    pts.append(bytes([0xc0]) + bytes(47))  # identity
    pts.append(pts[0])                     # repeated point
    ks = [rng.randrange(R) for _ in pts[:-3]] + [0, R - 1, 5]
    want = bytes([0xc0]) + bytes(47)
    for p, k in zip(pts, ks):
        want = oracle_mod.g1_add(want, oracle_mod.g1_mul(p, k))
    got = oracle_mod.g1_msm(b"".join(pts), b"".join(k.to_bytes(32, "big") for k in ks))
    assert got == want


def test_oracle_spec_weighted_verify(oracle_mod, seeded_verify_setup, rng):
    from tests.conftest import random_blob
    blob = random_blob(rng)
    from tests.conftest import SEEDED_TAU
    rc, proofs = oracle_mod.proofs_from_tau(blob, SEEDED_TAU)
    assert rc == 0
    com = seeded_verify_setup.commit(blob)
    cells = oracle_mod.blob_to_cells(blob)[1]
    idx = [5, 77, 0, 127]
    sel_cells = b"".join(cells[2048 * i:2048 * (i + 1)] for i in idx)
    sel_proofs = b"".join(proofs[48 * i:48 * (i + 1)] for i in idx)
    coms = com * len(idx)
    r = oracle_mod.verify_challenge(coms, idx, sel_cells, sel_proofs)
    ok, a, b = seeded_verify_setup.verify_r(coms, idx, sel_cells, sel_proofs, r)
    assert ok and seeded_verify_setup.verify(coms, idx, sel_cells, sel_proofs)
    # A = sum r^k pi_k, recomputed from scalar multiples
    rv = int.from_bytes(r, "big")
    want = bytes([0xc0]) + bytes(47)
    for k, i in enumerate(idx):
        want = oracle_mod.g1_add(want, oracle_mod.g1_mul(proofs[48 * i:48 * (i + 1)], pow(rv, k, R)))
    assert a == want
    bad_cells = bytearray(sel_cells)
    bad_cells[2048 + 31] ^= 1
    assert not seeded_verify_setup.verify_r(coms, idx, bytes(bad_cells), sel_proofs, r)[0]
    assert not seeded_verify_setup.verify(coms, idx, bytes(bad_cells), sel_proofs)


def test_seeded_setup_fixture(seeded_setup_bytes, oracle_mod):
    """tests/golden/seeded_setup.bin (read by bench.py) is the oracle's seeded setup."""
    from tests.conftest import GOLDEN, SEEDED_TAU
    data = open(os.path.join(GOLDEN, "seeded_setup.bin"), "rb").read()
    assert len(data) == 4096 * 48 + 65 * 96
    assert data[:4096 * 48] == seeded_setup_bytes
    raw = oracle_mod.g2_powers_from_tau(SEEDED_TAU, 65)
    for i in (0, 1, 63, 64):
        assert oracle_mod.g2_decompress(data[4096 * 48 + 96 * i:4096 * 48 + 96 * (i + 1)]) == raw[192 * i:192 * (i + 1)]
