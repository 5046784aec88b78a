"""CPU tests: pin the oracle against the reference's consensus vectors and
against what the mathematics fixes (closed forms, brute force, identities).

Each check is chosen so that a plausible mistake in the oracle (a wrong bit
reversal, a dropped 1/n, a transposed circulant, a missing sign) fails it.
"""
import hashlib
import random

import pytest

from conftest import SEEDED_TAU, random_blob

R = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
P = 0x1A0111EA397FE69A4B1BA7B6434BACD764774B84F38512BF6730D2A0F6B0F6241EABFFFEB153FFFFB9FEFFFFFFFFAAAB


def brev(x, bits):
    return int(format(x, "0%db" % bits)[::-1], 2)


# ------------------------------------------------------------------ fields

def test_field_mul_matches_bigint(oracle_mod):
    rng = random.Random(1)
    edge_r = [0, 1, R - 1, R - 2, (1 << 255) - 1 - R, 2]
    edge_p = [0, 1, P - 1, P - 2, 2]
    for a in edge_r + [rng.randrange(R) for _ in range(200)]:
        b = rng.randrange(R)
        assert oracle_mod.fr_mul(a % R, b) == a % R * b % R
    for a in edge_p + [rng.randrange(P) for _ in range(200)]:
        b = rng.randrange(P)
        assert oracle_mod.fp_mul(a, b) == a * b % P


def test_fr_inverse_and_reduction(oracle_mod):
    rng = random.Random(2)
    for _ in range(20):
        a = rng.randrange(1, R)
        assert oracle_mod.fr_inv(a) == pow(a, -1, R)
    # Values >= r are reduced mod r (k_cells_to_fr.comp semantics).
    big = (1 << 256) - 1
    assert oracle_mod.fr_mul(big, 1) == big % R


def test_root_of_unity(oracle_mod):
    w = oracle_mod.root_8192(1)
    assert w == pow(7, (R - 1) // 8192, R)
    assert pow(w, 8192, R) == 1 and pow(w, 4096, R) == R - 1
    assert oracle_mod.root_8192(5) == pow(w, 5, R)


@pytest.mark.parametrize("n", [2, 8, 32])
def test_ntt_matches_dft(oracle_mod, n):
    rng = random.Random(n)
    xs = [rng.randrange(R) for _ in range(n)]
    w = pow(7, (R - 1) // n, R)
    data = b"".join(x.to_bytes(32, "big") for x in xs)
    fwd = oracle_mod.fr_ntt(data, n, False)
    want = [sum(xs[t] * pow(w, t * k, R) for t in range(n)) % R for k in range(n)]
    assert [int.from_bytes(fwd[32 * k:32 * k + 32], "big") for k in range(n)] == want
    back = oracle_mod.fr_ntt(fwd, n, True)
    assert back == data


# --------------------------------------------------------------- G1 basics

def test_g1_group_laws(oracle_mod, mainnet_setup_bytes):
    g = mainnet_setup_bytes[:48]  # [tau^0]G1 = generator
    assert oracle_mod.g1_mul_gen(1) == g
    ident = bytes([0xC0]) + bytes(47)
    assert oracle_mod.g1_mul_gen(R) == ident
    assert oracle_mod.g1_mul_gen(0) == ident
    a, b = 0x1234567, 0xABCDEF12345
    assert oracle_mod.g1_add(oracle_mod.g1_mul_gen(a), oracle_mod.g1_mul_gen(b)) == \
        oracle_mod.g1_mul_gen(a + b)
    assert oracle_mod.g1_add(oracle_mod.g1_mul_gen(a), oracle_mod.g1_mul_gen(R - a)) == ident
    for k in range(0, 4096, 511):
        pt = mainnet_setup_bytes[48 * k:48 * k + 48]
        rc, back = oracle_mod.g1_roundtrip(pt)
        assert rc == 0 and back == pt


# -------------------------------------------------- consensus vectors (golden)

def test_cells_match_vectors(vectors, oracle_mod):
    for name in vectors.valid_compute():
        rc, cells = oracle_mod.blob_to_cells(vectors.blob(name))
        assert rc == 0
        assert hashlib.sha256(cells).hexdigest() == vectors.compute[name]["cells_sha256"], name
        # systematic code: the first 64 cells are the blob itself
        assert cells[:131072] == vectors.blob(name)


def test_proofs_match_vectors(vectors, oracle_mainnet):
    for name in vectors.valid_compute():
        rc, proofs = oracle_mainnet.compute_proofs(vectors.blob(name), 1)
        assert rc == 0
        assert proofs == vectors.proofs(name), name


def test_invalid_blobs_rejected(vectors, oracle_mainnet):
    names = vectors.invalid_compute_131072()
    assert len(names) == 2
    for name in names:
        rc, _ = oracle_mainnet.compute_proofs(vectors.blob(name), 1)
        assert rc == 6, name


def test_recover_vectors(vectors, oracle_mod, cells_of):
    for case in vectors.recover_cases():
        src = cells_of(case["source"])
        cells = bytearray(262144)
        present = bytearray(128)
        for k in case["cell_indices"]:
            cells[2048 * k:2048 * k + 2048] = src[2048 * k:2048 * k + 2048]
            present[k] = 1
        rc, out = oracle_mod.recover_cells(bytes(cells), bytes(present), 1)
        if case["valid"]:
            assert rc == 0 and hashlib.sha256(out).hexdigest() == case["cells_sha256"], case["name"]
        else:
            assert rc == 1, case["name"]


def test_recover_random_patterns(vectors, oracle_mod, cells_of):
    rng = random.Random(3)
    for name in ["compute_cells_and_kzg_proofs_case_valid_4",
                 "compute_cells_and_kzg_proofs_case_valid_6"]:
        src = cells_of(name)
        for npresent in (64, 97):
            keep = set(rng.sample(range(128), npresent))
            cells = bytearray(rng.randbytes(262144))  # garbage in missing cells
            present = bytearray(128)
            for k in keep:
                cells[2048 * k:2048 * k + 2048] = src[2048 * k:2048 * k + 2048]
                present[k] = 1
            rc, out = oracle_mod.recover_cells(bytes(cells), bytes(present), 1)
            assert rc == 0 and out == src


def test_recover_reduces_noncanonical_present_cells(vectors, oracle_mod, cells_of):
    src = bytearray(cells_of("compute_cells_and_kzg_proofs_case_valid_3"))
    present = bytes([1] * 64 + [0] * 64)
    # element 5 of cell 2 as value + r (still < 2^256) must behave as value
    off = 2 * 2048 + 5 * 32
    v = int.from_bytes(src[off:off + 32], "big")
    assert v + R < (1 << 256)
    bumped = bytearray(src)
    bumped[off:off + 32] = (v + R).to_bytes(32, "big")
    rc1, out1 = oracle_mod.recover_cells(bytes(src), present, 1)
    rc2, out2 = oracle_mod.recover_cells(bytes(bumped), present, 1)
    assert rc1 == rc2 == 0 and out1 == out2 == bytes(src)


# ----------------------------------------------------- closed forms / identities

def test_fk20_matches_closed_form_with_known_tau(oracle_mod, seeded_setup_bytes):
    """proof_c = [q_c(tau)]G1, q_c = p div (X^64 - w^(64 brev7(c))): pins the
    whole FK20 data flow (x_ext_fft layout, circulant, phase B, brp7)."""
    S = oracle_mod.Setup(seeded_setup_bytes)
    rng = random.Random(4)
    single = bytearray(131072)
    single[32 * 1000:32 * 1001] = rng.randrange(R).to_bytes(32, "big")
    for blob in (random_blob(rng), bytes(single)):
        rc, fk20 = S.compute_proofs(blob, 1)
        rc2, closed = oracle_mod.proofs_from_tau(blob, SEEDED_TAU)
        assert rc == rc2 == 0
        assert fk20 == closed


def test_fk20_matches_direct_quotient_sums(oracle_mod, oracle_mainnet, vectors):
    """h_m = sum_k c_{k+64m} s_k by plain MSMs, then FFT_128 -- no circulant."""
    rc, direct = oracle_mainnet.compute_proofs_direct(
        vectors.blob("compute_cells_and_kzg_proofs_case_valid_2"))
    assert rc == 0
    assert direct == vectors.proofs("compute_cells_and_kzg_proofs_case_valid_2")


def test_proofs_are_linear(oracle_mod, oracle_mainnet):
    rng = random.Random(5)
    a = [rng.randrange(R) for _ in range(4096)]
    b = [rng.randrange(R) for _ in range(4096)]
    enc = lambda xs: b"".join(x.to_bytes(32, "big") for x in xs)
    _, pa = oracle_mainnet.compute_proofs(enc(a), 1)
    _, pb = oracle_mainnet.compute_proofs(enc(b), 1)
    _, pab = oracle_mainnet.compute_proofs(enc([(x + y) % R for x, y in zip(a, b)]), 1)
    for c in (0, 77, 127):
        s = slice(48 * c, 48 * c + 48)
        assert oracle_mod.g1_add(pa[s], pb[s]) == pab[s]


def test_split_circulant_equals_ifft_trunc_fft():
    """The GPU's phase B (cyclic + negacyclic halves with K+- = (K[e] +- K[e+64])/2)
    equals FFT_128(trunc_64(IFFT_128(u))) as a linear map -- checked on scalars."""
    rng = random.Random(6)
    w = pow(7, (R - 1) // 128, R)
    inv128 = pow(128, -1, R)
    u = [rng.randrange(R) for _ in range(128)]
    h = [sum(u[j] * pow(w, -j * t, R) for j in range(128)) * inv128 % R for t in range(128)]
    h = h[:64] + [0] * 64
    want = [sum(h[t] * pow(w, t * k, R) for t in range(128)) % R for k in range(128)]

    def K(e):
        if e == 0:
            return pow(2, -1, R)
        if e % 2 == 0:
            return 0
        return pow(64 * (1 - pow(w, e, R)), -1, R)

    half = pow(2, -1, R)
    kp = [(K(e) + K(e + 64)) * half % R for e in range(64)]
    km = [(K(e) - K(e + 64)) * half % R for e in range(64)]
    assert sum(1 for x in kp if x) == 33 and sum(1 for x in km if x) == 33
    up = [(u[j] + u[j + 64]) % R for j in range(64)]
    um = [(u[j] - u[j + 64]) % R for j in range(64)]
    out = [0] * 128
    for a in range(64):
        cp = sum(kp[e] * up[(a - e) % 64] for e in range(64)) % R
        cm = sum((-1 if e > a else 1) * km[e] * um[(a - e) % 64] for e in range(64)) % R
        out[a] = (cp + cm) % R
        out[a + 64] = (cp - cm) % R
    assert out == want


def test_vanishing_values_closed_form():
    """Z(X) = z(X^64): its 8192-point evaluations on the domain and on the 7-coset
    take the 128 values z(w128^k) and z(7^64 w128^k) -- the identity the GPU
    recovery uses instead of two NTTs of Z."""
    rng = random.Random(7)
    missing = rng.sample(range(128), 40)
    w8192 = pow(7, (R - 1) // 8192, R)
    roots = [pow(w8192, 64 * brev(c, 7), R) for c in missing]

    def z(y):
        out = 1
        for rt in roots:
            out = out * (y - rt) % R
        return out

    for k in rng.sample(range(8192), 12):
        x = pow(w8192, k, R)
        assert z(pow(x, 64, R)) == z(pow(w8192, 64 * (k % 128), R))
        xc = 7 * x % R
        assert z(pow(xc, 64, R)) == z(pow(7, 64, R) * pow(w8192, 64 * (k % 128), R) % R)
    # zero exactly on the missing cells' cosets: wire cell c holds indices brev13(64c+t)
    for c in range(128):
        k = brev(64 * c, 13)
        is_zero = z(pow(pow(w8192, k, R), 64, R)) == 0
        assert is_zero == (c in missing)


def test_a_flipped_blob_byte_is_detected(vectors, oracle_mainnet):
    """The vector check discriminates: one changed element changes the proofs."""
    blob = bytearray(vectors.blob("compute_cells_and_kzg_proofs_case_valid_2"))
    blob[32 * 17 + 31] ^= 1
    rc, proofs = oracle_mainnet.compute_proofs(bytes(blob), 1)
    assert rc == 0 and proofs != vectors.proofs("compute_cells_and_kzg_proofs_case_valid_2")
