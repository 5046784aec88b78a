"""GPU test of vkzg_amd_write_v1_tables: the reference's tables file, version 1
(src/precomputed_tables.cpp:16-58), written for the mainnet setup, parsed and
checked section by section against independent computations:

  * header: magic 'VKZG', version 1, item counts;
  * position table: entries (j, i, d) converted from the reference's
    Montgomery form (x 2^384 mod p) and compressed equal the oracle's
    [2^(8d)] x_ext_fft[j][i]; (0, 0) nowhere, since no x_ext_fft multiple is
    the identity;
  * roots: Montgomery omega_8192^(+-k) (x 2^256 mod r) from Python integers;
  * items: per bucket |digit| - 1 and per (e, d, sign) the digits rebuild
    K+-[e] = (K[e] +- K[e+64]) / 2 with K[0] = 1/2, K[odd e] =
    1/(64 (1 - omega_128^e)), K[even e] = 0 (the split circulant of
    k_phase_b_split.comp), each digit in [-127, 128] and obeying the
    reference's byte-plus-carry recoding (k_phase_a_sort.comp:31-40); the
    counts are 1050 and 1046;
  * offsets, permutations and Montgomery 1/4096.
"""
import random
import struct

import pytest

pytestmark = pytest.mark.gpu

R = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
P = 0x1A0111EA397FE69A4B1BA7B6434BACD764774B84F38512BF6730D2A0F6B0F6241EABFFFEB153FFFFB9FEFFFFFFFFAAAB


def _taps():
    w128 = pow(pow(7, (R - 1) // 8192, R), 64, R)
    inv = lambda a: pow(a, R - 2, R)

    def tap(e):
        if e == 0:
            return inv(2)
        if e % 2 == 0:
            return 0
        return inv(64 * (1 - pow(w128, e, R)) % R)
    K = [tap(e) for e in range(128)]
    return ([(K[e] + K[e + 64]) * inv(2) % R for e in range(64)],
            [(K[e] - K[e + 64]) * inv(2) % R for e in range(64)])


def _compress(x, y):
    b = bytearray(x.to_bytes(48, "big"))
    b[0] |= 0x80 | (0x20 if y > (P - 1) // 2 else 0)
    return bytes(b)


def test_v1_tables_file(gpu_prover, oracle_mainnet, tmp_path):
    path = tmp_path / "precomputed_tables.bin"
    gpu_prover.write_v1_tables(path)
    data = path.read_bytes()
    magic, version, n_plus, n_minus = struct.unpack_from("<4I", data, 0)
    assert magic == 0x475A4B56 and version == 1
    assert (n_plus, n_minus) == (1050, 1046)
    pos_words = 128 * 64 * 32 * 24
    off = 16
    pos = memoryview(data)[off:off + 4 * pos_words]
    off += 4 * pos_words
    roots = data[off:off + 4 * 65536]
    off += 4 * 65536
    roots_inv = data[off:off + 4 * 65536]
    off += 4 * 65536
    items = [struct.unpack_from("<%dI" % n_plus, data, off)]
    off += 4 * n_plus
    items.append(struct.unpack_from("<%dI" % n_minus, data, off))
    off += 4 * n_minus
    offsets = [struct.unpack_from("<129I", data, off), struct.unpack_from("<129I", data, off + 516)]
    off += 2 * 516
    perms = [struct.unpack_from("<128I", data, off), struct.unpack_from("<128I", data, off + 512)]
    off += 2 * 512
    inv_blob = data[off:off + 32]
    off += 32
    assert off == len(data)

    # position table against the oracle's x_ext_fft
    rinv384 = pow(2, -384, P)
    rng = random.Random(3)
    picks = [(0, 0, 0), (127, 63, 31), (5, 17, 1), (64, 0, 16)] + [
        (rng.randrange(128), rng.randrange(64), rng.randrange(32)) for _ in range(12)]
    for j, i, d in picks:
        w = (((j * 64 + i) * 32 + d) * 24) * 4
        x = int.from_bytes(pos[w:w + 48], "little") * rinv384 % P
        y = int.from_bytes(pos[w + 48:w + 96], "little") * rinv384 % P
        assert (x, y) != (0, 0)
        assert (y * y - x ** 3 - 4) % P == 0
        want = oracle_mainnet.xext(j, i)
        if d:
            from oracle import oracle as O
            want = O.g1_mul(want, 1 << (8 * d))
        assert _compress(x, y) == want, (j, i, d)

    # roots
    w = pow(7, (R - 1) // 8192, R)
    for k in (0, 1, 2, 4095, 4096, 8191):
        assert int.from_bytes(roots[32 * k:32 * k + 32], "little") == pow(w, k, R) * (1 << 256) % R
        assert int.from_bytes(roots_inv[32 * k:32 * k + 32], "little") == pow(w, -k, R) * (1 << 256) % R
    assert int.from_bytes(inv_blob, "little") == pow(4096, -1, R) * (1 << 256) % R

    # phase-B items rebuild the split-circulant taps
    for half, taps in enumerate(_taps()):
        rebuilt = [0] * 64
        digits = [[0] * 32 for _ in range(64)]
        off_h, items_h = offsets[half], items[half]
        assert off_h[0] == 0 and off_h[128] == len(items_h)
        assert sorted(perms[half]) == list(range(128))
        loads = [off_h[k + 1] - off_h[k] for k in range(128)]
        assert [loads[k] for k in perms[half]] == sorted(loads, reverse=True)
        for k in range(128):
            for it in items_h[off_h[k]:off_h[k + 1]]:
                e, d, sign = it & 0xFF, (it >> 8) & 0xFF, (it >> 16) & 1
                assert e < 64 and d < 32 and it >> 17 == 0
                dig = -(k + 1) if sign else k + 1
                assert -127 <= dig <= 128 and digits[e][d] == 0
                digits[e][d] = dig
                rebuilt[e] = (rebuilt[e] + dig * (1 << (8 * d))) % R
        assert rebuilt == taps
        for e in range(64):  # byte-plus-carry recoding
            carry = 0
            for d in range(32):
                v = ((taps[e] >> (8 * d)) & 0xFF) + carry
                carry = 1 if v > 128 else 0
                assert digits[e][d] == (v - 256 if v > 128 else v)
