"""Device Montgomery arithmetic vs Python big integers (gfx950, via the
vkzg_amd_debug_* hooks).  Bit-exact."""
import random

import pytest

import vkzg_amd

pytestmark = pytest.mark.gpu

R = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
P = 0x1A0111EA397FE69A4B1BA7B6434BACD764774B84F38512BF6730D2A0F6B0F6241EABFFFEB153FFFFB9FEFFFFFFFFAAAB


def limbs(x, n):
    return tuple((x >> (32 * i)) & 0xFFFFFFFF for i in range(n))


def value(t):
    return sum(v << (32 * i) for i, v in enumerate(t))


def test_fp_mul_random_and_edges():
    rng = random.Random(11)
    edge = [0, 1, 2, P - 1, P - 2, (P - 1) // 2, (P + 1) // 2, 1 << 380, (1 << 32) - 1]
    a = edge + [rng.randrange(P) for _ in range(4000)]
    b = [rng.choice(edge) for _ in edge] + [rng.randrange(P) for _ in range(4000)]
    got = vkzg_amd.debug_fp_mul([limbs(x, 12) for x in a], [limbs(y, 12) for y in b])
    for x, y, g in zip(a, b, got):
        assert value(g) == x * y % P, (hex(x), hex(y))


def test_fr_mul_random_edges_and_unreduced_inputs():
    rng = random.Random(12)
    edge = [0, 1, 2, R - 1, R - 2, R, R + 1, 2 * R, (1 << 256) - 1, (1 << 255)]
    a = edge + [rng.randrange(1 << 256) for _ in range(4000)]
    b = [rng.choice(edge) for _ in edge] + [rng.randrange(R) for _ in range(4000)]
    got = vkzg_amd.debug_fr_mul([limbs(x, 8) for x in a], [limbs(y, 8) for y in b])
    for x, y, g in zip(a, b, got):
        assert value(g) == (x % R) * (y % R) % R, (hex(x), hex(y))


def test_fp_inverse_binary_gcd():
    """The device binary-GCD inverse against Python (and, inside the kernel,
    against the Fermat power): random values, 0, 1, p - 1, powers of two."""
    rng = random.Random(13)
    edge = [0, 1, 2, 3, P - 1, P - 2, (P - 1) // 2, (P + 1) // 2] + [1 << k for k in range(0, 381, 7)]
    a = edge + [rng.randrange(P) for _ in range(4000)]
    got = vkzg_amd.debug_fp_inv([limbs(x, 12) for x in a])
    for x, g in zip(a, got):
        assert value(g) == (pow(x, -1, P) if x else 0), hex(x)


def test_fp_mul_rate_hook(gpu_prover):
    """The issue-rate hook runs and reports a plausible rate: more waves per
    SIMD never lower the chip's product rate much, and one wave alone needs at
    least the ~590 VALU instructions of a product times the 4-cycle issue."""
    r1, cyc = gpu_prover.debug_fp_mul_rate(1, iters=200)
    r8, _ = gpu_prover.debug_fp_mul_rate(8, iters=200)
    assert 1e9 < r1 <= r8 * 1.2
    assert 2000 < cyc < 100000
