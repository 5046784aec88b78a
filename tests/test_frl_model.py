"""The 9 x 29-bit Fr form of the radix-4 recovery kernels (field_frl.h),
restated limb by limb on the CPU: its constants against their definitions,
the column and value bounds its header states (worst-case operands, not only
random ones), the Montgomery product mod r, pack/unpack, and the value growth
of the radix-4 butterfly (recover.hip dft4) over the stages of a pass.  A
wrong offset limb, a constant in the wrong Montgomery radix or a bound that
does not hold fails here; the GPU tests check the recovered cells
bit-exactly."""
import os
import random
import re

R = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
M = (1 << 29) - 1
RL = 1 << 261
HEADER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "vkzg_amd", "csrc", "field_frl.h")


def limbs(x):
    return [(x >> (29 * i)) & M for i in range(8)] + [x >> 232]


def val(l):
    return sum(v << (29 * i) for i, v in enumerate(l))


def header_consts():
    text = open(HEADER).read()
    out = {}
    for name, body in re.findall(r"constexpr uint32_t (\w+)\[\d+\] = \{([^}]*)\};", text):
        out[name] = [int(x.strip().rstrip("u"), 16) for x in body.split(",")]
    return out


C = header_consts()
P = C["P"]


def mul(a, b):
    """frl_mul, with its 64-bit accumulator checked."""
    m = [0] * 9
    acc = 0
    for i in range(9):
        for j in range(i):
            acc += a[j] * b[i - j] + m[j] * P[i - j]
        acc += a[i] * b[0]
        m[i] = (-acc) & M
        acc += m[i]
        assert acc < 1 << 64 and acc & M == 0
        acc >>= 29
    r = [0] * 9
    for i in range(9, 18):
        for j in range(i - 8, 9):
            acc += a[j] * b[i - j] + m[j] * P[i - j]
        assert acc < 1 << 64
        r[i - 9] = acc & M
        acc >>= 29
    assert acc == 0
    assert val(r) % R == val(a) * val(b) * pow(RL, -1, R) % R
    return r


def add(a, b):
    return [x + y for x, y in zip(a, b)]


def sub(a, b, off):
    assert all(y <= o for y, o in zip(b, off))
    return [x + o - y for x, o, y in zip(a, off, b)]


def norm(a):
    a = list(a)
    for i in range(8):
        a[i + 1] += a[i] >> 29
        a[i] &= M
    return a


def test_constants_match_definitions():
    K = (pow(7, 8192, R) - 1) % R
    assert val(P) == R and P[0] == 1
    assert val(C["OFF4"]) == 4 * R and all(1 << 29 <= x < 1 << 30 for x in C["OFF4"][:8])
    assert val(C["OFF8"]) == 8 * R and all(1 << 30 <= x < 1 << 31 for x in C["OFF8"][:8])
    # the offsets' top limbs cover the subtrahends' (below 2r and 4r)
    assert C["OFF4"][8] >= (2 * R - 1) >> 232 and C["OFF8"][8] >= 2 * ((2 * R - 1) >> 232)
    assert C["SEVEN_POW64_M"] == limbs(pow(7, 64, R) * RL % R)
    assert C["INV_SEVEN_POW8192_MINUS_1_M"] == limbs(pow(K, -1, R) * RL % R)
    assert C["SIXTY_FOUR_INV_SEVEN_POW8192_MINUS_1_M"] == limbs(64 * pow(K, -1, R) * RL % R)
    assert C["ONE_M"] == limbs(RL % R) and C["ONE_PLAIN"] == limbs(1)
    assert C["R2"] == limbs(RL * RL % R)
    assert C["FROM_R256"] == limbs(pow(2, 266, R))
    t = pow(2, 261, R)
    assert C["TO_R261"] == [(t >> (32 * i)) & 0xFFFFFFFF for i in range(8)]


def test_product_bounds_worst_case():
    """First operand limbs just below 3 * 2^30, second operand normalised:
    no column overflows, and A B < 70 keeps the result below 2r."""
    top = lambda bound: (bound * R) >> 232
    b_max = [M] * 8 + [top(1)]  # canonical second operand, limbs at their maximum
    for A in (2, 6, 36):
        a = [(3 << 30) - 1] * 8 + [min(top(A), (3 << 30) - 1)]
        while val(a) >= A * R:
            a[8] -= 1
        r = mul(a, b_max)
        assert val(r) < 2 * R and all(x <= M for x in r)
    # the 1/Z product: first operand below 32r, second below 2r
    a = [(3 << 30) - 1] * 8 + [top(32)]
    while val(a) >= 32 * R:
        a[8] -= 1
    b = limbs(2 * R - 1)
    assert val(mul(a, b)) < 2 * R


def test_product_random_and_pack_roundtrip():
    rng = random.Random(5)
    for _ in range(3000):
        a = [rng.randrange(3 << 30) for _ in range(8)] + [rng.randrange((36 * R) >> 232)]
        b = limbs(rng.randrange(2 * R))
        r = mul(a, b)
        assert val(r) < 2 * R
        x = val(r)
        w = [(x >> (32 * i)) & 0xFFFFFFFF for i in range(8)]
        packed = sum(v << (32 * i) for i, v in enumerate(w))
        assert limbs(packed) == r


def test_butterfly_bounds_over_a_pass():
    """dft4 as recover.hip writes it, on worst-case operands: v0 carried
    (normalised below V r), v1..v3 products (normalised below 2r).  Outputs
    stay below (V + 10) r with limbs below 3 * 2^30, so three stages end
    below 32r and every product's first operand stays below 36r."""
    rng = random.Random(6)
    off4, off8 = C["OFF4"], C["OFF8"]
    w4 = limbs(rng.randrange(R))

    def dft4(v0, v1, v2, v3):
        a, b = add(v0, v2), sub(v0, v2, off4)
        c, d = add(v1, v3), sub(v1, v3, off4)
        tw = mul(d, w4)
        return add(a, c), add(b, tw), sub(a, c, off8), sub(b, tw, off4)

    def near(bound):  # normalised, just below bound * r
        return limbs(bound * R - 1 - rng.randrange(1 << 200))

    V = 2
    for stage in range(3):
        v0 = near(V) if stage == 0 else norm(near(V))
        outs = dft4(v0, near(2), near(2), near(2))
        for o in outs:
            assert all(x < 3 << 30 for x in o)
            assert val(o) < (V + 10) * R
        V += 10
    assert V == 32
    # the radix-2 stage of pass A: u0 normalised below 32r, t a product
    u0, t = near(32), near(2)
    assert val(add(u0, t)) < 34 * R and val(sub(u0, t, off4)) < 36 * R
