"""Phase A's window plans and signed scalars, restated on the CPU (no GPU):
k_fk20_scalars writes each scalar as sign and magnitude (fr_signed_magnitude),
k_phase_a recodes the magnitude in signed digits over the plan's windows (the
top k one bit wider than c) and flips every digit of a negative scalar, and
k_table_multiples lays out each base's multiples by slots.  The digits must
rebuild the scalar mod r with no carry out of the top window, every table
entry must be written exactly once, and (14, 3) must be 18 windows in
168 GiB.  A wrong offset, width or sign fails here; the GPU tests then check
the proofs bit-exactly at every plan the prover picks."""
import random

R = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001


class Plan:
    def __init__(self, c, k):
        self.c = c
        self.W = (255 - k + c - 1) // c  # window_plan
        self.k = min(k, self.W)
        self.H = 1 << (c - 1)

    def wide(self, w):
        return 1 if w >= self.W - self.k else 0

    def below(self, w):
        return w - (self.W - self.k) if w > self.W - self.k else 0

    def bit(self, w):
        return w * self.c + self.below(w)

    def entry(self, w):
        return (w + self.below(w)) * self.H

    def entries(self):
        return (self.W + self.k) * self.H


def signed_magnitude(a):
    """fr_signed_magnitude: (r - a) | 2^255 when r - a < a, else a."""
    n = (R - a) % (1 << 256)
    return (n | 1 << 255) if n < a else a


def digits(s, plan):
    """k_phase_a's recoding: [(w, digit)] with the scalar's sign applied."""
    neg = s >> 255
    m = s & ((1 << 255) - 1)
    out, carry = [], 0
    for w in range(plan.W):
        width = plan.c + plan.wide(w)
        hw = 1 << (width - 1)
        raw = ((m >> plan.bit(w)) & ((1 << width) - 1)) + carry
        carry = 1 if raw > hw else 0
        mag = 2 * hw - raw if carry else raw
        assert mag <= hw
        d = -mag if carry else mag
        out.append((w, -d if neg else d))
    assert carry == 0, "carry out of the top window"
    return out


def test_plans_cover_255_bits():
    p = Plan(14, 3)
    assert (p.W, p.k) == (18, 3)
    assert [p.c + p.wide(w) for w in range(p.W)] == [14] * 15 + [15] * 3
    assert p.bit(p.W - 1) + p.c + p.wide(p.W - 1) == 255
    assert 8192 * p.entries() * 128 == 168 << 30
    assert (Plan(14, 0).W, 8192 * Plan(14, 0).entries() * 128) == (19, 152 << 30)
    assert [Plan(c, 0).W for c in (8, 10, 12, 13)] == [32, 26, 22, 20]
    for c, k in ((14, 3), (14, 0), (13, 0), (12, 0), (10, 0), (8, 0), (13, 8), (12, 15)):
        p = Plan(c, k)
        assert p.W * c + p.k >= 255
        # windows tile the bits and the entries without gaps
        for w in range(1, p.W):
            assert p.bit(w) == p.bit(w - 1) + c + p.wide(w - 1)
            assert p.entry(w) == p.entry(w - 1) + (p.H << p.wide(w - 1))
        assert p.entry(p.W - 1) + (p.H << p.wide(p.W - 1)) == p.entries()


def test_chooser_plans():
    """choose_window's candidates, fewest windows first: c = floor(255 / W)
    with the top k = 255 - W c windows widened; each is a valid plan of W
    windows, and the table shrinks as W grows."""
    prev = None
    for W in range(18, 37):
        c, k = 255 // W, 255 - W * (255 // W)
        p = Plan(c, k)
        assert (p.W, p.k) == (W, k) and p.W * c + p.k == 255
        size = 8192 * p.entries() * 128
        assert prev is None or size < prev
        prev = size
    assert 8192 * Plan(13, 8).entries() * 128 == 108 << 30
    assert 8192 * Plan(12, 3).entries() * 128 == 48 << 30


def test_signed_digits_rebuild_the_scalar():
    rng = random.Random(11)
    edge = [0, 1, 2, (R - 1) // 2, (R + 1) // 2, R - 2, R - 1, 1 << 253, (1 << 254) - 1]
    for plan in (Plan(14, 3), Plan(14, 0), Plan(12, 0), Plan(10, 0), Plan(13, 8), Plan(12, 3), Plan(11, 2),
                 Plan(9, 21), Plan(7, 31)):
        for a in edge + [rng.randrange(R) for _ in range(400)]:
            s = signed_magnitude(a)
            assert (s & ((1 << 255) - 1)) < 1 << 254
            got = sum(d << plan.bit(w) for w, d in digits(s, plan))
            assert got % R == a


def test_table_slots_write_every_entry_once():
    """k_table_multiples' thread -> (base, window, multiples) map, on a few
    bases, for a plain and a widened plan with small c."""
    for plan in (Plan(7, 0), Plan(7, 3), Plan(8, 2)):
        S = min(plan.H, 64)
        nseg = plan.H // S
        slots = plan.W + plan.k
        narrow = plan.W - plan.k
        nbases = 3
        seen = {}
        for g in range(nbases * slots * nseg):
            seg = g % nseg
            bv = g // nseg
            v, b = bv % slots, bv // slots
            w = v if v < narrow else narrow + (v - narrow) // 2
            half = 0 if v < narrow else (v - narrow) % 2
            d0 = (half * nseg + seg) * S + 1
            for m in range(S):
                idx = bv * plan.H + seg * S + m
                assert idx == b * plan.entries() + plan.entry(w) + d0 - 1 + m
                assert idx not in seen
                seen[idx] = (b, w, d0 + m)
        assert len(seen) == nbases * plan.entries()
        for (b, w, d) in seen.values():
            assert 1 <= d <= plan.H << plan.wide(w)
