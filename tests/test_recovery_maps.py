"""The recovery kernels' algorithm, restated on the CPU and checked against
the oracle (no GPU): the five fused kernels of recover.hip with their lane,
register and LDS index maps, for both register radices, on plain integers
mod r.  A wrong twiddle exponent, transposed index or misplaced pointwise
factor in this model fails here; the GPU tests then check the kernels
against the oracle bit-exactly."""
import random

import pytest

from conftest import random_blob

R = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
W = pow(7, (R - 1) // 8192, R)
ROOTS = [pow(W, k, R) for k in range(8192)]
ROOTS_INV = [pow(W, (8192 - k) % 8192, R) for k in range(8192)]
INV8192 = pow(8192, R - 2, R)
SHIFT_FWD = [pow(7, k, R) * INV8192 % R for k in range(8192)]
SHIFT_INV = [pow(pow(7, R - 2, R), k, R) * INV8192 % R for k in range(8192)]
SEVEN64 = pow(7, 64, R)
KINV = pow((pow(7, 8192, R) - 1) % R, R - 2, R)
CELLS = 262144


def brev(x, bits):
    return int(format(x, "0%db" % bits)[::-1], 2)


def dft(v, tab):
    """out[e] = sum_b v[b] w_N^(b e), radix-2 DIT on a register array (as dft8/dft4)."""
    n = len(v)
    lg = n.bit_length() - 1
    x = [v[brev(i, lg)] for i in range(n)]
    ln = 2
    while ln <= n:
        for st in range(0, n, ln):
            for j in range(ln // 2):
                t = x[st + j + ln // 2] * tab[(8192 // ln) * j] % R
                x[st + j], x[st + j + ln // 2] = (x[st + j] + t) % R, (x[st + j] - t) % R
        ln *= 2
    return x


def vanish(present, xs, want):
    out = []
    for x in xs:
        p = 1
        for c in range(128):
            if bool(present[c]) == want:
                p = p * (x - ROOTS[64 * brev(c, 7)]) % R
        out.append(p)
    return out


# ------------------------------------------------------------------ radix 8

def r8_pass_b(v, tab):
    """v[L][b] = y[a + 8b][cc], L = 8a + cc -> out[L][f] = X[e + 8f], L = 8e + cc."""
    lds = {}
    for L in range(64):
        cc, a = L & 7, L >> 3
        z = dft(v[L], tab)
        for e in range(8):
            lds[cc * 64 + e * 8 + a] = z[e] * (tab[128 * a * e % 8192] if e else 1) % R
    return [dft([lds[(L & 7) * 64 + (L >> 3) * 8 + a] for a in range(8)], tab) for L in range(64)]


def r8_k2(L, r):
    m = L & 15
    return (m >> 1) + 8 * (2 * (m & 1) + (r >> 2)) + 32 * (r & 3)


def r8_pass_a(v, tab, n1_of):
    """v[L][b] = x[a + 16b] of column c, L = 16c + a -> out[L][r] = y[n1][r8_k2(L, r)]."""
    lds = {}
    for L in range(64):
        c, a = L >> 4, L & 15
        z = dft(v[L], tab)
        for e in range(8):
            lds[c * 128 + a * 8 + e] = z[e] * (tab[64 * a * e % 8192] if e else 1) % R
    lds2 = {}
    for L in range(64):
        c, m = L >> 4, L & 15
        e, odd = m >> 1, m & 1
        for j in range(2):
            a0 = 2 * odd + j
            g4 = dft([lds[c * 128 + (a0 + 4 * a1) * 8 + e] for a1 in range(4)], tab)
            for g in range(4):
                lds2[c * 128 + e * 16 + a0 * 4 + g] = g4[g] * (tab[512 * a0 * g] if g else 1) % R
    out = []
    for L in range(64):
        c, m = L >> 4, L & 15
        e, odd = m >> 1, m & 1
        res = []
        for t in range(2):
            g = 2 * odd + t
            res += dft([lds2[c * 128 + e * 16 + a0 * 4 + g] for a0 in range(4)], tab)
        out.append([res[r] * tab[n1_of(L) * r8_k2(L, r) % 8192] % R for r in range(8)])
    return out


def r8_s_index(n1, k2):
    return ((k2 & 63) >> 2) * 512 + n1 * 8 + (k2 & 3) + 4 * (k2 >> 6)


# ------------------------------------------------------------------ radix 4

def stage4(lds, col, N, NS, j, tab):
    v = [lds[col + j + r * (N // 4)] for r in range(4)]
    if NS > 1:
        v = [v[r] * tab[(8192 // (NS * 4)) * (j % NS) * r % 8192] % R for r in range(4)]
    return dft(v, tab)


def store4(lds, col, NS, j, v):
    base = (j // NS) * NS * 4 + j % NS
    for r in range(4):
        lds[col + base + r * NS] = v[r]


def r4_pass_b(v, tab):
    """v[L][r] = y[j + 16r][cc], L = 4j + cc -> out[L][r] = X[j + 16r] of column cc."""
    lds = {}
    for L in range(64):
        store4(lds, (L & 3) * 64, 1, L >> 2, dft(v[L], tab))
    lds2 = {}
    for L in range(64):
        store4(lds2, (L & 3) * 64, 4, L >> 2, stage4(lds, (L & 3) * 64, 64, 4, L >> 2, tab))
    return [stage4(lds2, (L & 3) * 64, 64, 16, L >> 2, tab) for L in range(64)]


def r4_k2(j, t):
    return j + 32 * (t >> 1) + 64 * (t & 1)


def r4_pass_a(v, tab, n1_of):
    """v[L][r] = x[j + 32r] of column c, L = 32c + j -> out[L][t] = y[n1][r4_k2(j, t)]."""
    lds = {}
    for L in range(64):
        store4(lds, (L >> 5) * 128, 1, L & 31, dft(v[L], tab))
    for ns, nxt in ((4, 4), (16, 16)):
        new = {}
        for L in range(64):
            store4(new, (L >> 5) * 128, nxt, L & 31, stage4(lds, (L >> 5) * 128, 128, ns, L & 31, tab))
        lds = new
    out = []
    for L in range(64):
        col, j = (L >> 5) * 128, L & 31
        u = [lds[col + j], lds[col + j + 64], lds[col + j + 32], lds[col + j + 96]]
        t0, t1 = u[1] * tab[64 * j] % R, u[3] * tab[64 * (j + 32)] % R
        y = [(u[0] + t0) % R, (u[0] - t0) % R, (u[2] + t1) % R, (u[2] - t1) % R]
        out.append([y[t] * tab[n1_of(L) * r4_k2(j, t) % 8192] % R for t in range(4)])
    return out


def r4_s_index(n1, k2):
    return ((k2 & 63) >> 1) * 256 + n1 * 4 + (k2 & 1) + 2 * (k2 >> 6)


# ------------------------------------------------------------------ the five kernels

def recover_model(cells, present, radix):
    E = [int.from_bytes(cells[32 * w:32 * w + 32], "big") % R for w in range(8192)]
    per = 16 if radix == 8 else 32  # items per blob
    cw = 4 if radix == 8 else 2      # pass-A columns per item
    nv = 8 if radix == 8 else 4      # values per lane
    pass_a = r8_pass_a if radix == 8 else r4_pass_a
    pass_b = r8_pass_b if radix == 8 else r4_pass_b
    s_index = r8_s_index if radix == 8 else r4_s_index
    k2_of = r8_k2 if radix == 8 else (lambda L, t: r4_k2(L & 31, t))
    lanes_a = 16 if radix == 8 else 32
    S = [None] * 8192
    for q in range(per):
        c0 = cw * q
        cols = [c0 + v % cw + 64 * (v // cw) for v in range(2 * cw)]
        zd = vanish(present, [ROOTS[64 * k] for k in cols], False)
        v = []
        for L in range(64):
            c, a = L // lanes_a, L % lanes_a
            n1 = c0 + c
            row = []
            for b in range(nv):
                n2 = a + lanes_a * b
                row.append(E[brev(n1 + 64 * n2, 13)] * zd[c + cw * (n2 & 1)] % R)
            v.append(row)
        out = pass_a(v, ROOTS_INV, lambda L, c0=c0: c0 + L // lanes_a)
        for L in range(64):
            for t in range(nv):
                S[s_index(c0 + L // lanes_a, k2_of(L, t))] = out[L][t]
    for stage in range(3):
        tb = ROOTS_INV if stage != 1 else ROOTS
        ta = ROOTS if stage != 1 else ROOTS_INV
        S2 = [None] * 8192
        for q in range(per):
            c0 = cw * q
            ncb = 2 * cw  # pass-B columns
            X = pass_b([[S[q * 64 * ncb + 64 * b + L] for b in range(nv)] for L in range(64)], tb)
            if stage == 1:
                ks = [c0 + v % cw + 64 * (v // cw) for v in range(ncb)]
                izc = [p * KINV % R for p in vanish(present, [SEVEN64 * ROOTS[64 * k] % R for k in ks], True)]
            lds = {}
            for L in range(64):
                cc = L % ncb
                jb = L // ncb
                k2 = c0 + cc % cw + 64 * (cc // cw)
                for f in range(nv):
                    k1 = jb + (64 // nv) * f
                    k = k2 + 128 * k1
                    fac = SHIFT_FWD[k] if stage == 0 else izc[cc] if stage == 1 else SHIFT_INV[k]
                    lds[(cc % cw) * 128 + 2 * k1 + cc // cw] = X[L][f] * fac % R
            v = [[lds[(L // lanes_a) * 128 + L % lanes_a + lanes_a * b] for b in range(nv)] for L in range(64)]
            out = pass_a(v, ta, lambda L, c0=c0: c0 + L // lanes_a)
            for L in range(64):
                for t in range(nv):
                    S2[s_index(c0 + L // lanes_a, k2_of(L, t))] = out[L][t]
        S = S2
    res = bytearray(CELLS)
    for q in range(per):
        c0 = cw * q
        ncb = 2 * cw
        X = pass_b([[S[q * 64 * ncb + 64 * b + L] for b in range(nv)] for L in range(64)], ROOTS)
        for L in range(64):
            cc, jb = L % ncb, L // ncb
            k2 = c0 + cc % cw + 64 * (cc // cw)
            for f in range(nv):
                w = brev(k2 + 128 * (jb + (64 // nv) * f), 13)
                res[32 * w:32 * w + 32] = X[L][f].to_bytes(32, "big")
    return bytes(res)


@pytest.mark.parametrize("radix", [4, 8])
@pytest.mark.parametrize("consistent", [True, False])
def test_recovery_maps_match_oracle(oracle_mod, radix, consistent):
    rng = random.Random(radix * 10 + consistent)
    if consistent:
        rc, cells = oracle_mod.blob_to_cells(random_blob(rng))
        assert rc == 0
        keep = set(rng.sample(range(128), 64))
    else:
        cells = rng.randbytes(CELLS)  # not a codeword; some elements >= r
        keep = set(rng.sample(range(128), 100))
    present = bytes(1 if c in keep else 0 for c in range(128))
    masked = bytearray(rng.randbytes(CELLS))
    for c in keep:
        masked[2048 * c:2048 * (c + 1)] = cells[2048 * c:2048 * (c + 1)]
    rc, want = oracle_mod.recover_cells(bytes(masked), present, 1)
    assert rc == 0
    got = recover_model(bytes(masked), present, radix)
    assert got == want
    if consistent:
        assert got == cells


# ------------------------------------------------------------------ the three kernels

def r4_pass_af(v, tab, fac):
    """r4_pass_a with output (L, t) multiplied by fac(L, t) instead of the twiddle."""
    out = r4_pass_a(v, tab, lambda L: 0)
    return [[out[L][t] * fac(L, t) % R for t in range(4)] for L in range(64)]


def m3_index(c1, k2):
    return ((c1 >> 1) * 32 + (k2 >> 2)) * 8 + (c1 & 1) * 4 + (k2 & 3)


def recover3_model(cells, present):
    """k_rec3_in, k_rec3_cols, k_rec3_out with their lane and tile maps."""
    E = [int.from_bytes(cells[32 * w:32 * w + 32], "big") % R for w in range(8192)]
    M = [None] * 8192
    izc = [None] * 128
    for q in range(32):  # k_rec3_in: cells k2 = 4q + cc
        ks = [4 * q + vv for vv in range(4)]
        zd = vanish(present, [ROOTS[64 * k] for k in ks], False)
        pc = vanish(present, [SEVEN64 * ROOTS[64 * k] % R for k in ks], True)
        for vv in range(4):
            izc[4 * q + vv] = 64 * pc[vv] * KINV % R
        v = []
        for L in range(64):
            cc, j = L & 3, L >> 2
            k2 = 4 * q + cc
            # element brev6(k1) of cell brev7(k2), k1 = j + 16 r
            v.append([E[64 * brev(k2, 7) + brev(j + 16 * r, 6)] * zd[cc] % R for r in range(4)])
        X = r4_pass_b(v, ROOTS_INV)
        for L in range(64):
            cc, j = L & 3, L >> 2
            k2 = 4 * q + cc
            for r in range(4):
                c1 = j + 16 * r
                M[m3_index(c1, k2)] = X[L][r] * ROOTS_INV[k2 * c1 % 8192] % R
    for q in range(32):  # k_rec3_cols: columns c1 = 2q + c
        def c1_of(L, q=q):
            return 2 * q + (L >> 5)
        v = [[M[m3_index(c1_of(L), (L & 31) + 32 * r)] for r in range(4)] for L in range(64)]

        def reorder(o):
            return [[row[0], row[2], row[1], row[3]] for row in o]
        o = r4_pass_af(v, ROOTS_INV, lambda L, t: SHIFT_FWD[c1_of(L) + 64 * r4_k2(L & 31, t)])
        o = r4_pass_af(reorder(o), ROOTS, lambda L, t: izc[r4_k2(L & 31, t)])
        o = r4_pass_af(reorder(o), ROOTS_INV, lambda L, t: SHIFT_INV[c1_of(L) + 64 * r4_k2(L & 31, t)])
        o = r4_pass_af(reorder(o), ROOTS, lambda L, t: ROOTS[c1_of(L) * r4_k2(L & 31, t) % 8192])
        for L in range(64):
            for t in range(4):
                M[m3_index(c1_of(L), r4_k2(L & 31, t))] = o[L][t]
    res = bytearray(CELLS)
    for q in range(32):  # k_rec3_out
        v = [[M[m3_index((L >> 2) + 16 * r, 4 * q + (L & 3))] for r in range(4)] for L in range(64)]
        X = r4_pass_b(v, ROOTS)
        for L in range(64):
            cc, j = L & 3, L >> 2
            k2 = 4 * q + cc
            for r in range(4):
                w = 64 * brev(k2, 7) + brev(j + 16 * r, 6)
                res[32 * w:32 * w + 32] = X[L][r].to_bytes(32, "big")
    return bytes(res)


@pytest.mark.parametrize("consistent", [True, False])
def test_three_kernel_recovery_matches_oracle(oracle_mod, consistent):
    """The cell-first split: T2's per-cell DFT_64, the division by the
    per-cell constant Z on the coset and T3's per-cell inverse DFT_64 cancel
    to the factor 64; the rest is per cell, per column, per cell."""
    rng = random.Random(70 + consistent)
    if consistent:
        rc, cells = oracle_mod.blob_to_cells(random_blob(rng))
        assert rc == 0
        keep = set(rng.sample(range(128), 64))
    else:
        cells = rng.randbytes(CELLS)
        keep = set(rng.sample(range(128), 97))
    present = bytes(1 if c in keep else 0 for c in range(128))
    masked = bytearray(rng.randbytes(CELLS))
    for c in keep:
        masked[2048 * c:2048 * (c + 1)] = cells[2048 * c:2048 * (c + 1)]
    rc, want = oracle_mod.recover_cells(bytes(masked), present, 1)
    assert rc == 0
    got = recover3_model(bytes(masked), present)
    assert got == want
    if consistent:
        assert got == cells


def test_m3_tiles():
    """Every (c1, k2) has its own slot; a column pair is one contiguous
    256-value block and 4 cells x 2 columns one 8-value tile."""
    idx = sorted(m3_index(c1, k2) for c1 in range(64) for k2 in range(128))
    assert idx == list(range(8192))
    for q in range(32):
        assert sorted(m3_index(2 * q + c, k2) for c in range(2) for k2 in range(128)) == \
            list(range(256 * q, 256 * q + 256))


def nth_set_bit(w, k):
    """recover.hip's nth_set_bit: binary search on the low halves' popcounts."""
    pos = 0
    for half in (32, 16, 8, 4, 2, 1):
        c = bin(w & ((1 << half) - 1)).count("1")
        if k >= c:
            k -= c
            w >>= half
            pos += half
    return pos


def test_vanishing_parts_cover_wanted_cells():
    """vanishing_product's work split: the two ballots' 128-bit mask, part p
    of a value taking the wanted cells of rank p, p + PARTS, ... -- every
    wanted cell exactly once, no other cell."""
    rng = random.Random(12)
    for trial in range(300):
        present = [rng.random() < rng.random() for _ in range(128)]
        for want in (False, True):
            m0 = sum(1 << l for l in range(64) if present[l] == want)
            m1 = sum(1 << l for l in range(64) if present[64 + l] == want)
            n0 = bin(m0).count("1")
            n = n0 + bin(m1).count("1")
            for parts in (8, 16):
                got = []
                for p in range(parts):
                    for t in range(p, n, parts):
                        got.append(nth_set_bit(m0, t) if t < n0 else 64 + nth_set_bit(m1, t - n0))
                assert sorted(got) == [c for c in range(128) if present[c] == want]
