"""Phase B's three convolutions, restated on the CPU and checked against the
definition (no GPU): the folds k_ladder doubles (Fc, Fn, F-), the taps and signed
8-bit GLV digit items of build_phase_b_items, the slot and wrap rules of
k_phase_b, the bucket weights of k_bucket_reduce and the recombination of
k_finalize, on scalars standing in for points (phase B is linear; phi is
multiplication by lambda).  The result must equal FFT_128(trunc_64(IFFT_128(u))),
the map the reference's split circulant computes.  A wrong tap, slot, wrap
sign or recombination sign fails here; the GPU tests then check the proofs
bit-exactly."""
import random

R = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
LAMBDA = 0xAC45A4010001A40200000000FFFFFFFF
DIGIT_BITS = 8
LAD_POS = 16  # signed bytes of a balanced GLV half's magnitude (below 2^127: no carry position)
BUCKETS = 128


def circulant_taps():
    w = pow(7, (R - 1) // 128, R)

    def K(e):
        if e == 0:
            return pow(2, -1, R)
        if e % 2 == 0:
            return 0
        return pow(64 * (1 - pow(w, e, R)), -1, R)

    half = pow(2, -1, R)
    kp = [(K(e) + K(e + 64)) * half % R for e in range(64)]
    km = [(K(e) - K(e + 64)) * half % R for e in range(64)]
    return kp, km


def segment_taps(kp, km):
    half = pow(2, -1, R)
    kc = [(kp[e] + kp[e + 32]) * half % R for e in range(32)]
    kn = [(kp[e] - kp[e + 32]) * half % R for e in range(32)]
    return [kc, kn, km]


def glv_split_signed(k):
    """glv_split_signed: balanced signed halves with k = h1 + lambda h2 mod r."""
    k1, k2 = k % LAMBDA, k // LAMBDA
    half = LAMBDA // 2
    if k1 > half:
        k1, k2 = k1 - LAMBDA, k2 + 1
    if k2 > half:
        k1, k2 = k1 - 1, k2 - (LAMBDA + 1)
    assert (k1 + LAMBDA * k2 - k) % R == 0
    assert abs(k1) <= half + 1 and abs(k2) <= half + 1 and half + 1 < 1 << 127
    return k1, k2


def items_of(taps):
    """(bucket k) -> [(e, pos, neg)]: build_phase_b_items' balanced GLV split
    and signed 8-bit digits of the halves' magnitudes, each digit's sign
    flipped with its half's."""
    lists = [[] for _ in range(BUCKETS)]
    win = 1 << DIGIT_BITS
    for e, t in enumerate(taps):
        for g, h in enumerate(glv_split_signed(t)):
            part, hneg = abs(h), h < 0
            carry = 0
            for d in range(LAD_POS):
                v = ((part >> (DIGIT_BITS * d)) & (win - 1)) + carry
                carry = 0
                if v > win // 2:
                    v -= win
                    carry = 1
                if v:
                    lists[abs(v) - 1].append((e, d + g * LAD_POS, (v < 0) != hneg))
            assert carry == 0
    return lists


def output_rules(a):
    """k_phase_b's PhaseBOut: (segment, aa, base)."""
    seg = 0 if a < 32 else (1 if a < 64 else 2)
    aa = a & 31 if a < 64 else a & 63
    base = a & 32 if a < 64 else 64
    return seg, aa, base


def phase_b_model(u):
    kp, km = circulant_taps()
    taps = segment_taps(kp, km)
    lists = [items_of(t) for t in taps]
    fp = [(u[j] + u[j + 64]) % R for j in range(64)]
    fm = [(u[j] - u[j + 64]) % R for j in range(64)]
    # k_ladder's chains: Fc (slots 0..31), Fn (32..63), F- (64..127)
    f = [(fp[i] + fp[i + 32]) % R for i in range(32)] + [(fp[i] - fp[i + 32]) % R for i in range(32)] + fm
    # ladder positions: 2^(8d) x and phi(2^(8d) x) = lambda 2^(8d) x
    pts = [[x * pow(2, DIGIT_BITS * d, R) % R for d in range(LAD_POS)] for x in f]
    cvals = []
    for a in range(128):
        seg, aa, base = output_rules(a)
        total = 0
        for k in range(BUCKETS):
            bucket = 0
            for e, pos, neg in lists[seg][k]:
                if seg != 0 and e > aa:
                    neg = not neg
                slot = base + ((aa - e) & (63 if seg == 2 else 31))
                x = pts[slot][pos] if pos < LAD_POS else pts[slot][pos - LAD_POS] * LAMBDA % R
                bucket += -x if neg else x
            total += (k + 1) * bucket
        cvals.append(total % R)
    out = [0] * 128
    for a in range(64):
        yc, yn = cvals[a & 31], cvals[32 + (a & 31)]
        cp = (yc + yn) % R if a < 32 else (yc - yn) % R
        cm = cvals[64 + a]
        out[a] = (cp + cm) % R
        out[a + 64] = (cp - cm) % R
    return out, taps


def test_phase_b_segments_equal_ifft_trunc_fft():
    rng = random.Random(8)
    w = pow(7, (R - 1) // 128, R)
    inv128 = pow(128, -1, R)
    u = [rng.randrange(R) for _ in range(128)]
    h = [sum(u[j] * pow(w, -j * t, R) for j in range(128)) * inv128 % R for t in range(128)]
    h = h[:64] + [0] * 64
    want = [sum(h[t] * pow(w, t * k, R) for t in range(128)) % R for k in range(128)]
    got, taps = phase_b_model(u)
    assert got == want
    # the cyclic half's sub-convolutions keep 17 nonzero taps each (e = 0, odd e)
    assert [sum(1 for x in t if x) for t in taps] == [17, 17, 33]


def work_order(lists):
    """build_phase_b_items' class order: (segment, bucket) by item count,
    longest first (stable), and the segment-2 classes before each class."""
    cls = sorted(((seg, k) for seg in range(3) for k in range(BUCKETS)),
                 key=lambda c: -len(lists[c[0]][c[1]]))
    n2 = [0]
    for seg, _ in cls:
        n2.append(n2[-1] + (seg == 2))
    return cls, n2


def phase_b_item(item, nblobs, cls, n2, band=32):
    """k_phase_b's item -> (segment, bucket, [(lane, blob, a)]) in bands of
    `band` blobs, classes longest first within a band."""
    band_items = BUCKETS * 2 * band
    nfull, rest = divmod(nblobs, band)
    b, j = divmod(item, band_items)
    b0 = b * band
    bb = band if b < nfull else rest
    pb = (bb + 1) // 2
    start = lambda c: n2[c] * bb + (c - n2[c]) * pb
    c = 0
    step = 1 << (len(cls).bit_length() - 1)
    while step:
        t = c + step
        if t < len(cls) and start(t) <= j:
            c = t
        step >>= 1
    assert start(c) <= j and (c + 1 == len(cls) or start(c + 1) > j)
    g = j - start(c)
    seg, k = cls[c]
    lanes = []
    for lane in range(64):
        if seg == 2:
            blob, a = b0 + g, 64 + lane
        else:
            blob, a = b0 + 2 * g + (lane >> 5), seg * 32 + (lane & 31)
        lanes.append((lane, blob, a))
    return seg, k, lanes


def test_phase_b_work_order_covers_every_bucket_once():
    """k_phase_b's item -> (blob, output, bucket) map over bands of 32 blobs
    with the classes longest first: every bucket of every output exactly
    once, one item list per wave, odd blob counts leaving the last pair's
    upper half idle, and within a band the item lengths never increase."""
    kp, km = circulant_taps()
    lists = [items_of(t) for t in segment_taps(kp, km)]
    cls, n2 = work_order(lists)
    for nblobs in (1, 2, 3, 5, 31, 32, 33, 64, 67):
        rest = nblobs % 32
        nitems = (nblobs // 32) * BUCKETS * 64 + BUCKETS * (2 * ((rest + 1) // 2) + rest)
        seen = []
        prev = None
        for item in range(nitems):
            seg, k, lanes = phase_b_item(item, nblobs, cls, n2)
            for lane, blob, a in lanes:
                if blob < nblobs:
                    assert output_rules(a)[0] == seg
                    seen.append((blob, a, k))
            band, j = divmod(item, BUCKETS * 64)
            n = len(lists[seg][k])
            if j and prev is not None:
                assert n <= prev
            prev = n
        assert sorted(seen) == [(b, a, k) for b in range(nblobs) for a in range(128) for k in range(BUCKETS)]


def reduce_lane_model(B, per, l, handle_doubling):
    """k_bucket_reduce's per-lane running sums on scalars standing in for
    points (0 = the identity): returns (rs, s, flagged), flagged when an
    addition meets a == +-b with both non-identity (the case the fast
    additions flag)."""
    flagged = False

    def add(a, b):
        nonlocal flagged
        if a and b and (a == b or (a + b) % R == 0):
            flagged = True
        return (a + b) % R

    s = rs = 0
    single = False
    for i in range(per - 1, -1, -1):
        b = B[l * per + i]
        s = add(s, b)
        if handle_doubling and b == 0 and single:
            rs = 2 * rs % R
            single = False
        else:
            single = rs == 0 and s != 0
            rs = add(rs, s)
    return rs, s, flagged


def test_reduction_empty_buckets_take_the_doubling():
    """Empty buckets (the identity) are the same for every blob, since the
    item lists are.  An empty bucket right after a lane's first non-empty one
    leaves the running sum equal to s: without the doubling step the fast
    additions would flag it and send every output of the segment to the
    checked recompute.  With it no lane count flags, on the real lists'
    empty buckets and on an empty bucket at every position, and the weighted
    sum stays exact."""
    kp, km = circulant_taps()
    taps = segment_taps(kp, km)
    lists = [items_of(t) for t in taps]
    patterns = [{k for k in range(BUCKETS) if not lists[s][k]} for s in range(3)]
    patterns += [{k} for k in range(BUCKETS)] + [{k, k + 1} for k in range(BUCKETS - 1)]
    rng = random.Random(9)
    naive_flags = set()
    for empty in patterns:
        B = [0 if k in empty else rng.randrange(1, R) for k in range(BUCKETS)]
        want = sum((k + 1) * B[k] for k in range(BUCKETS)) % R
        for lanes in (4, 8, 16, 32):
            per = BUCKETS // lanes
            got = 0
            for l in range(lanes):
                rs, s, flagged = reduce_lane_model(B, per, l, True)
                assert not flagged, (sorted(empty), lanes, l)
                got += rs + per * l * s
                if reduce_lane_model(B, per, l, False)[2]:
                    naive_flags.add(lanes)
            assert got % R == want
    assert naive_flags == {4, 8, 16, 32}  # the case the doubling step exists for


def test_balanced_glv_split_edges():
    """glv_split_signed on the edge scalars and random ones: the halves
    rebuild k mod r and stay below 2^127 (16 signed bytes, no carry)."""
    rng = random.Random(12)
    half = LAMBDA // 2
    edge = [0, 1, half, half + 1, LAMBDA - 1, LAMBDA, LAMBDA + 1, LAMBDA * LAMBDA, LAMBDA * (LAMBDA + 1),
            R - 1, R - 2, LAMBDA * half, LAMBDA * (half + 1) + half + 1]
    for k in edge + [rng.randrange(R) for _ in range(2000)]:
        glv_split_signed(k % R)
