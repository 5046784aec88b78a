// Host runtime behind the C ABI of include/vkzg.h and include/vkzg_amd.h.
//
// One prover = one HIP device, one stream (plus sub-streams a large batch is
// spread over), resident tables and a working set sized for max_batch_size
// blobs.  Calls are serialised by a mutex (reference src/vkzg.cpp:159,1188,
// 1206).  Every HIP failure maps to VKZG_ERR_GPU.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <system_error>
#include <thread>

#include "../../include/vkzg_amd.h"
#include "runtime.h"

extern "C" const uint8_t vkzg_amd_mainnet_g1[];
extern "C" const uint8_t vkzg_amd_mainnet_g1_end[];

using namespace vkzg_amd;

// Fp is 14 x 28-bit limbs in 32-bit words (56 bytes) in memory as in registers.
static_assert(sizeof(Fr) == 32 && sizeof(Fp) == 56, "field layout");
static_assert(sizeof(G1A) == 112 && sizeof(G1J) == 168 && sizeof(G1X) == 224 && sizeof(G1AT) == 128,
              "point layout");

namespace vkzg_amd {

constexpr size_t kSetupBytes = 4096 * 48;
constexpr uint32_t kDefaultBatch = 256;

uint64_t table_bytes_for(uint32_t c, uint32_t k = 0) {
    return 8192ull * window_plan(c, k).entries() * sizeof(G1AT);
}

// Window plan of the resident phase-A table: $VKZG_AMD_WINDOW = c if set
// (6..14, all windows c bits); else the fewest windows W whose cheapest plan
// fits in 65% of the free memory, c = floor(255 / W) bits with the top
// k = 255 - W c windows one bit wider.  On a fresh 288 GB MI355X that is
// W = 18 (14 and 15 bits, 168 GiB); W = 19 is 108 GiB, 20 is 70 GiB, 21 is
// 48 GiB, 23 is 25 GiB: a second prover in the same process gets more windows.
WindowPlan choose_window(size_t free_bytes) {
    if (const char *env = std::getenv("VKZG_AMD_WINDOW")) {
        const int c = std::atoi(env);
        if (c >= 6 && c <= 14) return window_plan((uint32_t)c, 0);
    }
    for (uint32_t W = 18; W <= 36; W++) {
        const uint32_t c = 255 / W, k = 255 - W * c;
        if (c > 14 || c < 6) continue;
        if (table_bytes_for(c, k) <= free_bytes / 20 * 13) return window_plan(c, k);
    }
    return window_plan(6, 0);
}

// GLV split of a phase-B scalar (its halves are written in signed digits
// below, item = e | d << 8 | (v < 0) << 16, grouped by bucket |v|-1):
// k = k1 + lambda k2 with k1 = k mod lambda and k2 = k div lambda, for the
// G1 endomorphism phi(x, y) = (beta x, y) = [lambda] P, lambda = z^2 - 1 =
// 0xac45a4010001a40200000000ffffffff.  k < r = lambda^2 + lambda + 1 keeps
// both halves below 2^128.
void glv_split(const uint32_t k[8], unsigned __int128 &k1, unsigned __int128 &k2) {
    const unsigned __int128 lambda =
        ((unsigned __int128)0xac45a4010001a402ull << 64) | (unsigned __int128)0x00000000ffffffffull;
    unsigned __int128 rem = 0, q = 0;
    for (int bit = 255; bit >= 0; bit--) {
        const bool top = (rem >> 127) != 0;  // rem < lambda < 2^128 before the shift
        rem = (rem << 1) | ((k[bit / 32] >> (bit % 32)) & 1u);
        if (top || rem >= lambda) {
            rem -= lambda;
            if (bit >= 128) throw VKZG_ERR_SETUP;  // impossible for k < 2^255: k / lambda < 2^128
            q |= (unsigned __int128)1 << bit;
        }
    }
    k1 = rem;
    k2 = q;
}

// Balanced signed GLV halves: k = s1 m1 + lambda s2 m2 (mod r), signs s1, s2
// and magnitudes m1, m2 <= lambda/2 + 1 < 2^127 (so 16 signed bytes hold each
// half with no carry position).  From glv_split's k1 < lambda, k2 <= lambda + 1:
// k1 > lambda/2 takes k1 - lambda and k2 + 1; then k2 > lambda/2 takes
// k2 - (lambda + 1) and k1 - 1, since lambda (lambda + 1) = r - 1 = -1 mod r.
void glv_split_signed(const uint32_t k[8], unsigned __int128 m[2], bool neg[2]) {
    const unsigned __int128 lambda =
        ((unsigned __int128)0xac45a4010001a402ull << 64) | (unsigned __int128)0x00000000ffffffffull;
    const unsigned __int128 half = lambda >> 1;
    glv_split(k, m[0], m[1]);
    neg[0] = neg[1] = false;
    if (m[0] > half) {
        m[0] = lambda - m[0];
        neg[0] = true;
        m[1] += 1;  // k2 <= lambda here: k < r
    }
    if (m[1] > half) {
        m[1] = lambda + 1 - m[1];  // k2 <= lambda + 1
        neg[1] = true;
        if (neg[0]) {
            m[0] += 1;
        } else if (m[0] == 0) {
            m[0] = 1;
            neg[0] = true;
        } else {
            m[0] -= 1;
        }
    }
    if ((m[0] >> 127) != 0 || (m[1] >> 127) != 0) throw VKZG_ERR_SETUP;  // cannot happen
}

// Canonical Fr arithmetic on the host (little-endian 32-bit limbs).
void fr_host_add(uint32_t r[8], const uint32_t a[8], const uint32_t b[8]) {
    uint32_t t[8], u[8];
    uint64_t c = 0;
    for (int i = 0; i < 8; i++) {
        c += (uint64_t)a[i] + b[i];
        t[i] = (uint32_t)c;
        c >>= 32;
    }
    int64_t br = 0;
    for (int i = 0; i < 8; i++) {
        const int64_t d = (int64_t)t[i] - cfr::P[i] + br;
        u[i] = (uint32_t)d;
        br = d >> 32;
    }
    const bool keep = br < 0;  // a + b < r (a + b < 2^256: no carry out)
    for (int i = 0; i < 8; i++) r[i] = keep ? t[i] : u[i];
}
void fr_host_neg(uint32_t r[8], const uint32_t a[8]) {
    bool zero = true;
    for (int i = 0; i < 8; i++) zero = zero && a[i] == 0;
    int64_t br = 0;
    for (int i = 0; i < 8; i++) {
        const int64_t d = (int64_t)cfr::P[i] - a[i] + br;
        r[i] = zero ? 0u : (uint32_t)d;
        br = d >> 32;
    }
}
// a / 2 mod r: (a + r) / 2 for odd a (r is odd).
void fr_host_half(uint32_t r[8], const uint32_t a[8]) {
    uint32_t t[9];
    const bool odd = (a[0] & 1u) != 0;
    uint64_t c = 0;
    for (int i = 0; i < 8; i++) {
        c += (uint64_t)a[i] + (odd ? cfr::P[i] : 0u);
        t[i] = (uint32_t)c;
        c >>= 32;
    }
    t[8] = (uint32_t)c;
    for (int i = 0; i < 8; i++) r[i] = (t[i] >> 1) | (t[i + 1] << 31);
}

// Phase-B taps of the three convolutions (canonical), from the split
// circulant K+ (cyclic, 64 taps) and K- (negacyclic, 64 taps) of
// k_phase_b_scalars.  The cyclic half splits once more by
// X^64 - 1 = (X^32 - 1)(X^32 + 1):
//   segment 0: Kc[e] = (K+[e] + K+[e+32]) / 2, cyclic over Fc = F+lo + F+hi,
//   segment 1: Kn[e] = (K+[e] - K+[e+32]) / 2, negacyclic over Fn = F+lo - F+hi,
//   segment 2: K-[e], negacyclic over F- (64 taps),
// so that C+[a] = Yc[a] + Yn[a] and C+[a+32] = Yc[a] - Yn[a] for a < 32
// (k_finalize).  K+ has only 33 nonzero taps (e = 0 and odd e), Kc and Kn 17
// each: the cyclic half costs 64 x 17 scalar-point terms instead of 64 x 33.
void phase_b_taps(const Fr *kpm, std::vector<Fr> taps[3]) {
    taps[0].assign(32, Fr{});
    taps[1].assign(32, Fr{});
    for (int e = 0; e < 32; e++) {
        uint32_t sum[8], neg[8], diff[8];
        fr_host_add(sum, kpm[e].v, kpm[e + 32].v);
        fr_host_neg(neg, kpm[e + 32].v);
        fr_host_add(diff, kpm[e].v, neg);
        fr_host_half(taps[0][e].v, sum);
        fr_host_half(taps[1][e].v, diff);
    }
    taps[2].assign(kpm + 64, kpm + 128);
}

// Phase-B item lists: for each segment s and bucket k, the (e, d, sign) with
// signed kDigitBits-bit digit of magnitude k+1 at position d of the tap
// taps[s][e] (the byte-plus-carry recoding of k_phase_b_split.comp:25-61, on
// the magnitudes of the balanced signed GLV halves, each digit's sign flipped
// with its half's), where d < kLadPos indexes the digits of k1 and
// kLadPos <= d < 2 kLadPos those of k2.  Offsets: kBuckets + 1 per segment.
void build_phase_b_items(const Fr *kpm, std::vector<uint32_t> &items, std::vector<uint32_t> &offsets) {
    constexpr uint32_t kWin = 1u << kDigitBits, kHalf = kWin / 2;
    std::vector<Fr> taps[3];
    phase_b_taps(kpm, taps);
    items.clear();
    offsets.assign(3 * (kBuckets + 1), 0);
    for (int seg = 0; seg < 3; seg++) {
        std::vector<uint32_t> lists[kBuckets];
        for (uint32_t e = 0; e < (uint32_t)taps[seg].size(); e++) {
            unsigned __int128 parts[2];
            bool half_neg[2];
            glv_split_signed(taps[seg][e].v, parts, half_neg);
            for (int g = 0; g < 2; g++) {
                int carry = 0;
                for (uint32_t d = 0; d < kLadPos; d++) {
                    const uint32_t sh = kDigitBits * d;  // the top position may start at bit 128
                    const uint32_t bits = sh < 128 ? (uint32_t)(parts[g] >> sh) & (kWin - 1) : 0u;
                    int v = (int)bits + carry;
                    carry = 0;
                    if (v > (int)kHalf) {
                        v -= (int)kWin;
                        carry = 1;
                    }
                    if (v == 0) continue;
                    const uint32_t mag = (uint32_t)(v < 0 ? -v : v);
                    const uint32_t pos = d + (uint32_t)g * kLadPos;
                    const bool dneg = (v < 0) != half_neg[g];
                    lists[mag - 1].push_back(e | (pos << 8) | ((dneg ? 1u : 0u) << 16));
                }
                if (carry) throw VKZG_ERR_SETUP;  // cannot happen: magnitudes are below 2^127
            }
        }
        for (uint32_t k = 0; k < kBuckets; k++) {
            offsets[seg * (kBuckets + 1) + k] = (uint32_t)items.size();
            items.insert(items.end(), lists[k].begin(), lists[k].end());
        }
        offsets[seg * (kBuckets + 1) + kBuckets] = (uint32_t)items.size();
    }
    // Work order of k_phase_b: the 3 kBuckets (segment, bucket) classes by
    // item count, longest first (so that a band of blobs ends on its
    // shortest buckets), as seg << 16 | k, then the number of segment-2
    // classes before each class (3 kBuckets + 1 entries).
    std::vector<uint32_t> cls;
    for (uint32_t seg = 0; seg < 3; seg++)
        for (uint32_t k = 0; k < kBuckets; k++) cls.push_back(seg << 16 | k);
    auto len = [&](uint32_t c) {
        const uint32_t seg = c >> 16, k = c & 0xffffu;
        return offsets[seg * (kBuckets + 1) + k + 1] - offsets[seg * (kBuckets + 1) + k];
    };
    std::stable_sort(cls.begin(), cls.end(), [&](uint32_t x, uint32_t y) { return len(x) > len(y); });
    offsets.insert(offsets.end(), cls.begin(), cls.end());
    uint32_t n2 = 0;
    for (uint32_t c : cls) {
        offsets.push_back(n2);
        n2 += (c >> 16) == 2 ? 1u : 0u;
    }
    offsets.push_back(n2);
}

// Fills t with a freshly built phase-A table of the window plan: the W chains
// 2^bit(w) xext (affine), then every multiple d * 2^bit(w) xext,
// d <= 2^(width(w)-1).
void build_phase_a_table(vkzg_prover *p, PhaseATable &t, WindowPlan plan) {
    hipStream_t s = p->stream;
    const uint32_t c = plan.c;
    t.c = c;
    t.W = plan.W;
    t.H = plan.H();
    t.k = plan.k;
    t.bytes = 8192ull * plan.entries() * sizeof(G1AT);
    t.table = p->arena.dalloc<G1AT>(t.bytes / sizeof(G1AT));
    G1J *pwj = p->arena.dalloc<G1J>((size_t)t.W * 8192);
    Fp *pw_prefix = p->arena.dalloc<Fp>((size_t)t.W * 8192);
    G1A *pw = p->arena.dalloc<G1A>((size_t)t.W * 8192);
    ck(launch_table_chains(pwj, p->xext, plan, s));
    ck(launch_jac_to_affine(pw, pwj, pw_prefix, (uint64_t)t.W * 8192, 16, s));
    ck(hipStreamSynchronize(s));
    p->arena.dfree(pwj);
    p->arena.dfree(pw_prefix);
    const uint32_t S = t.H < 64 ? t.H : 64;
    const uint64_t total_threads = 8192ull * (t.W + t.k) * (t.H / S);  // S entries each
    const uint32_t slice = 1u << 18;
    G1J *scratch = p->arena.dalloc<G1J>((size_t)S * slice);
    Fp *pref = p->arena.dalloc<Fp>((size_t)S * slice);
    for (uint64_t t0 = 0; t0 < total_threads; t0 += slice) {
        const uint32_t n = (uint32_t)std::min<uint64_t>(slice, total_threads - t0);
        ck(launch_table_multiples(t.table, pw, plan, t0, n, scratch, pref, s));
    }
    ck(hipStreamSynchronize(s));
    p->arena.dfree(scratch);
    p->arena.dfree(pref);
    p->arena.dfree(pw);
}

PhaseATable phase_a_table(const vkzg_prover *p, uint32_t window_bits) {
    if (window_bits >= kExtraWindowMin && window_bits <= kExtraWindowMax && window_bits != p->c) {
        const PhaseATable &t = p->extra_tables[window_bits - kExtraWindowMin];
        if (t.table) return t;
    }
    PhaseATable m;
    m.c = p->c;
    m.W = p->W;
    m.H = p->H;
    m.k = p->wide;
    m.bytes = p->table_bytes;
    m.table = p->table;
    return m;
}

void build_tables(vkzg_prover *p, const uint8_t *setup) {
    hipStream_t s = p->stream;
    p->roots = p->arena.dalloc<Fr>(8192);
    p->roots_inv = p->arena.dalloc<Fr>(8192);
    p->shift_fwd = p->arena.dalloc<Fr>(8192);
    p->shift_inv = p->arena.dalloc<Fr>(8192);
    ck(launch_roots(p->roots, p->roots_inv, p->shift_fwd, p->shift_inv, s));
    // the same tables as FrL entries, for the radix-4 recovery kernels
    Fr *const fr_tabs[4] = {p->roots, p->roots_inv, p->shift_fwd, p->shift_inv};
    for (int k = 0; k < 4; k++) {
        p->frl_tabs[k] = p->arena.dalloc<uint4>(3 * 8192);
        ck(launch_frl_table(p->frl_tabs[k], fr_tabs[k], 8192, s));
    }

    // Decompress the setup.
    uint8_t *d_setup = p->arena.dalloc<uint8_t>(kSetupBytes);
    uint32_t *d_flag = p->arena.dalloc<uint32_t>(1);
    ck(hipMemcpyAsync(d_setup, setup, kSetupBytes, hipMemcpyHostToDevice, s));
    ck(hipMemsetAsync(d_flag, 0, 4, s));
    p->srs = p->arena.dalloc<G1A>(4096);
    ck(launch_decompress(p->srs, d_setup, 4096, d_flag, s));
    uint32_t flag = 0;
    ck(hipMemcpyAsync(&flag, d_flag, 4, hipMemcpyDeviceToHost, s));
    ck(hipStreamSynchronize(s));
    p->arena.dfree(d_setup);
    if (flag) throw VKZG_ERR_SETUP;

    // x_ext_fft.
    G1J *xj = p->arena.dalloc<G1J>(8192);
    Fp *prefix = p->arena.dalloc<Fp>(8192);
    p->xext = p->arena.dalloc<G1A>(8192);
    ck(launch_xext_fft(xj, p->srs, p->roots, s));
    ck(launch_jac_to_affine(p->xext, xj, prefix, 8192, 8, s));

    // Phase-B scalars -> item lists.
    Fr *d_kpm = p->arena.dalloc<Fr>(128);
    ck(launch_phase_b_scalars(d_kpm, p->roots, s));
    std::vector<Fr> kpm(128);
    ck(hipMemcpyAsync(kpm.data(), d_kpm, 128 * sizeof(Fr), hipMemcpyDeviceToHost, s));
    ck(hipStreamSynchronize(s));
    std::vector<uint32_t> items, offsets;
    build_phase_b_items(kpm.data(), items, offsets);
    p->kb_count = (uint32_t)items.size();
    // segments 0 and 1 serve 32 outputs per blob each, segment 2 serves 64
    p->kb_per_blob = 0;
    for (int seg = 0; seg < 3; seg++)
        p->kb_per_blob += (seg == 2 ? 64u : 32u) *
                          (offsets[seg * (kBuckets + 1) + kBuckets] - offsets[seg * (kBuckets + 1)]);
    p->kb_items = p->arena.dalloc<uint32_t>(items.size());
    p->kb_offsets = p->arena.dalloc<uint32_t>(offsets.size());
    ck(hipMemcpyAsync(p->kb_items, items.data(), items.size() * 4, hipMemcpyHostToDevice, s));
    ck(hipMemcpyAsync(p->kb_offsets, offsets.data(), offsets.size() * 4, hipMemcpyHostToDevice, s));

    // Phase-A multiples table.
    size_t free_b = 0, total_b = 0;
    ck(hipMemGetInfo(&free_b, &total_b));
    PhaseATable main_table;
    build_phase_a_table(p, main_table, choose_window(free_b));
    p->c = main_table.c;
    p->W = main_table.W;
    p->H = main_table.H;
    p->wide = main_table.k;
    p->table_bytes = main_table.bytes;
    p->table = main_table.table;
    p->arena.dfree(xj);
    p->arena.dfree(prefix);
    p->arena.dfree(d_kpm);
    p->arena.dfree(d_flag);
}


uint64_t WorkSet::bytes_per_blob() {
    const uint64_t npos = 128ull * kLadPos;
    return kBlobBytes + kProofBytes + 4 * 4 + 128 + kMaxLadSplits * 128 * 64 * sizeof(Fr) +
           kMaxLadSplits * 128 * (kMaxGroups + 1) * sizeof(G1X) +
           npos * sizeof(G1J) + 2 * npos * sizeof(G1A) + npos * sizeof(Fp) +
           128 * kBuckets * (sizeof(G1X) + 1) + 128 * sizeof(G1X) + 2 * kCellsBytes + 128 +
           2 * 8192 * sizeof(Fr);
}

void WorkSet::allocate(uint32_t B) {
    cap = B;
    blobs = arena.dalloc<uint8_t>((size_t)B * kBlobBytes);
    proofs = arena.dalloc<uint8_t>((size_t)B * kProofBytes);
    err = arena.dalloc<uint32_t>(B);
    queue = arena.dalloc<uint32_t>(B);
    ck(hipMemset(queue, 0, (size_t)B * sizeof(uint32_t)));
    queue_b = arena.dalloc<uint32_t>(B);
    ck(hipMemset(queue_b, 0, (size_t)B * sizeof(uint32_t)));
    queue_r = arena.dalloc<uint32_t>(B);
    ck(hipMemset(queue_r, 0, (size_t)B * sizeof(uint32_t)));
    scal = arena.dalloc<Fr>((size_t)B * 128 * 64 * kMaxLadSplits);
    u = arena.dalloc<G1X>((size_t)B * 128 * (kMaxGroups + 1) * kMaxLadSplits);
    lad = arena.dalloc<G1J>((size_t)B * 128 * kLadPos);
    lad_aff = arena.dalloc<G1A>((size_t)B * 128 * kLadPos);
    lad_phi = arena.dalloc<G1A>((size_t)B * 128 * kLadPos);
    prefix = arena.dalloc<Fp>((size_t)B * 128 * kLadPos);
    buckets = arena.dalloc<G1X>((size_t)B * 128 * kBuckets);
    bucket_flags = arena.dalloc<uint8_t>((size_t)B * 128 * kBuckets);
    cval_flags = arena.dalloc<uint8_t>((size_t)B * 128);
    cvals = arena.dalloc<G1X>((size_t)B * 128);
    cells = arena.dalloc<uint8_t>((size_t)B * kCellsBytes);
    cells_out = arena.dalloc<uint8_t>((size_t)B * kCellsBytes);
    present = arena.dalloc<uint8_t>((size_t)B * 128);
    bufA = arena.dalloc<Fr>((size_t)B * 8192);
    bufB = arena.dalloc<Fr>((size_t)B * 8192);
}

// Spreading a batch over sub-streams was meant to overlap one sub-batch's
// latency-bound stages (ladder, inversions) with another's phase A.  Measured
// at 256 blobs it loses (28.8 ms in one stream, 32.4 ms over two, 33.1 ms over
// three): the ladder's workgroups wait for slots held by phase A, so the
// dependent chains get longer, not hidden.  The default tunings therefore use
// one stream; the autotuner may pick two, and $VKZG_AMD_SPLITS (1-4) forces a
// count.
Tuning tuning_for(const vkzg_prover *p, uint32_t nb) {
    Tuning t;
    {
        std::lock_guard<std::mutex> lock(p->tune_mu);
        t = p->tune[tune_class(nb)];
    }
    t.force_checked = p->force_checked.load();
    if (p->forced_splits > 0) t.splits = (uint32_t)p->forced_splits;
    if (t.splits > (uint32_t)kMaxSplits) t.splits = kMaxSplits;
    if (t.splits > nb) t.splits = nb;
    if (t.splits < 1) t.splits = 1;
    return t;
}

// Defaults before (or without) autotuning: small batches split each MSM into
// more window groups and the ladder into segments, so that more lanes work on
// the few MSMs and chains there are (one wave per SIMD already saturates
// phase A at 64 blobs); the normalisation chunk is fitted to the chip
// (norm_chunk 0).
void default_tunings(vkzg_prover *p) {
    static const uint32_t groups[kTuneClasses] = {10, 10, 4, 1, 1};
    static const uint32_t ladder_splits[kTuneClasses] = {4, 2, 1, 1, 1};
    static const uint32_t reduce_lanes[kTuneClasses] = {64, 64, 32, 16, 4};
    for (int c = 0; c < kTuneClasses; c++) {
        p->tune[c] = Tuning{};
        p->tune[c].groups = groups[c];
        p->tune[c].ladder_splits = ladder_splits[c];
        p->tune[c].reduce_lanes = reduce_lanes[c];
        p->tune[c].norm_chunk = 0;
    }
}

// Points per lane of the normalisation for norm_chunk 0: the smallest chunk
// whose lanes fit in one wave per SIMD.  Each lane's work is one inversion
// plus 7 products per point, issued at one wave per SIMD; a second wave on a
// SIMD doubles the stage, so 557,056 points (256 blobs) take chunks of 9 on
// 1024 SIMDs, not 8.
// Lanes per ladder chain: four while four-lane chains still fit one wave per
// SIMD (the chain's length sets the time there), two above that (the issue
// rate does).
uint32_t ladder_lanes(uint64_t chains, uint32_t cus) {
    return chains * 4 <= (uint64_t)cus * 4 * 64 ? 4u : 2u;
}

// Phase B on quads (four lanes per bucket) while the quad lanes come to at
// most kPhaseBQuadWaves waves per SIMD (8 blobs on 256 CUs): there the
// longest bucket's chain, more than the issue rate, sets the time.  Measured
// limits of 1, 2, 4 and 8 waves: each size up to 8 blobs was fastest on quads.
constexpr uint64_t kPhaseBQuadWaves = 8;
bool phase_b_quad(uint32_t nb, uint32_t cus) {
    return (uint64_t)nb * 128 * kBuckets * 4 <= (uint64_t)cus * 4 * 64 * kPhaseBQuadWaves;
}

uint32_t fit_norm_chunk(uint64_t count, uint32_t cus) {
    const uint64_t lanes = (uint64_t)cus * 4 * 64;
    const uint64_t c = (count + lanes - 1) / lanes;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(c, 256));
}

bool tuning_valid(const Tuning &t) {
    return t.groups >= 1 && t.groups <= kMaxGroups && t.blocks_per_cu >= 1 && t.blocks_per_cu <= 2 &&
           t.norm_chunk <= 256 && t.splits >= 1 && t.splits <= (uint32_t)kMaxSplits &&
           (t.rec_waves == 1 || t.rec_waves == 2 || t.rec_waves == 4) && (t.rec_radix == 2 || t.rec_radix == 4 || t.rec_radix == 8) &&
           (t.ladder_splits == 1 || t.ladder_splits == 2 || t.ladder_splits == 4) &&
           (t.reduce_lanes == 4 || t.reduce_lanes == 8 || t.reduce_lanes == 16 || t.reduce_lanes == 32 ||
            t.reduce_lanes == 64) &&
           (t.window_bits == 0 || (t.window_bits >= kExtraWindowMin && t.window_bits <= kExtraWindowMax));
}

RecTables rec_tables(const vkzg_prover *p) {
    return RecTables{p->roots,       p->roots_inv,   p->shift_fwd,   p->shift_inv,
                     p->frl_tabs[0], p->frl_tabs[1], p->frl_tabs[2], p->frl_tabs[3]};
}

void enqueue_recovery_into(vkzg_prover *p, Fr *A, Fr *B, uint8_t *d_out, const uint8_t *d_cells,
                           const uint8_t *d_present, uint32_t nb, hipStream_t s, StageClock &clk,
                           uint32_t radix, uint32_t waves) {
    const RecTables t = rec_tables(p);
    if (radix == 2) {
        // three kernels: per cell, per column (in place in A), per cell;
        // B holds the coset factors
        ck(launch_rec3(d_out, A, B, d_cells, d_present, t, nb, waves, s, 0));
        clk.mark("cells_in");
        ck(launch_rec3(d_out, A, B, d_cells, d_present, t, nb, waves, s, 1));
        clk.mark("columns");
        ck(launch_rec3(d_out, A, B, d_cells, d_present, t, nb, waves, s, 2));
        clk.mark("cells_out");
        return;
    }
    // T1 pass A: (E * Z) -> inverse
    ck(launch_rec_first(A, d_cells, 0, d_present, t, nb, radix, waves, s));
    clk.mark("t1_a");
    // T1 pass B, * 7^k / 8192 (coefficients on the coset), T2 pass A
    ck(launch_rec_mid(B, A, 0, d_present, t, nb, radix, waves, s));
    clk.mark("t1_b_t2_a");
    // T2 pass B (coset evaluations), / Z on the coset, T3 pass A
    ck(launch_rec_mid(A, B, 1, d_present, t, nb, radix, waves, s));
    clk.mark("t2_b_t3_a");
    // T3 pass B, * 7^-k / 8192 (coefficients of P), T4 pass A
    ck(launch_rec_mid(B, A, 2, d_present, t, nb, radix, waves, s));
    clk.mark("t3_b_t4_a");
    // T4 pass B: P on the whole domain, as cells
    ck(launch_rec_last(d_out, B, t, nb, radix, waves, s));
    clk.mark("t4_b");
}

namespace {

// One sub-batch: blobs [off, off + nb) of working set ws.
void proofs_range(vkzg_prover *p, WorkSet &ws, uint32_t off, uint32_t nb, uint8_t *d_proofs,
                  const uint8_t *d_blobs, hipStream_t s, StageClock &clk, const Tuning &tn,
                  size_t blob_stride = kBlobBytes, Fr *coeffs = nullptr) {
    uint32_t *err = ws.err + off;
    uint32_t *queue = ws.queue + off;
    Fr *scal = ws.scal + (size_t)off * 8192 * kMaxLadSplits;
    // phase-A partials [nb][splits][128][groups], then their sums [nb][splits][128]
    G1X *u = ws.u + (size_t)off * 128 * (kMaxGroups + 1) * kMaxLadSplits;
    G1X *usum = u + (size_t)nb * 128 * kMaxGroups * kMaxLadSplits;
    const size_t npos = (size_t)128 * kLadPos;  // ladder points per blob
    G1J *lad = ws.lad + off * npos;
    G1A *lad_aff = ws.lad_aff + off * npos, *lad_phi = ws.lad_phi + off * npos;
    Fp *prefix = ws.prefix + off * npos;
    G1X *buckets = ws.buckets + (size_t)off * 128 * kBuckets, *cvals = ws.cvals + (size_t)off * 128;
    ck(launch_fk20_scalars(scal, d_blobs, blob_stride, err, p->roots, p->roots_inv, queue, coeffs,
                           tn.ladder_splits, nb, s));
    clk.mark("fk20_scalars");
    const PhaseATable tab = phase_a_table(p, tn.window_bits);
    ck(launch_phase_a(u, scal, tab.table, tab.plan(), nb * 128 * tn.ladder_splits, tn.groups,
                      p->cus * tn.blocks_per_cu, queue, tn.force_checked, s));
    clk.mark("phase_a");
    if (tn.groups > 1) {
        const uint64_t nsum = (uint64_t)nb * 128 * tn.ladder_splits;  // quads while they fit a wave per SIMD
        ck(launch_sum_groups(usum, u, tn.groups, (uint32_t)nsum, nsum * 32 <= (uint64_t)p->cus * 4 * 64,
                             tn.force_checked, s));
        clk.mark("sum_groups");
    }
    ck(launch_ladder(lad, tn.groups > 1 ? usum : u, tn.ladder_splits, nb,
                     ladder_lanes((uint64_t)nb * 128 * tn.ladder_splits, p->cus), s));
    clk.mark("ladder");
    // Small unsplit batches skip the normalisation: phase B's quads add the
    // ladder points in XYZZ form, which take twice the bytes of affine ones
    // and sit in the affine buffers' memory (so 2 nb <= the workset's blobs).
    const bool quad_b = phase_b_quad(nb, p->cus);
    static_assert(sizeof(G1X) <= 2 * sizeof(G1A), "XYZZ ladder points must fit the affine buffers");
    const bool xyzz = quad_b && tn.splits == 1 && off == 0 && 2 * nb <= ws.cap;
    if (xyzz) {
        G1X *lx = reinterpret_cast<G1X *>(ws.lad_aff), *lpx = reinterpret_cast<G1X *>(ws.lad_phi);
        ck(launch_ladder_xyzz(lx, lpx, lad, (uint64_t)nb * npos, ws.queue_b + off, ws.queue_r + off, s));
        clk.mark("normalize");
        ck(launch_phase_b(buckets, ws.bucket_flags + (size_t)off * 128 * kBuckets, lx, lpx, p->kb_items,
                          p->kb_offsets, nb, p->cus * 2, true, true, ws.queue_b + off, tn.force_checked, s));
    } else {
        const uint32_t chunk = tn.norm_chunk ? tn.norm_chunk : fit_norm_chunk((uint64_t)nb * npos, p->cus);
        ck(launch_jac_to_affine(lad_aff, lad, prefix, (uint64_t)nb * npos, chunk, s));
        ck(launch_glv_phi(lad_phi, lad_aff, (uint64_t)nb * npos, ws.queue_b + off, ws.queue_r + off, s));
        clk.mark("normalize");
        ck(launch_phase_b(buckets, ws.bucket_flags + (size_t)off * 128 * kBuckets, lad_aff, lad_phi,
                          p->kb_items, p->kb_offsets, nb, p->cus * 2, quad_b, false, ws.queue_b + off,
                          tn.force_checked, s));
    }
    clk.mark("phase_b");
    ck(launch_bucket_reduce(cvals, ws.cval_flags + off * 128, buckets, nb, tn.reduce_lanes, p->cus * 2,
                            ws.queue_r + off, tn.force_checked, s));
    clk.mark("bucket_reduce");
    ck(launch_finalize(d_proofs, cvals, nb, (uint64_t)nb * 256 <= (uint64_t)p->cus * 4 * 64, tn.force_checked, s));
    clk.mark("finalize");
}

void recovery_range(vkzg_prover *p, WorkSet &ws, uint32_t off, uint32_t nb, uint8_t *d_out,
                    const uint8_t *d_cells, const uint8_t *d_present, hipStream_t s,
                    StageClock &clk, const Tuning &tn) {
    enqueue_recovery_into(p, ws.bufA + (size_t)off * 8192, ws.bufB + (size_t)off * 8192, d_out, d_cells,
                          d_present, nb, s, clk, tn.rec_radix, tn.rec_waves);
}

// Runs body(k, off, n, stream) for each of `splits` sub-batches, forked from
// and joined back into s.  The clock follows sub-batch 0.
template <typename Body>
void split_over_streams(StreamSet &ss, uint32_t nb, uint32_t splits, hipStream_t s,
                        StageClock &clk, Body body) {
    if (splits <= 1) {
        body(0u, 0u, nb, s);
        return;
    }
    ck(hipEventRecord(ss.fork, s));
    uint32_t off = 0;
    for (uint32_t k = 0; k < splits; k++) {
        const uint32_t n = nb / splits + (k < nb % splits ? 1u : 0u);
        hipStream_t q = ss.sub[k];
        ck(hipStreamWaitEvent(q, ss.fork, 0));
        if (k == 0 && clk.on) {
            clk.s = q;
            ck(hipEventRecord(clk.p->ev[0], q));
        }
        body(k, off, n, q);
        ck(hipEventRecord(ss.join[k], q));
        off += n;
    }
    for (uint32_t k = 0; k < splits; k++) ck(hipStreamWaitEvent(s, ss.join[k], 0));
}

}  // namespace

namespace {
// Every launch-shape field in its own bit range (graph cache key).
uint64_t tuning_word(const Tuning &tn) {
    return (uint64_t)tn.splits | ((uint64_t)tn.groups << 8) | ((uint64_t)tn.blocks_per_cu << 16) |
           ((uint64_t)tn.rec_waves << 20) | ((uint64_t)tn.norm_chunk << 24) |
           ((uint64_t)tn.ladder_splits << 36) | ((uint64_t)tn.reduce_lanes << 40) |
           ((uint64_t)tn.rec_radix << 48) | ((uint64_t)tn.window_bits << 52) |
           ((uint64_t)(tn.force_checked != 0) << 56);
}
}  // namespace

void enqueue_proofs(vkzg_prover *p, WorkSet &ws, StreamSet &ss, uint8_t *d_proofs,
                    const uint8_t *d_blobs, uint32_t nb, hipStream_t s, StageClock &clk,
                    const Tuning &tn, GraphCache *graphs) {
    StageClock off_clock(p, s, false);
    auto body = [&](hipStream_t s0) {
        split_over_streams(ss, nb, tn.splits, s0, clk, [&](uint32_t k, uint32_t off, uint32_t n, hipStream_t q) {
            proofs_range(p, ws, off, n, d_proofs + (size_t)off * kProofBytes,
                         d_blobs + (size_t)off * kBlobBytes, q, k == 0 ? clk : off_clock, tn);
        });
    };
    if (!graphs) return body(s);
    const GraphKey key{{1u, nb, tuning_word(tn), (uint64_t)(uintptr_t)d_proofs,
                        (uint64_t)(uintptr_t)d_blobs, (uint64_t)(uintptr_t)&ws}};
    enqueue_maybe_graph(p, *graphs, key, s, clk, body);
}

void enqueue_recovery(vkzg_prover *p, WorkSet &ws, StreamSet &ss, uint8_t *d_out,
                      const uint8_t *d_cells, const uint8_t *d_present, uint32_t nb, hipStream_t s,
                      StageClock &clk, const Tuning &tn, GraphCache *graphs) {
    StageClock off_clock(p, s, false);
    auto body = [&](hipStream_t s0) {
        split_over_streams(ss, nb, tn.splits, s0, clk, [&](uint32_t k, uint32_t off, uint32_t n, hipStream_t q) {
            recovery_range(p, ws, off, n, d_out + (size_t)off * kCellsBytes,
                           d_cells + (size_t)off * kCellsBytes, d_present + (size_t)off * 128, q,
                           k == 0 ? clk : off_clock, tn);
        });
    };
    if (!graphs) return body(s);
    const GraphKey key{{2u | ((uint64_t)(uintptr_t)d_present << 2), nb, tuning_word(tn),
                        (uint64_t)(uintptr_t)d_out, (uint64_t)(uintptr_t)d_cells,
                        (uint64_t)(uintptr_t)&ws}};
    enqueue_maybe_graph(p, *graphs, key, s, clk, body);
}

// Cells (the 2x extension, as compute_cells_and_kzg_proofs returns them) and
// proofs of nb blobs read at blob_stride: the proof pipeline also writes the
// blob coefficients to ws.bufA, and the forward evaluation passes of the
// recovery turn them into canonical cell bytes.  d_cells may be null.
void enqueue_cells_and_proofs(vkzg_prover *p, WorkSet &ws, uint8_t *d_cells, uint8_t *d_proofs,
                              const uint8_t *d_blobs, size_t blob_stride, uint32_t nb, hipStream_t s,
                              const Tuning &tn, GraphCache *graphs) {
    StageClock off_clock(p, s, false);
    auto body = [&](hipStream_t q) {
        proofs_range(p, ws, 0, nb, d_proofs, d_blobs, q, off_clock, tn, blob_stride,
                     d_cells ? ws.bufA : nullptr);
        if (d_cells) {
            const RecTables t = rec_tables(p);
            ck(launch_rec_first(ws.bufB, ws.bufA, 1, nullptr, t, nb, tn.rec_radix, tn.rec_waves, q));
            ck(launch_rec_last(d_cells, ws.bufB, t, nb, tn.rec_radix, tn.rec_waves, q));
        }
    };
    if (!graphs) return body(s);
    const GraphKey key{{3u | ((uint64_t)(uintptr_t)d_cells << 2), nb | ((uint64_t)blob_stride << 32), tuning_word(tn),
                        (uint64_t)(uintptr_t)d_proofs, (uint64_t)(uintptr_t)d_blobs,
                        (uint64_t)(uintptr_t)&ws}};
    enqueue_maybe_graph(p, *graphs, key, s, off_clock, body);
}

// Host-buffer staging of the synchronous entry points.  A large copy is
// split over kCopyThreads host threads (one host thread copies pageable
// memory at about a third of what the PCIe DMA takes); an upload also sends
// each thread's part by DMA in kCopySlice slices as soon as a slice is
// staged, so the copy of one slice overlaps the DMA of the previous one.
// Small copies stay on the calling thread.
constexpr unsigned kCopyThreads = 4;
constexpr size_t kCopySlice = (size_t)2 << 20;
constexpr size_t kThreadedCopyMin = (size_t)4 << 20;

void parallel_copy(uint8_t *dst, const uint8_t *src, size_t bytes) {
    if (bytes < kThreadedCopyMin) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const size_t part = (bytes + kCopyThreads - 1) / kCopyThreads;
    const auto copy_part = [=](size_t off) { std::memcpy(dst + off, src + off, std::min(part, bytes - off)); };
    std::thread helpers[kCopyThreads - 1];
    for (unsigned k = 1; k < kCopyThreads; k++) {
        const size_t off = k * part;
        if (off >= bytes) break;
        try {
            helpers[k - 1] = std::thread(copy_part, off);
        } catch (const std::system_error &) {
            copy_part(off);  // no thread to spare: copy it here
        }
    }
    copy_part(0);
    for (std::thread &t : helpers)
        if (t.joinable()) t.join();
}

// host (pageable) -> pinned staging -> dev, enqueued on s.
void stage_upload(int device, uint8_t *dev, uint8_t *pinned, const uint8_t *host, size_t bytes, hipStream_t s) {
    if (bytes < kThreadedCopyMin) {
        std::memcpy(pinned, host, bytes);
        ck(hipMemcpyAsync(dev, pinned, bytes, hipMemcpyHostToDevice, s));
        return;
    }
    // parts in whole slices, so every slice but the last is full
    const size_t slices = (bytes + kCopySlice - 1) / kCopySlice;
    const size_t per = (slices + kCopyThreads - 1) / kCopyThreads;
    std::atomic<bool> failed{false};
    const auto work = [&](size_t first, bool own_thread) {
        if (own_thread && hipSetDevice(device) != hipSuccess) {
            failed = true;
            return;
        }
        for (size_t k = first; k < std::min(slices, first + per); k++) {
            const size_t off = k * kCopySlice, n = std::min(kCopySlice, bytes - off);
            std::memcpy(pinned + off, host + off, n);
            if (hipMemcpyAsync(dev + off, pinned + off, n, hipMemcpyHostToDevice, s) != hipSuccess) {
                failed = true;
                return;
            }
        }
    };
    std::thread helpers[kCopyThreads - 1];
    for (unsigned t = 1; t < kCopyThreads; t++) {
        if (t * per >= slices) break;
        try {
            helpers[t - 1] = std::thread(work, t * per, true);
        } catch (const std::system_error &) {
            work(t * per, false);  // no thread to spare: stage it here
        }
    }
    work(0, false);
    for (std::thread &t : helpers)
        if (t.joinable()) t.join();
    if (failed) throw HipFail{};
}

bool chunk_missing_ok(const uint8_t *present, uint32_t nb) {
    for (uint32_t b = 0; b < nb; b++) {
        uint32_t missing = 0;
        for (uint32_t i = 0; i < 128; i++) missing += present[(size_t)b * 128 + i] == 0;
        if (missing > 64) return false;
    }
    return true;
}


// Autotuner: for each batch-size class, time the proof pipeline on synthetic
// blobs of the class's largest size under candidate tunings and keep the
// fastest, one knob at a time (phase-A window groups, persistent grid,
// normalisation chunk, sub-streams).  Uses the prover's own working set;
// the caller holds p->mu (or owns p exclusively).
void autotune(vkzg_prover *p) {
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = p->stream;
    // Random canonical blobs: 32-byte big-endian elements below 2^254 < r.
    const uint32_t nmax = p->max_batch;
    uint64_t x = 0x9E3779B97F4A7C15ull;
    for (size_t k = 0; k < (size_t)nmax * kBlobBytes; k += 8) {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        std::memcpy(p->h_in + k, &x, 8);
    }
    for (size_t e = 0; e < (size_t)nmax * 4096; e++) p->h_in[32 * e] &= 0x3f;
    ck(hipMemcpyAsync(p->ws.blobs, p->h_in, (size_t)nmax * kBlobBytes, hipMemcpyHostToDevice, s));
    ck(hipStreamSynchronize(s));
    hipEvent_t e0, e1;
    ck(hipEventCreate(&e0));
    ck(hipEventCreate(&e1));
    auto time_one = [&](uint32_t nb, const Tuning &tn) {
        StageClock clk(p, s, false);
        float best = 1e30f;
        for (int rep = 0; rep < 3; rep++) {
            ck(hipEventRecord(e0, s));
            enqueue_proofs(p, p->ws, p->streams, p->ws.proofs, p->ws.blobs, nb, s, clk, tn, nullptr);
            ck(hipEventRecord(e1, s));
            ck(hipEventSynchronize(e1));
            float ms = 0;
            ck(hipEventElapsedTime(&ms, e0, e1));
            if (rep > 0 && ms < best) best = ms;  // rep 0 warms up
        }
        return best;
    };
    // Candidate smaller phase-A windows (c - 2 and c - 4 of the main table, within
    // 10..13): built now when they fit in half the free memory, kept only if
    // some class picks one.
    std::vector<uint32_t> windows;
    for (uint32_t w : {p->c - 2, p->c - 4}) {
        if (w < kExtraWindowMin || w > kExtraWindowMax || w >= p->c) continue;
        PhaseATable &slot = p->extra_tables[w - kExtraWindowMin];
        if (!slot.table) {
            size_t free_b = 0, total_b = 0;
            ck(hipMemGetInfo(&free_b, &total_b));
            if (table_bytes_for(w) > free_b / 2) continue;
            build_phase_a_table(p, slot, window_plan(w, 0));
        }
        windows.push_back(w);
    }
    for (int c = 0; c < kTuneClasses; c++) {
        const uint32_t nb = std::min<uint32_t>(kTuneClassTop[c], nmax);
        if (c > 0 && kTuneClassTop[c - 1] >= nmax) {
            std::lock_guard<std::mutex> lock(p->tune_mu);
            p->tune[c] = p->tune[c - 1];  // class beyond the batch size: never used alone
            continue;
        }
        // A class below the largest is also timed at half its top size (32
        // blobs for the 17-64 class), so that a shape which suits only the
        // top size (one window group at 64 blobs, half the SIMDs idle at 32)
        // does not take the whole class; the largest class is timed at its
        // top only.
        const uint32_t half = nb / 2;
        const bool two = c + 1 < kTuneClasses && c > 0 && half > kTuneClassTop[c - 1];
        auto time_class = [&](const Tuning &tn) { return time_one(nb, tn) + (two ? time_one(half, tn) : 0.0f); };
        Tuning best = p->tune[c];
        float best_ms = time_class(best);
        auto consider = [&](Tuning cand) {
            if (!tuning_valid(cand)) return;
            const float ms = time_class(cand);
            if (ms < best_ms * 0.98f) {
                best_ms = ms;
                best = cand;
            }
        };
        for (uint32_t g : {1u, 2u, 4u, 5u, 10u, 20u}) { Tuning t = best; t.groups = g; consider(t); }
        if (nb <= 64)
            for (uint32_t m : {2u, 4u}) { Tuning t = best; t.ladder_splits = m; consider(t); }
        for (uint32_t r : {4u, 16u, 32u, 64u}) { Tuning t = best; t.reduce_lanes = r; consider(t); }
        for (uint32_t b : {1u, 2u}) { Tuning t = best; t.blocks_per_cu = b; consider(t); }
        for (uint32_t ch : {0u, 8u, 16u, 32u, 64u}) { Tuning t = best; t.norm_chunk = ch; consider(t); }
        if (nb >= 8) { Tuning t = best; t.splits = 2; consider(t); }
        for (uint32_t w : windows) { Tuning t = best; t.window_bits = w; consider(t); }
        { Tuning t = best; t.window_bits = 0; consider(t); }
        if (best.window_bits) p->extra_named[best.window_bits - kExtraWindowMin] = true;
        std::lock_guard<std::mutex> lock(p->tune_mu);
        p->tune[c] = best;
    }
    // Smaller tables no tuning has ever named go (a named one may be in use by
    // a tuning another thread has copied, so it stays until the prover goes).
    for (uint32_t w : windows) {
        PhaseATable &slot = p->extra_tables[w - kExtraWindowMin];
        if (!p->extra_named[w - kExtraWindowMin] && slot.table) {
            p->arena.dfree(slot.table);
            slot = PhaseATable{};
        }
    }
    // Recovery: NTT sub-transforms per pass block, on random cells with the
    // first 64 of 128 present.
    for (size_t k = 0; k < (size_t)nmax * kCellsBytes; k += 8) {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        std::memcpy(p->h_in + k, &x, 8);
    }
    for (uint32_t b = 0; b < nmax; b++)
        for (uint32_t c = 0; c < 128; c++) p->h_present[128 * b + c] = c < 64 ? 1 : 0;
    ck(hipMemcpyAsync(p->ws.cells, p->h_in, (size_t)nmax * kCellsBytes, hipMemcpyHostToDevice, s));
    ck(hipMemcpyAsync(p->ws.present, p->h_present, (size_t)nmax * 128, hipMemcpyHostToDevice, s));
    auto time_rec = [&](uint32_t nb, const Tuning &tn) {
        StageClock clk(p, s, false);
        float best = 1e30f;
        for (int rep = 0; rep < 4; rep++) {
            ck(hipEventRecord(e0, s));
            enqueue_recovery(p, p->ws, p->streams, p->ws.cells_out, p->ws.cells, p->ws.present, nb, s,
                             clk, tn, nullptr);
            ck(hipEventRecord(e1, s));
            ck(hipEventSynchronize(e1));
            float ms = 0;
            ck(hipEventElapsedTime(&ms, e0, e1));
            if (rep > 0 && ms < best) best = ms;
        }
        return best;
    };
    for (int c = 0; c < kTuneClasses; c++) {
        const uint32_t nb = std::min<uint32_t>(kTuneClassTop[c], nmax);
        if (c > 0 && kTuneClassTop[c - 1] >= nmax) {
            std::lock_guard<std::mutex> lock(p->tune_mu);
            p->tune[c].rec_waves = p->tune[c - 1].rec_waves;
            p->tune[c].rec_radix = p->tune[c - 1].rec_radix;
            continue;
        }
        Tuning t = p->tune[c];
        float best_ms = 1e30f;
        uint32_t best_waves = t.rec_waves, best_radix = t.rec_radix;
        for (uint32_t radix : {2u, 4u, 8u}) {
            for (uint32_t waves : {1u, 2u, 4u}) {
                t.rec_radix = radix;
                t.rec_waves = waves;
                const float ms = time_rec(nb, t);
                if (ms < best_ms * 0.98f) {
                    best_ms = ms;
                    best_waves = waves;
                    best_radix = radix;
                }
            }
        }
        std::lock_guard<std::mutex> lock(p->tune_mu);
        p->tune[c].rec_waves = best_waves;
        p->tune[c].rec_radix = best_radix;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    p->autotuned = true;
    p->autotune_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

vkzg_result create(vkzg_prover **out, const vkzg_options *opts, const uint8_t *setup) {
    const auto t0 = std::chrono::steady_clock::now();
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return VKZG_ERR_GPU;
    vkzg_prover *p = new (std::nothrow) vkzg_prover();
    if (!p) return VKZG_ERR_MALLOC;
    try {
        ck(hipGetDevice(&p->device));
        noalloc_scratch_setup(p->device);
        hipDeviceProp_t prop;
        ck(hipGetDeviceProperties(&prop, p->device));
        p->name = prop.name;
        if (p->name.empty()) {
            // Some driver stacks leave the marketing name blank; fall back to the ISA.
            std::string arch = prop.gcnArchName;
            p->name = "AMD " + arch.substr(0, arch.find(':')) + " (" +
                      std::to_string(prop.multiProcessorCount) + " CUs)";
        }
        p->cus = (uint32_t)prop.multiProcessorCount;
        if (const char *env = std::getenv("VKZG_AMD_SPLITS")) p->forced_splits = std::atoi(env);
        if (const char *env = std::getenv("VKZG_AMD_GRAPHS")) p->use_graphs = std::atoi(env) != 0;
        ck(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
        p->streams.create();
        for (auto &e : p->ev) ck(hipEventCreate(&e));
        p->max_batch = opts && opts->max_batch_size ? opts->max_batch_size : kDefaultBatch;
        default_tunings(p);
        build_tables(p, setup);
        p->ws.allocate(p->max_batch);
        const size_t B = p->max_batch;
        p->h_in = p->arena.halloc<uint8_t>(B * kCellsBytes);
        p->h_out = p->arena.halloc<uint8_t>(B * kCellsBytes);
        p->h_present = p->arena.halloc<uint8_t>(B * 128);
        p->h_err = p->arena.halloc<uint32_t>(B);
        ck(hipStreamSynchronize(p->stream));
        if (const char *env = std::getenv("VKZG_AMD_AUTOTUNE"))
            if (std::atoi(env) != 0) autotune(p);
    } catch (HipFail &) {
        delete p;
        return VKZG_ERR_GPU;
    } catch (vkzg_result r) {
        delete p;
        return r;
    }
    p->setup_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = p;
    return VKZG_OK;
}

void reset_stage_times(vkzg_prover *p) {
    for (float &v : p->stage_ms) v = 0;
    p->nstages = 0;
}

}  // namespace vkzg_amd

vkzg_prover::~vkzg_prover() {
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    (void)hipDeviceSynchronize();
    vkzg_amd::destroy_verify_state(vfy);
    graphs.clear();
    ws.arena.release();
    arena.release();
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    streams.destroy();
    if (stream) (void)hipStreamDestroy(stream);
    (void)hipSetDevice(cur);
}

// ------------------------------------------------------------------ C ABI

extern "C" {

const char *vkzg_error_string(vkzg_result r) {
    switch (r) {
        case VKZG_OK: return "ok";
        case VKZG_ERR_BADARGS: return "invalid argument";
        case VKZG_ERR_MALLOC: return "allocation failed";
        case VKZG_ERR_IO: return "i/o error";
        case VKZG_ERR_SETUP: return "malformed trusted setup";
        case VKZG_ERR_GPU: return "gpu error";
        case VKZG_ERR_INVALID_BLOB: return "blob contains a non-canonical field element";
    }
    return "unknown error";
}

void vkzg_options_default(vkzg_options *opts) {
    if (!opts) return;
    opts->max_batch_size = 0;
}

vkzg_result vkzg_prover_new(vkzg_prover **out, const vkzg_options *opts) {
    if (!out) return VKZG_ERR_BADARGS;
    *out = nullptr;
    if ((size_t)(vkzg_amd_mainnet_g1_end - vkzg_amd_mainnet_g1) != kSetupBytes) return VKZG_ERR_SETUP;
    return create(out, opts, vkzg_amd_mainnet_g1);
}

vkzg_result vkzg_amd_prover_new_with_setup(vkzg_prover **out, const vkzg_options *opts,
                                           const uint8_t *g1, size_t len) {
    if (!out) return VKZG_ERR_BADARGS;
    *out = nullptr;
    if (!g1 || len != kSetupBytes) return VKZG_ERR_BADARGS;
    return create(out, opts, g1);
}

void vkzg_prover_free(vkzg_prover *p) { delete p; }

const char *vkzg_prover_device_name(const vkzg_prover *p) { return p ? p->name.c_str() : ""; }

uint32_t vkzg_prover_gpu_core_count(const vkzg_prover *p) { return p ? p->cus : 0; }

vkzg_result vkzg_compute_proofs_batch(vkzg_prover *p, uint8_t *proofs, const uint8_t *blobs,
                                      size_t num_blobs) {
    if (!p || !blobs || !proofs) return VKZG_ERR_BADARGS;
    if (num_blobs == 0) return VKZG_OK;
    std::lock_guard<std::mutex> lock(p->mu);
    DeviceGuard guard(p->device);
    reset_stage_times(p);
    try {
        for (size_t done = 0; done < num_blobs;) {
            const uint32_t nb = (uint32_t)std::min<size_t>(p->max_batch, num_blobs - done);
            hipStream_t s = p->stream;
            stage_upload(p->device, p->ws.blobs, p->h_in, blobs + done * kBlobBytes, nb * kBlobBytes, s);
            StageClock clk(p, s, true);
            const Tuning tn = tuning_for(p, nb);
            p->last_splits = tn.splits;
            enqueue_proofs(p, p->ws, p->streams, p->ws.proofs, p->ws.blobs, nb, s, clk, tn,
                           &p->graphs);
            ck(hipMemcpyAsync(p->h_err, p->ws.err, nb * 4, hipMemcpyDeviceToHost, s));
            ck(hipMemcpyAsync(p->h_out, p->ws.proofs, nb * kProofBytes, hipMemcpyDeviceToHost, s));
            ck(hipStreamSynchronize(s));
            clk.finish();
            for (uint32_t b = 0; b < nb; b++)
                if (p->h_err[b]) return VKZG_ERR_INVALID_BLOB;
            std::memcpy(proofs + done * kProofBytes, p->h_out, nb * kProofBytes);
            done += nb;
        }
    } catch (HipFail &) {
        return VKZG_ERR_GPU;
    }
    return VKZG_OK;
}

vkzg_result vkzg_recover_cells_batch(vkzg_prover *p, uint8_t *cells_out, const uint8_t *cells,
                                     const uint8_t *cell_present, size_t num_blobs) {
    if (!p || !cells_out || !cells || !cell_present) return VKZG_ERR_BADARGS;
    if (num_blobs == 0) return VKZG_OK;
    std::lock_guard<std::mutex> lock(p->mu);
    DeviceGuard guard(p->device);
    reset_stage_times(p);
    try {
        for (size_t done = 0; done < num_blobs;) {
            const uint32_t nb = (uint32_t)std::min<size_t>(p->max_batch, num_blobs - done);
            const uint8_t *pm = cell_present + done * 128;
            if (!chunk_missing_ok(pm, nb)) return VKZG_ERR_BADARGS;
            hipStream_t s = p->stream;
            stage_upload(p->device, p->ws.cells, p->h_in, cells + done * kCellsBytes, nb * kCellsBytes, s);
            std::memcpy(p->h_present, pm, nb * 128);
            ck(hipMemcpyAsync(p->ws.present, p->h_present, nb * 128, hipMemcpyHostToDevice, s));
            StageClock clk(p, s, true);
            const Tuning tn = tuning_for(p, nb);
            p->last_splits = tn.splits;
            enqueue_recovery(p, p->ws, p->streams, p->ws.cells_out, p->ws.cells, p->ws.present, nb,
                             s, clk, tn, &p->graphs);
            ck(hipMemcpyAsync(p->h_out, p->ws.cells_out, nb * kCellsBytes, hipMemcpyDeviceToHost, s));
            ck(hipStreamSynchronize(s));
            clk.finish();
            parallel_copy(cells_out + done * kCellsBytes, p->h_out, nb * kCellsBytes);
            done += nb;
        }
    } catch (HipFail &) {
        return VKZG_ERR_GPU;
    }
    return VKZG_OK;
}

vkzg_result vkzg_amd_compute_cells_and_proofs(vkzg_prover *p, uint8_t *cells, uint8_t *proofs,
                                              const uint8_t *blobs, size_t num_blobs) {
    if (!p || !blobs || (!cells && !proofs)) return VKZG_ERR_BADARGS;
    if (num_blobs == 0) return VKZG_OK;
    std::lock_guard<std::mutex> lock(p->mu);
    DeviceGuard guard(p->device);
    reset_stage_times(p);
    try {
        for (size_t done = 0; done < num_blobs;) {
            const uint32_t nb = (uint32_t)std::min<size_t>(p->max_batch, num_blobs - done);
            hipStream_t s = p->stream;
            stage_upload(p->device, p->ws.blobs, p->h_in, blobs + done * kBlobBytes, nb * kBlobBytes, s);
            const Tuning tn = tuning_for(p, nb);
            Tuning one = tn;
            one.splits = 1;
            enqueue_cells_and_proofs(p, p->ws, cells ? p->ws.cells_out : nullptr, p->ws.proofs,
                                     p->ws.blobs, kBlobBytes, nb, s, one, &p->graphs);
            ck(hipMemcpyAsync(p->h_err, p->ws.err, nb * 4, hipMemcpyDeviceToHost, s));
            if (cells)
                ck(hipMemcpyAsync(p->h_out, p->ws.cells_out, nb * kCellsBytes, hipMemcpyDeviceToHost, s));
            if (proofs)  // h_in's blobs were consumed by the copy above, in stream order
                ck(hipMemcpyAsync(p->h_in, p->ws.proofs, nb * kProofBytes, hipMemcpyDeviceToHost, s));
            ck(hipStreamSynchronize(s));
            for (uint32_t b = 0; b < nb; b++)
                if (p->h_err[b]) return VKZG_ERR_INVALID_BLOB;
            if (cells) parallel_copy(cells + done * kCellsBytes, p->h_out, nb * kCellsBytes);
            if (proofs) std::memcpy(proofs + done * kProofBytes, p->h_in, nb * kProofBytes);
            done += nb;
        }
    } catch (HipFail &) {
        return VKZG_ERR_GPU;
    }
    return VKZG_OK;
}

vkzg_result vkzg_amd_recover_cells_and_proofs(vkzg_prover *p, uint8_t *cells_out, uint8_t *proofs_out,
                                              const uint8_t *cells, const uint8_t *cell_present,
                                              size_t num_blobs) {
    if (!p || !cells_out || !proofs_out || !cells || !cell_present) return VKZG_ERR_BADARGS;
    if (num_blobs == 0) return VKZG_OK;
    std::lock_guard<std::mutex> lock(p->mu);
    DeviceGuard guard(p->device);
    reset_stage_times(p);
    try {
        for (size_t done = 0; done < num_blobs;) {
            const uint32_t nb = (uint32_t)std::min<size_t>(p->max_batch, num_blobs - done);
            const uint8_t *pm = cell_present + done * 128;
            if (!chunk_missing_ok(pm, nb)) return VKZG_ERR_BADARGS;
            hipStream_t s = p->stream;
            stage_upload(p->device, p->ws.cells, p->h_in, cells + done * kCellsBytes, nb * kCellsBytes, s);
            std::memcpy(p->h_present, pm, nb * 128);
            ck(hipMemcpyAsync(p->ws.present, p->h_present, nb * 128, hipMemcpyHostToDevice, s));
            const Tuning tn = tuning_for(p, nb);
            Tuning one = tn;
            one.splits = 1;
            StageClock clk(p, s, false);
            enqueue_recovery(p, p->ws, p->streams, p->ws.cells_out, p->ws.cells, p->ws.present, nb, s,
                             clk, one, &p->graphs);
            // The first 64 recovered cells of a blob are the blob itself (cell c
            // holds p(w^brev13(64c+t)) and brev13(i) = 2 brev12(i) for i < 4096).
            enqueue_cells_and_proofs(p, p->ws, nullptr, p->ws.proofs, p->ws.cells_out, kCellsBytes, nb,
                                     s, one, &p->graphs);
            ck(hipMemcpyAsync(p->h_out, p->ws.cells_out, nb * kCellsBytes, hipMemcpyDeviceToHost, s));
            ck(hipMemcpyAsync(p->h_in, p->ws.proofs, nb * kProofBytes, hipMemcpyDeviceToHost, s));
            ck(hipStreamSynchronize(s));
            parallel_copy(cells_out + done * kCellsBytes, p->h_out, nb * kCellsBytes);
            std::memcpy(proofs_out + done * kProofBytes, p->h_in, nb * kProofBytes);
            done += nb;
        }
    } catch (HipFail &) {
        return VKZG_ERR_GPU;
    }
    return VKZG_OK;
}

vkzg_result vkzg_amd_compute_proofs_device(vkzg_prover *p, void *d_proofs, const void *d_blobs,
                                           size_t num_blobs, void *hip_stream) {
    if (!p || !d_proofs || !d_blobs) return VKZG_ERR_BADARGS;
    if (num_blobs == 0) return VKZG_OK;
    std::lock_guard<std::mutex> lock(p->mu);
    DeviceGuard guard(p->device);
    reset_stage_times(p);
    hipStream_t s = hip_stream ? (hipStream_t)hip_stream : p->stream;
    try {
        for (size_t done = 0; done < num_blobs;) {
            const uint32_t nb = (uint32_t)std::min<size_t>(p->max_batch, num_blobs - done);
            StageClock clk(p, s, true);
            const Tuning tn = tuning_for(p, nb);
            p->last_splits = tn.splits;
            enqueue_proofs(p, p->ws, p->streams, (uint8_t *)d_proofs + done * kProofBytes,
                           (const uint8_t *)d_blobs + done * kBlobBytes, nb, s, clk, tn,
                           &p->graphs);
            ck(hipMemcpyAsync(p->h_err, p->ws.err, nb * 4, hipMemcpyDeviceToHost, s));
            ck(hipStreamSynchronize(s));
            clk.finish();
            for (uint32_t b = 0; b < nb; b++)
                if (p->h_err[b]) return VKZG_ERR_INVALID_BLOB;
            done += nb;
        }
    } catch (HipFail &) {
        return VKZG_ERR_GPU;
    }
    return VKZG_OK;
}

vkzg_result vkzg_amd_recover_cells_device(vkzg_prover *p, void *d_cells_out, const void *d_cells,
                                          const void *d_present, const uint8_t *h_present,
                                          size_t num_blobs, void *hip_stream) {
    if (!p || !d_cells_out || !d_cells || !d_present || !h_present) return VKZG_ERR_BADARGS;
    if (num_blobs == 0) return VKZG_OK;
    if (!aligned16(d_cells_out) || !aligned16(d_cells) || !aligned16(d_present)) return VKZG_ERR_BADARGS;
    std::lock_guard<std::mutex> lock(p->mu);
    DeviceGuard guard(p->device);
    reset_stage_times(p);
    hipStream_t s = hip_stream ? (hipStream_t)hip_stream : p->stream;
    try {
        for (size_t done = 0; done < num_blobs;) {
            const uint32_t nb = (uint32_t)std::min<size_t>(p->max_batch, num_blobs - done);
            if (!chunk_missing_ok(h_present + done * 128, nb)) return VKZG_ERR_BADARGS;
            StageClock clk(p, s, true);
            const Tuning tn = tuning_for(p, nb);
            p->last_splits = tn.splits;
            enqueue_recovery(p, p->ws, p->streams, (uint8_t *)d_cells_out + done * kCellsBytes,
                             (const uint8_t *)d_cells + done * kCellsBytes,
                             (const uint8_t *)d_present + done * 128, nb, s, clk, tn, &p->graphs);
            ck(hipStreamSynchronize(s));
            clk.finish();
            done += nb;
        }
    } catch (HipFail &) {
        return VKZG_ERR_GPU;
    }
    return VKZG_OK;
}

vkzg_result vkzg_amd_get_tuning(const vkzg_prover *p, size_t num_blobs, vkzg_amd_tuning *out) {
    if (!p || !out || num_blobs == 0) return VKZG_ERR_BADARGS;
    Tuning t;
    {
        std::lock_guard<std::mutex> lock(p->tune_mu);
        t = p->tune[tune_class((uint32_t)std::min<size_t>(num_blobs, 0xffffffffu))];
    }
    out->phase_a_groups = t.groups;
    out->phase_a_blocks_per_cu = t.blocks_per_cu;
    out->normalize_chunk = t.norm_chunk;
    out->splits = t.splits;
    out->recovery_waves = t.rec_waves;
    out->recovery_radix = t.rec_radix;
    out->ladder_splits = t.ladder_splits;
    out->reduce_lanes = t.reduce_lanes;
    out->phase_a_window_bits = t.window_bits;
    return VKZG_OK;
}

vkzg_result vkzg_amd_set_tuning(vkzg_prover *p, size_t num_blobs, const vkzg_amd_tuning *in) {
    if (!p || !in || num_blobs == 0) return VKZG_ERR_BADARGS;
    Tuning t;
    t.groups = in->phase_a_groups;
    t.blocks_per_cu = in->phase_a_blocks_per_cu;
    t.norm_chunk = in->normalize_chunk;
    t.splits = in->splits;
    t.rec_waves = in->recovery_waves;
    t.rec_radix = in->recovery_radix;
    t.ladder_splits = in->ladder_splits;
    t.reduce_lanes = in->reduce_lanes;
    t.window_bits = in->phase_a_window_bits == p->c ? 0u : in->phase_a_window_bits;
    if (!tuning_valid(t)) return VKZG_ERR_BADARGS;
    std::lock_guard<std::mutex> lock(p->mu);
    if (t.window_bits) {  // only a resident smaller table
        if (!p->extra_tables[t.window_bits - kExtraWindowMin].table) return VKZG_ERR_BADARGS;
        p->extra_named[t.window_bits - kExtraWindowMin] = true;
    }
    std::lock_guard<std::mutex> tlock(p->tune_mu);
    p->tune[tune_class((uint32_t)std::min<size_t>(num_blobs, 0xffffffffu))] = t;
    return VKZG_OK;
}

vkzg_result vkzg_amd_load_window_table(vkzg_prover *p, uint32_t window_bits) {
    if (!p || window_bits < kExtraWindowMin || window_bits > kExtraWindowMax) return VKZG_ERR_BADARGS;
    std::lock_guard<std::mutex> lock(p->mu);
    if (window_bits >= p->c) return VKZG_ERR_BADARGS;
    DeviceGuard guard(p->device);
    PhaseATable &slot = p->extra_tables[window_bits - kExtraWindowMin];
    if (slot.table) {
        p->extra_named[window_bits - kExtraWindowMin] = true;
        return VKZG_OK;
    }
    try {
        size_t free_b = 0, total_b = 0;
        ck(hipMemGetInfo(&free_b, &total_b));
        if (table_bytes_for(window_bits) > free_b) return VKZG_ERR_MALLOC;
        build_phase_a_table(p, slot, window_plan(window_bits, 0));
    } catch (HipFail &) {
        if (slot.table) p->arena.dfree(slot.table);
        slot = PhaseATable{};
        return VKZG_ERR_GPU;
    }
    p->extra_named[window_bits - kExtraWindowMin] = true;
    return VKZG_OK;
}

vkzg_result vkzg_amd_set_graphs(vkzg_prover *p, int enable) {
    if (!p) return VKZG_ERR_BADARGS;
    std::lock_guard<std::mutex> lock(p->mu);
    p->use_graphs = enable != 0;
    return VKZG_OK;
}

vkzg_result vkzg_amd_debug_force_checked(vkzg_prover *p, int enable) {
    if (!p) return VKZG_ERR_BADARGS;
    p->force_checked.store(enable ? 1u : 0u);
    return VKZG_OK;
}

vkzg_result vkzg_amd_autotune(vkzg_prover *p, double *ms_out) {
    if (!p) return VKZG_ERR_BADARGS;
    std::lock_guard<std::mutex> lock(p->mu);
    DeviceGuard guard(p->device);
    try {
        autotune(p);
    } catch (HipFail &) {
        return VKZG_ERR_GPU;
    }
    if (ms_out) *ms_out = p->autotune_ms;
    return VKZG_OK;
}

vkzg_result vkzg_amd_prover_info(const vkzg_prover *p, vkzg_amd_info *out) {
    if (!p || !out) return VKZG_ERR_BADARGS;
    out->window_bits = p->c;
    out->windows = p->W;
    out->table_bytes = p->table_bytes;
    out->resident_table_bytes = p->table_bytes;
    out->extra_table_windows = 0;
    for (uint32_t i = 0; i <= kExtraWindowMax - kExtraWindowMin; i++)
        if (p->extra_tables[i].table) {
            out->resident_table_bytes += p->extra_tables[i].bytes;
            out->extra_table_windows |= 1u << (kExtraWindowMin + i);
        }
    out->working_bytes = WorkSet::bytes_per_blob() * p->max_batch;
    out->last_splits = p->last_splits;
    out->graphs_enabled = p->use_graphs ? 1u : 0u;
    out->graph_captures = p->graphs.captures;
    out->graph_replays = p->graphs.replays;
    out->autotuned = p->autotuned ? 1u : 0u;
    out->autotune_ms = p->autotune_ms;
    out->phase_b_items = p->kb_per_blob;
    out->max_batch_size = p->max_batch;
    out->setup_ms = p->setup_ms;
    return VKZG_OK;
}

vkzg_result vkzg_amd_set_stage_timing(vkzg_prover *p, int enable) {
    if (!p) return VKZG_ERR_BADARGS;
    std::lock_guard<std::mutex> lock(p->mu);
    p->timing = enable != 0;
    return VKZG_OK;
}

int vkzg_amd_last_stage_times(vkzg_prover *p, float *ms, const char **names, int max) {
    if (!p) return 0;
    std::lock_guard<std::mutex> lock(p->mu);
    const int n = p->nstages < max ? p->nstages : max;
    for (int i = 0; i < n; i++) {
        if (ms) ms[i] = p->stage_ms[i];
        if (names) names[i] = p->stage_names[i];
    }
    return n;
}

// op 0: Fp product, 1: Fr product, 2: Fp inverse of a (b unused).
static vkzg_result debug_op(vkzg_prover *p, uint32_t *out, const uint32_t *a, const uint32_t *b,
                            size_t n, int op) {
    const int limbs = op == 1 ? 8 : 12;
    if (!out || !a || !b) return VKZG_ERR_BADARGS;
    if (n == 0) return VKZG_OK;
    // A prover is optional: without one the current device and its null
    // stream are used.
    std::unique_lock<std::mutex> lock;
    if (p) lock = std::unique_lock<std::mutex>(p->mu);
    int dev = 0;
    if (p) dev = p->device;
    else if (hipGetDevice(&dev) != hipSuccess) return VKZG_ERR_GPU;
    DeviceGuard guard(dev);
    hipStream_t stream = p ? p->stream : nullptr;
    void *da = nullptr, *db = nullptr, *dc = nullptr;
    const size_t bytes = n * limbs * 4;
    vkzg_result rc = VKZG_OK;
    try {
        ck(hipMalloc(&da, bytes));
        ck(hipMalloc(&db, bytes));
        ck(hipMalloc(&dc, bytes));
        ck(hipMemcpyAsync(da, a, bytes, hipMemcpyHostToDevice, stream));
        ck(hipMemcpyAsync(db, b, bytes, hipMemcpyHostToDevice, stream));
        if (op == 0)
            ck(launch_debug_fp_mul((uint32_t *)dc, (const uint32_t *)da, (const uint32_t *)db,
                                   (uint32_t)n, stream));
        else if (op == 2)
            ck(launch_debug_fp_inv((uint32_t *)dc, (const uint32_t *)da, (uint32_t)n, stream));
        else
            ck(launch_debug_fr_mul((uint32_t *)dc, (const uint32_t *)da, (const uint32_t *)db,
                                   (uint32_t)n, stream));
        ck(hipMemcpyAsync(out, dc, bytes, hipMemcpyDeviceToHost, stream));
        ck(hipStreamSynchronize(stream));
    } catch (HipFail &) {
        rc = VKZG_ERR_GPU;
    }
    (void)hipFree(da);
    (void)hipFree(db);
    (void)hipFree(dc);
    return rc;
}

vkzg_result vkzg_amd_debug_fp_mul(vkzg_prover *p, uint32_t *out, const uint32_t *a,
                                  const uint32_t *b, size_t n) {
    return debug_op(p, out, a, b, n, 0);
}

vkzg_result vkzg_amd_debug_fp_mul_rate(vkzg_prover *p, uint32_t waves_per_simd, uint32_t iters,
                                       double *muls_per_s, double *one_wave_cycles) {
    if (!p || !muls_per_s || !one_wave_cycles || waves_per_simd == 0 || waves_per_simd > 8 || iters == 0)
        return VKZG_ERR_BADARGS;
    std::lock_guard<std::mutex> lock(p->mu);
    DeviceGuard guard(p->device);
    void *buf = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    vkzg_result rc = VKZG_OK;
    try {
        ck(hipMalloc(&buf, 1024 * sizeof(Fp)));
        ck(hipEventCreate(&e0));
        ck(hipEventCreate(&e1));
        hipStream_t s = p->stream;
        int clock_khz = 0;
        ck(hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeClockRate, p->device));
        const uint32_t blocks = p->cus * 4 * waves_per_simd;  // one wave per block, 4 SIMDs per CU
        float ms[2] = {0, 0};
        for (int pass = 0; pass < 2; pass++) {
            const uint32_t nb = pass == 0 ? blocks : 1;
            ck(launch_debug_fp_mul_chain(static_cast<Fp *>(buf), nb, iters, s));  // warm-up
            ck(hipEventRecord(e0, s));
            ck(launch_debug_fp_mul_chain(static_cast<Fp *>(buf), nb, iters, s));
            ck(hipEventRecord(e1, s));
            ck(hipEventSynchronize(e1));
            ck(hipEventElapsedTime(&ms[pass], e0, e1));
        }
        *muls_per_s = (double)blocks * 64.0 * iters / (ms[0] * 1e-3);
        *one_wave_cycles = (double)ms[1] * 1e-3 * clock_khz * 1e3 / iters;
    } catch (HipFail &) {
        rc = VKZG_ERR_GPU;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(buf);
    return rc;
}

vkzg_result vkzg_amd_debug_mad_rate(vkzg_prover *p, uint32_t waves_per_simd, uint32_t iters,
                                    double *mads_per_s, double *clock_mhz) {
    if (!p || !mads_per_s || !clock_mhz || waves_per_simd == 0 || waves_per_simd > 8 || iters == 0)
        return VKZG_ERR_BADARGS;
    std::lock_guard<std::mutex> lock(p->mu);
    DeviceGuard guard(p->device);
    void *buf = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    vkzg_result rc = VKZG_OK;
    try {
        ck(hipMalloc(&buf, 1024 * sizeof(uint64_t)));
        ck(hipEventCreate(&e0));
        ck(hipEventCreate(&e1));
        hipStream_t s = p->stream;
        int clock_khz = 0;
        ck(hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeClockRate, p->device));
        const uint32_t blocks = p->cus * 4 * waves_per_simd;  // one wave per block, 4 SIMDs per CU
        ck(launch_debug_mad_chain(static_cast<uint64_t *>(buf), blocks, iters, s));  // warm-up
        ck(hipEventRecord(e0, s));
        ck(launch_debug_mad_chain(static_cast<uint64_t *>(buf), blocks, iters, s));
        ck(hipEventRecord(e1, s));
        ck(hipEventSynchronize(e1));
        float ms = 0;
        ck(hipEventElapsedTime(&ms, e0, e1));
        *mads_per_s = (double)blocks * 64.0 * iters * 128.0 / (ms * 1e-3);
        *clock_mhz = clock_khz / 1e3;
    } catch (HipFail &) {
        rc = VKZG_ERR_GPU;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(buf);
    return rc;
}

vkzg_result vkzg_amd_debug_fp_inv(vkzg_prover *p, uint32_t *out, const uint32_t *a, size_t n) {
    return debug_op(p, out, a, a, n, 2);
}

vkzg_result vkzg_amd_debug_fr_mul(vkzg_prover *p, uint32_t *out, const uint32_t *a,
                                  const uint32_t *b, size_t n) {
    return debug_op(p, out, a, b, n, 1);
}

}  // extern "C"
