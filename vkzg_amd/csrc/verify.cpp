// Host side of batch cell-proof verification and of commitments (C ABI in
// include/vkzg_amd.h).  The reference has neither (include/vkzg.h:14-18); the
// functions follow the consensus specs' polynomial-commitments-sampling
// (verify_cell_kzg_proof_batch) and polynomial-commitments
// (blob_to_kzg_commitment), with the argument conventions of the reference's
// entry points (src/vkzg.cpp:1183-1218: NULL checks first, n == 0 is OK).
//
// A verification call:
//   host    argument checks, commitment deduplication, grouping of cells by
//           commitment and by cell index, pinned staging, one H2D copy;
//   device  decompression of commitments and proofs and parsing of cells,
//           with the G1 subgroup checks on a second stream;
//   host    the Fiat-Shamir challenge (SHA-256 over every input), computed
//           while the device works;
//   device  r^k, interpolation sums, the two MSMs and the pairing check;
//   host    one 8-byte readback (verdict, error bits).
#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/vkzg_amd.h"
#include "msm.h"
#include "pairing.h"
#include "runtime.h"
#include "sha256.h"
#include "verify.h"

namespace vkzg_amd {

constexpr size_t kG2Bytes = 96;
constexpr size_t kG2SetupPoints = 65;  // [tau^i]G2, i <= 64
constexpr size_t kCellBytes = VKZG_BYTES_PER_CELL;
constexpr uint32_t kInterpChunk = 32;           // cells per interpolation block
constexpr size_t kMaxVerifyCells = 1u << 22;    // keeps every index below 2^30

struct VerifyState {
    bool ready = false;
    DeviceArena setup;      // lines of [tau^64]G2 and [tau^0]G2
    Fp2 *lines = nullptr;   // [2][kPairSteps][2]
    G1A *identity = nullptr;
    void *dev = nullptr;
    size_t dev_bytes = 0;
    uint8_t *host = nullptr;
    size_t host_bytes = 0;
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;

    void ensure(size_t dbytes, size_t hbytes) {
        if (dbytes > dev_bytes) {
            if (dev) (void)hipFree(dev);
            dev = nullptr;
            dev_bytes = 0;
            ck(hipMalloc(&dev, dbytes));
            dev_bytes = dbytes;
        }
        if (hbytes > host_bytes) {
            if (host) (void)hipHostFree(host);
            host = nullptr;
            host_bytes = 0;
            ck(hipHostMalloc(reinterpret_cast<void **>(&host), hbytes, hipHostMallocDefault));
            host_bytes = hbytes;
        }
    }
    void streams() {
        if (side) return;
        ck(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        ck(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
        ck(hipEventCreateWithFlags(&join, hipEventDisableTiming));
    }
    ~VerifyState() {
        if (dev) (void)hipFree(dev);
        if (host) (void)hipHostFree(host);
        if (join) (void)hipEventDestroy(join);
        if (fork) (void)hipEventDestroy(fork);
        if (side) (void)hipStreamDestroy(side);
        setup.release();
    }
};

void destroy_verify_state(VerifyState *v) { delete v; }

namespace {

// Bump layout of one allocation: every region 256-byte aligned.
struct Layout {
    size_t size = 0;
    size_t take(size_t bytes) {
        const size_t at = size;
        size += (bytes + 255) & ~size_t(255);
        return at;
    }
};

struct Key48 {
    const uint8_t *p;
    bool operator==(const Key48 &o) const { return std::memcmp(p, o.p, 48) == 0; }
};
struct Key48Hash {
    size_t operator()(const Key48 &k) const {
        uint64_t a, b;
        std::memcpy(&a, k.p + 8, 8);  // x bytes: uniformly distributed for valid points
        std::memcpy(&b, k.p + 40, 8);
        return (size_t)(a * 0x9e3779b97f4a7c15ULL ^ b);
    }
};

void put_be64(uint8_t *d, uint64_t v) {
    for (int i = 0; i < 8; i++) d[i] = (uint8_t)(v >> (56 - 8 * i));
}

// r as 8 little-endian words: SHA-256 digest as a big-endian integer mod r.
void digest_mod_r(uint32_t *r, const uint8_t *digest) {
    for (int i = 0; i < 8; i++)
        r[i] = (uint32_t)digest[31 - 4 * i] | (uint32_t)digest[30 - 4 * i] << 8 |
               (uint32_t)digest[29 - 4 * i] << 16 | (uint32_t)digest[28 - 4 * i] << 24;
    for (int pass = 0; pass < 2; pass++) {
        uint32_t t[8];
        uint64_t borrow = 0;
        for (int i = 0; i < 8; i++) {
            const uint64_t d = (uint64_t)r[i] - cfr::P[i] - borrow;
            t[i] = (uint32_t)d;
            borrow = (d >> 63) & 1u;
        }
        if (borrow) break;
        std::memcpy(r, t, sizeof(t));
    }
}

// Deduplicated commitments (first occurrence order) and each cell's index
// into them (verify_cell_kzg_proof_batch's deduplicated_commitments and
// commitment_indices).
uint32_t dedup(std::vector<uint32_t> &uniq, std::vector<uint32_t> &cidx, const uint8_t *commitments, size_t n) {
    std::unordered_map<Key48, uint32_t, Key48Hash> seen;
    seen.reserve(n * 2);
    uniq.clear();
    cidx.resize(n);
    for (size_t k = 0; k < n; k++) {
        const Key48 key{commitments + 48 * k};
        auto it = seen.find(key);
        if (it == seen.end()) {
            const uint32_t j = (uint32_t)uniq.size();
            seen.emplace(key, j);
            uniq.push_back((uint32_t)k);
            cidx[k] = j;
        } else {
            cidx[k] = it->second;
        }
    }
    return (uint32_t)uniq.size();
}

// compute_verify_cell_kzg_proof_batch_challenge: SHA-256 of the domain
// "RCKZGCBATCH__V1_", FIELD_ELEMENTS_PER_BLOB, FIELD_ELEMENTS_PER_CELL, the
// numbers of commitments and cells (8-byte big-endian each), the
// deduplicated commitments, then per cell its commitment index and cell index
// (8-byte big-endian), its 64 elements (32-byte big-endian; the cell bytes
// once they are known to be canonical) and its proof; reduced mod r.
void challenge(uint32_t *r, const std::vector<uint32_t> &uniq, const std::vector<uint32_t> &cidx,
               const uint64_t *cell_indices, const uint8_t *commitments, const uint8_t *cells,
               const uint8_t *proofs, size_t n, bool allow_hw = true) {
    sha256::Ctx c;
    sha256::init(c, allow_hw);
    uint8_t head[48];
    std::memcpy(head, "RCKZGCBATCH__V1_", 16);
    put_be64(head + 16, 4096);
    put_be64(head + 24, 64);
    put_be64(head + 32, uniq.size());
    put_be64(head + 40, n);
    sha256::update(c, head, sizeof(head));
    for (uint32_t k : uniq) sha256::update(c, commitments + 48ull * k, 48);
    for (size_t k = 0; k < n; k++) {
        uint8_t idx[16];
        put_be64(idx, cidx[k]);
        put_be64(idx + 8, cell_indices[k]);
        sha256::update(c, idx, 16);
        sha256::update(c, cells + kCellBytes * k, kCellBytes);
        sha256::update(c, proofs + 48 * k, 48);
    }
    uint8_t digest[32];
    sha256::finish(c, digest);
    digest_mod_r(r, digest);
}

// Counting sort of 0..n-1 by key (< nkeys): off[nkeys + 1], order[n].
void group_by(std::vector<uint32_t> &off, std::vector<uint32_t> &order, const uint32_t *key, size_t n,
              uint32_t nkeys) {
    off.assign(nkeys + 1, 0);
    for (size_t k = 0; k < n; k++) off[key[k] + 1]++;
    for (uint32_t i = 0; i < nkeys; i++) off[i + 1] += off[i];
    std::vector<uint32_t> cur(off.begin(), off.end() - 1);
    order.resize(n);
    for (size_t k = 0; k < n; k++) order[cur[key[k]]++] = (uint32_t)k;
}

// terms (may be null): the two G1 inputs of the pairing check, compressed
// (sum r^k pi_k, then RL), for the tests.
vkzg_result verify_impl(vkzg_prover *p, int *ok, const uint8_t *commitments, const uint64_t *cell_indices,
                        const uint8_t *cells, const uint8_t *proofs, size_t n, uint8_t *terms = nullptr) {
    VerifyState &v = *p->vfy;
    std::vector<uint32_t> uniq, cidx, cgrp_off, cgrp, iord_off, iord;
    const uint32_t u = dedup(uniq, cidx, commitments, n);
    std::vector<uint32_t> cell_idx(n);
    for (size_t k = 0; k < n; k++) cell_idx[k] = (uint32_t)cell_indices[k];
    group_by(cgrp_off, cgrp, cidx.data(), n, u);
    group_by(iord_off, iord, cell_idx.data(), n, 128);
    std::vector<uint32_t> chunks;
    for (uint32_t c = 0; c < 128; c++)
        for (uint32_t lo = iord_off[c]; lo < iord_off[c + 1]; lo += kInterpChunk) {
            chunks.push_back(c);
            chunks.push_back(lo);
            chunks.push_back(std::min(lo + kInterpChunk, iord_off[c + 1]));
        }
    const uint32_t nch = (uint32_t)(chunks.size() / 3);
    const uint32_t n32 = (uint32_t)n;
    const uint32_t ne = 2 * n32 + u + 64;
    const MsmShape sh = msm_shape(ne, 2);

    // Host staging (mirrors the start of the device layout): points, cells, index arrays.
    Layout hl;
    const size_t h_pts = hl.take(48ull * (u + n)), h_cells = hl.take(kCellBytes * n);
    const size_t h_cidx = hl.take(4ull * n), h_coff = hl.take(4ull * (u + 1)), h_cgrp = hl.take(4ull * n);
    const size_t h_iord = hl.take(4ull * n), h_chunks = hl.take(12ull * nch);
    const size_t in_bytes = hl.size;
    const size_t h_r = hl.take(32), h_res = hl.take(16), h_terms = hl.take(96);
    Layout dl;
    dl.size = in_bytes;
    const size_t d_r = dl.take(32), d_res = dl.take(16), d_terms = dl.take(96);
    const size_t d_pts = dl.take(sizeof(G1A) * (u + n + 64ull)), d_ev = dl.take(sizeof(Fr) * 64ull * n);
    const size_t d_rp = dl.take(sizeof(Fr) * (size_t)n), d_part = dl.take(sizeof(Fr) * 64ull * nch);
    const size_t d_agg = dl.take(sizeof(Fr) * 64), d_sc = dl.take(sizeof(Fr) * (size_t)ne);
    const size_t d_pidx = dl.take(4ull * ne), d_mid = dl.take(ne), d_out = dl.take(sizeof(G1A) * 2);
    const size_t d_L = dl.take(sizeof(Fp) * 12 * kPairSteps), d_msm = dl.take(msm_scratch_bytes(sh));
    v.ensure(dl.size, hl.size);
    v.streams();
    uint8_t *H = v.host;
    uint8_t *D = static_cast<uint8_t *>(v.dev);
    for (uint32_t j = 0; j < u; j++) std::memcpy(H + h_pts + 48ull * j, commitments + 48ull * uniq[j], 48);
    std::memcpy(H + h_pts + 48ull * u, proofs, 48 * n);
    std::memcpy(H + h_cells, cells, kCellBytes * n);
    std::memcpy(H + h_cidx, cell_idx.data(), 4 * n);
    std::memcpy(H + h_coff, cgrp_off.data(), 4ull * (u + 1));
    std::memcpy(H + h_cgrp, cgrp.data(), 4 * n);
    std::memcpy(H + h_iord, iord.data(), 4 * n);
    std::memcpy(H + h_chunks, chunks.data(), 12ull * nch);

    hipStream_t s = p->stream;
    auto U32 = [&](size_t off) { return reinterpret_cast<uint32_t *>(D + off); };
    uint32_t *res = U32(d_res);  // [0] verdict, [1] error bits
    G1A *pts = reinterpret_cast<G1A *>(D + d_pts);
    Fr *ev = reinterpret_cast<Fr *>(D + d_ev), *rp = reinterpret_cast<Fr *>(D + d_rp);
    Fr *agg = reinterpret_cast<Fr *>(D + d_agg), *sc = reinterpret_cast<Fr *>(D + d_sc);
    G1A *out = reinterpret_cast<G1A *>(D + d_out);
    ck(hipMemcpyAsync(D, H, in_bytes, hipMemcpyHostToDevice, s));
    ck(hipMemsetAsync(res, 0, 16, s));
    ck(launch_vfy_points(pts, D + h_pts, u + n32, p->srs, 64, res + 1, s));
    ck(launch_vfy_cells(ev, D + h_cells, n32, res + 1, s));
    // G1 membership of commitments and proofs, beside the rest of the pipeline
    ck(hipEventRecord(v.fork, s));
    ck(hipStreamWaitEvent(v.side, v.fork, 0));
    ck(launch_vfy_subgroup(pts, u + n32, res + 1, v.side));
    ck(hipEventRecord(v.join, v.side));

    uint32_t r[8];
    challenge(r, uniq, cidx, cell_indices, commitments, cells, proofs, n);
    std::memcpy(H + h_r, r, 32);
    ck(hipMemcpyAsync(D + d_r, H + h_r, 32, hipMemcpyHostToDevice, s));
    ck(launch_vfy_weights(rp, reinterpret_cast<const Fr *>(D + d_r), n32, s));
    ck(launch_vfy_interp(reinterpret_cast<Fr *>(D + d_part), agg, U32(h_chunks), nch, U32(h_iord), ev, rp,
                         p->roots_inv, s));
    ck(launch_vfy_entries(sc, U32(d_pidx), D + d_mid, rp, U32(h_coff), U32(h_cgrp), U32(h_cidx), p->roots, agg,
                          n32, u, s));
    ck(launch_msm(sh, D + d_msm, pts, U32(d_pidx), sc, D + d_mid, out, s));
    if (terms) {
        ck(launch_compress_points(D + d_terms, out, 2, s));
        ck(hipMemcpyAsync(H + h_terms, D + d_terms, 96, hipMemcpyDeviceToHost, s));
    }
    // e(sum r^k pi_k, [tau^64]G2) e(-RL, G2) == 1
    ck(launch_pairing_check(res, nullptr, reinterpret_cast<Fp *>(D + d_L), v.lines, v.lines + 2 * kPairSteps,
                            out, out + 1, true, s));
    ck(hipStreamWaitEvent(s, v.join, 0));
    ck(hipMemcpyAsync(H + h_res, res, 8, hipMemcpyDeviceToHost, s));
    ck(hipStreamSynchronize(s));
    uint32_t rr[2];
    std::memcpy(rr, H + h_res, 8);
    if (rr[1]) return VKZG_ERR_BADARGS;
    *ok = rr[0] ? 1 : 0;
    if (terms) std::memcpy(terms, H + h_terms, 96);
    return VKZG_OK;
}

// Runs `body` with the prover's lock held and its device current, mapping
// device failures to VKZG_ERR_GPU.
template <typename Body>
vkzg_result locked(vkzg_prover *p, Body body) {
    std::lock_guard<std::mutex> lock(p->mu);
    DeviceGuard guard(p->device);
    try {
        if (!p->vfy) p->vfy = new VerifyState();
        return body();
    } catch (HipFail &) {
        return VKZG_ERR_GPU;
    } catch (std::bad_alloc &) {
        return VKZG_ERR_MALLOC;
    }
}

// Device -> 48-byte big-endian canonical encodings of n Fp values.
void fp_to_be(uint8_t *out, const Fp *vals, size_t n) {
    for (size_t i = 0; i < n; i++) {
        // vals are canonical integers in 28-bit limbs
        uint8_t le[56] = {0};
        uint64_t acc = 0;
        int bits = 0, o = 0;
        for (int l = 0; l < 14; l++) {
            acc |= (uint64_t)vals[i].v[l] << bits;
            bits += 28;
            while (bits >= 8) {
                le[o++] = (uint8_t)acc;
                acc >>= 8;
                bits -= 8;
            }
        }
        for (int b = 0; b < 48; b++) out[48 * i + b] = le[47 - b];
    }
}

}  // namespace

}  // namespace vkzg_amd

using namespace vkzg_amd;

extern "C" {

vkzg_result vkzg_amd_load_g2_setup(vkzg_prover *p, const uint8_t *g2_monomial, size_t len) {
    if (!p || !g2_monomial || len != kG2SetupPoints * kG2Bytes) return VKZG_ERR_BADARGS;
    return locked(p, [&]() -> vkzg_result {
        VerifyState &v = *p->vfy;
        v.ready = false;
        v.setup.release();
        DeviceArena tmp;
        hipStream_t s = p->stream;
        vkzg_result rc = VKZG_OK;
        try {
            uint8_t *d_in = tmp.dalloc<uint8_t>(len);
            G2A *pts = tmp.dalloc<G2A>(kG2SetupPoints);
            G2A *q = tmp.dalloc<G2A>(2);
            uint32_t *inf = tmp.dalloc<uint32_t>(kG2SetupPoints + 2);
            uint32_t *err = inf + kG2SetupPoints;
            uint32_t *h = tmp.halloc<uint32_t>(kG2SetupPoints + 2);
            v.lines = v.setup.dalloc<Fp2>(2 * 2 * kPairSteps);
            v.identity = v.setup.dalloc<G1A>(1);
            ck(hipMemsetAsync(v.identity, 0, sizeof(G1A), s));
            ck(hipMemcpyAsync(d_in, g2_monomial, len, hipMemcpyHostToDevice, s));
            ck(hipMemsetAsync(err, 0, 8, s));
            ck(launch_g2_decompress(pts, inf, d_in, kG2SetupPoints, err, s));
            ck(hipMemcpyAsync(q, pts + 64, sizeof(G2A), hipMemcpyDeviceToDevice, s));
            ck(hipMemcpyAsync(q + 1, pts, sizeof(G2A), hipMemcpyDeviceToDevice, s));
            ck(launch_g2_lines(v.lines, q, 2, err, s));
            ck(hipMemcpyAsync(h, inf, 4 * (kG2SetupPoints + 2), hipMemcpyDeviceToHost, s));
            ck(hipStreamSynchronize(s));
            if (h[kG2SetupPoints] || h[0] || h[64]) rc = VKZG_ERR_SETUP;
        } catch (HipFail &) {
            rc = VKZG_ERR_GPU;
        }
        tmp.release();
        if (rc != VKZG_OK) {
            v.setup.release();
            return rc;
        }
        v.ready = true;
        return VKZG_OK;
    });
}

vkzg_result vkzg_amd_verify_cell_kzg_proof_batch(vkzg_prover *p, int *ok, const uint8_t *commitments,
                                                 const uint64_t *cell_indices, const uint8_t *cells,
                                                 const uint8_t *proofs, size_t num_cells) {
    if (!p || !ok) return VKZG_ERR_BADARGS;
    *ok = 0;
    if (num_cells && (!commitments || !cell_indices || !cells || !proofs)) return VKZG_ERR_BADARGS;
    if (num_cells > kMaxVerifyCells) return VKZG_ERR_BADARGS;
    for (size_t k = 0; k < num_cells; k++)
        if (cell_indices[k] >= 128) return VKZG_ERR_BADARGS;
    return locked(p, [&]() -> vkzg_result {
        if (!p->vfy->ready) return VKZG_ERR_SETUP;
        if (num_cells == 0) {
            *ok = 1;
            return VKZG_OK;
        }
        return verify_impl(p, ok, commitments, cell_indices, cells, proofs, num_cells);
    });
}

vkzg_result vkzg_amd_debug_verify_terms(vkzg_prover *p, int *ok, uint8_t terms[96], const uint8_t *commitments,
                                        const uint64_t *cell_indices, const uint8_t *cells, const uint8_t *proofs,
                                        size_t num_cells) {
    if (!p || !ok || !terms || !num_cells || !commitments || !cell_indices || !cells || !proofs)
        return VKZG_ERR_BADARGS;
    *ok = 0;
    if (num_cells > kMaxVerifyCells) return VKZG_ERR_BADARGS;
    for (size_t k = 0; k < num_cells; k++)
        if (cell_indices[k] >= 128) return VKZG_ERR_BADARGS;
    return locked(p, [&]() -> vkzg_result {
        if (!p->vfy->ready) return VKZG_ERR_SETUP;
        return verify_impl(p, ok, commitments, cell_indices, cells, proofs, num_cells, terms);
    });
}

vkzg_result vkzg_amd_blob_to_commitments(vkzg_prover *p, uint8_t *commitments, const uint8_t *blobs,
                                         size_t num_blobs) {
    if (!p || !commitments || !blobs) return VKZG_ERR_BADARGS;
    if (num_blobs == 0) return VKZG_OK;
    return locked(p, [&]() -> vkzg_result {
        VerifyState &v = *p->vfy;
        hipStream_t s = p->stream;
        const uint32_t chunk = 64;  // blobs per launch sequence (one MSM each)
        for (size_t done = 0; done < num_blobs;) {
            const uint32_t nb = (uint32_t)std::min<size_t>(chunk, num_blobs - done);
            const uint32_t ne = nb * 4096;
            const MsmShape sh = msm_shape(ne, nb);
            Layout l;
            const size_t o_blobs = l.take(kBlobBytes * nb), o_err = l.take(4ull * nb), o_cm = l.take(48ull * nb);
            const size_t host_bytes = l.size;
            const size_t o_sc = l.take(sizeof(Fr) * (size_t)ne), o_pidx = l.take(4ull * ne), o_mid = l.take(ne);
            const size_t o_out = l.take(sizeof(G1A) * nb), o_msm = l.take(msm_scratch_bytes(sh));
            v.ensure(l.size, host_bytes);
            uint8_t *H = v.host, *D = static_cast<uint8_t *>(v.dev);
            std::memcpy(H + o_blobs, blobs + done * kBlobBytes, kBlobBytes * nb);
            ck(hipMemcpyAsync(D + o_blobs, H + o_blobs, kBlobBytes * nb, hipMemcpyHostToDevice, s));
            Fr *sc = reinterpret_cast<Fr *>(D + o_sc);
            uint32_t *pidx = reinterpret_cast<uint32_t *>(D + o_pidx);
            G1A *out = reinterpret_cast<G1A *>(D + o_out);
            ck(launch_commit_scalars(sc, pidx, D + o_mid, reinterpret_cast<uint32_t *>(D + o_err), D + o_blobs, nb,
                                     p->roots_inv, s));
            ck(launch_msm(sh, D + o_msm, p->srs, pidx, sc, D + o_mid, out, s));
            ck(launch_compress_points(D + o_cm, out, nb, s));
            ck(hipMemcpyAsync(H + o_err, D + o_err, 4ull * nb + 0, hipMemcpyDeviceToHost, s));
            ck(hipMemcpyAsync(H + o_cm, D + o_cm, 48ull * nb, hipMemcpyDeviceToHost, s));
            ck(hipStreamSynchronize(s));
            for (uint32_t b = 0; b < nb; b++) {
                uint32_t e;
                std::memcpy(&e, H + o_err + 4ull * b, 4);
                if (e) return VKZG_ERR_INVALID_BLOB;
            }
            std::memcpy(commitments + 48 * done, H + o_cm, 48ull * nb);
            done += nb;
        }
        return VKZG_OK;
    });
}

vkzg_result vkzg_amd_verify_challenge(uint8_t r_out[32], const uint8_t *commitments, const uint64_t *cell_indices,
                                      const uint8_t *cells, const uint8_t *proofs, size_t num_cells,
                                      int portable_sha) {
    if (!r_out || (num_cells && (!commitments || !cell_indices || !cells || !proofs))) return VKZG_ERR_BADARGS;
    std::vector<uint32_t> uniq, cidx;
    dedup(uniq, cidx, commitments, num_cells);
    uint32_t r[8];
    challenge(r, uniq, cidx, cell_indices, commitments, cells, proofs, num_cells, portable_sha == 0);
    for (int i = 0; i < 8; i++)
        for (int b = 0; b < 4; b++) r_out[31 - 4 * i - b] = (uint8_t)(r[i] >> (8 * b));
    return VKZG_OK;
}

vkzg_result vkzg_amd_sha256(uint8_t out[32], const uint8_t *data, size_t len, int portable) {
    if (!out || (len && !data)) return VKZG_ERR_BADARGS;
    sha256::Ctx c;
    sha256::init(c, portable == 0);
    // feed in uneven pieces so the buffering path is exercised as well
    size_t at = 0, step = 1;
    while (at < len) {
        const size_t take = std::min(step, len - at);
        sha256::update(c, data + at, take);
        at += take;
        step = step * 3 + 1;
    }
    sha256::finish(c, out);
    return VKZG_OK;
}

int vkzg_amd_sha256_hw(void) { return sha256::hw_available() ? 1 : 0; }

vkzg_result vkzg_amd_debug_pairing(vkzg_prover *p, uint8_t *out, const uint8_t *g1, const uint8_t *g2, size_t n) {
    if (!p || !out || !g1 || !g2) return VKZG_ERR_BADARGS;
    return locked(p, [&]() -> vkzg_result {
        DeviceArena tmp;
        hipStream_t s = p->stream;
        vkzg_result rc = VKZG_OK;
        try {
            uint8_t *d1 = tmp.dalloc<uint8_t>(48);
            uint8_t *d2 = tmp.dalloc<uint8_t>(96);
            G1A *P = tmp.dalloc<G1A>(2);
            G2A *Q = tmp.dalloc<G2A>(1);
            Fp2 *lines = tmp.dalloc<Fp2>(2 * kPairSteps);
            Fp *L = tmp.dalloc<Fp>(12 * kPairSteps);
            Fp *f = tmp.dalloc<Fp>(12);
            uint32_t *w = tmp.dalloc<uint32_t>(4);
            Fp *hf = tmp.halloc<Fp>(12);
            uint32_t *hw = tmp.halloc<uint32_t>(4);
            for (size_t i = 0; i < n && rc == VKZG_OK; i++) {
                ck(hipMemcpyAsync(d1, g1 + 48 * i, 48, hipMemcpyHostToDevice, s));
                ck(hipMemcpyAsync(d2, g2 + 96 * i, 96, hipMemcpyHostToDevice, s));
                ck(hipMemsetAsync(w, 0, 16, s));
                ck(hipMemsetAsync(P + 1, 0, sizeof(G1A), s));
                ck(launch_vfy_points(P, d1, 1, nullptr, 0, w + 1, s));
                ck(launch_g2_decompress(Q, w + 2, d2, 1, w + 1, s));
                ck(launch_g2_lines(lines, Q, 1, w + 1, s));
                ck(launch_pairing_check(w, f, L, lines, lines, P, P + 1, false, s));
                ck(launch_fp_canonical(f, 12, s));
                ck(hipMemcpyAsync(hf, f, sizeof(Fp) * 12, hipMemcpyDeviceToHost, s));
                ck(hipMemcpyAsync(hw, w, 16, hipMemcpyDeviceToHost, s));
                ck(hipStreamSynchronize(s));
                if (hw[1] || hw[2]) rc = VKZG_ERR_BADARGS;  // malformed point or G2 point at infinity
                else fp_to_be(out + 576 * i, hf, 12);
            }
        } catch (HipFail &) {
            rc = VKZG_ERR_GPU;
        }
        tmp.release();
        return rc;
    });
}

vkzg_result vkzg_amd_debug_g1_msm(vkzg_prover *p, uint8_t out[48], const uint8_t *points, const uint8_t *scalars,
                                  size_t n) {
    if (!p || !out || !points || !scalars || n == 0 || n > kMaxVerifyCells) return VKZG_ERR_BADARGS;
    std::vector<uint32_t> sc(8 * n);
    for (size_t i = 0; i < n; i++) {
        for (int w = 0; w < 8; w++) {
            const uint8_t *b = scalars + 32 * i + 28 - 4 * w;
            sc[8 * i + w] = (uint32_t)b[3] | (uint32_t)b[2] << 8 | (uint32_t)b[1] << 16 | (uint32_t)b[0] << 24;
        }
        bool lt = false;  // scalar < r
        for (int w = 7; w >= 0; w--)
            if (sc[8 * i + w] != cfr::P[w]) {
                lt = sc[8 * i + w] < cfr::P[w];
                break;
            }
        if (!lt) return VKZG_ERR_BADARGS;
    }
    return locked(p, [&]() -> vkzg_result {
        DeviceArena tmp;
        hipStream_t s = p->stream;
        vkzg_result rc = VKZG_OK;
        try {
            const uint32_t n32 = (uint32_t)n;
            const MsmShape sh = msm_shape(n32, 1);
            uint8_t *dp = tmp.dalloc<uint8_t>(48 * n);
            G1A *pts = tmp.dalloc<G1A>(n);
            Fr *dsc = tmp.dalloc<Fr>(n);
            uint32_t *pidx = tmp.dalloc<uint32_t>(n);
            uint8_t *mid = tmp.dalloc<uint8_t>(n);
            uint8_t *scratch = tmp.dalloc<uint8_t>(msm_scratch_bytes(sh));
            G1A *res = tmp.dalloc<G1A>(1);
            uint8_t *dout = tmp.dalloc<uint8_t>(48);
            uint32_t *err = tmp.dalloc<uint32_t>(1);
            uint32_t *herr = tmp.halloc<uint32_t>(1);
            std::vector<uint32_t> idx(n);
            for (uint32_t i = 0; i < n32; i++) idx[i] = i;
            ck(hipMemcpyAsync(dp, points, 48 * n, hipMemcpyHostToDevice, s));
            ck(hipMemcpyAsync(dsc, sc.data(), 32 * n, hipMemcpyHostToDevice, s));
            ck(hipMemcpyAsync(pidx, idx.data(), 4 * n, hipMemcpyHostToDevice, s));
            ck(hipMemsetAsync(mid, 0, n, s));
            ck(hipMemsetAsync(err, 0, 4, s));
            ck(launch_vfy_points(pts, dp, n32, nullptr, 0, err, s));
            ck(launch_msm(sh, scratch, pts, pidx, dsc, mid, res, s));
            ck(launch_compress_points(dout, res, 1, s));
            ck(hipMemcpyAsync(out, dout, 48, hipMemcpyDeviceToHost, s));
            ck(hipMemcpyAsync(herr, err, 4, hipMemcpyDeviceToHost, s));
            ck(hipStreamSynchronize(s));
            if (*herr) rc = VKZG_ERR_BADARGS;
        } catch (HipFail &) {
            rc = VKZG_ERR_GPU;
        }
        tmp.release();
        return rc;
    });
}

}  // extern "C"
