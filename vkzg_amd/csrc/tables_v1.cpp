// The reference's tables file, version 1 (src/precomputed_tables.cpp:16-58),
// written from a prover's device-resident setup: the file the reference's
// loader reads from $VKZG_TABLES_PATH, e.g. to run the reference (or its Java
// binding) on the same setup as this build.  This build itself never reads
// it: its own tables are derived on the device at prover creation.
//
// Layout, little-endian u32 throughout:
//   header        {0x475A4B56 'VKZG', 1, items_plus_words, items_minus_words}
//   position      128 x 64 x 32 entries [2^(8d)] x_ext_fft[j][i], ((j*64+i)*32+d)*24
//                 words, affine x then y in the reference's Montgomery form
//                 (x 2^384 mod p, 12 words each), the identity (0, 0)
//   roots_fwd     omega_8192^k, Montgomery (x 2^256 mod r), 8192 x 8 words
//   roots_inv     omega_8192^-k, likewise
//   items +/-     phase-B bucket items e | d << 8 | sign << 16 for the 8-bit
//                 signed digits of K+ and K- (taps e < 64, positions d < 32),
//                 grouped by bucket |digit| - 1 (k_phase_b_split.comp:25-61)
//   offsets +/-   129 bucket starts each
//   perm +/-      the 128 buckets in descending load (any order is valid)
//   inv_blob      Montgomery 1/4096 (8 words)
// Digits follow the reference's recoding (k_phase_a_sort.comp:31-40): byte
// plus carry, minus 256 with a carry when above 128, so digits lie in
// [-127, 128].
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/vkzg_amd.h"
#include "runtime.h"

namespace vkzg_amd {

namespace {

constexpr uint32_t kV1Magic = 0x475A4B56u;
constexpr uint32_t kV1Version = 1u;
constexpr uint32_t kV1Positions = 32;

// Items, offsets and permutation of one phase-B half from its 64 canonical tap scalars.
void v1_items(const Fr *taps, std::vector<uint32_t> &items, std::vector<uint32_t> &offsets,
              std::vector<uint32_t> &perm) {
    std::vector<std::vector<uint32_t>> lists(128);
    for (uint32_t e = 0; e < 64; e++) {
        uint32_t carry = 0;
        for (uint32_t d = 0; d < kV1Positions; d++) {
            const uint32_t byte = (taps[e].v[d / 4] >> (8 * (d % 4))) & 0xffu;
            const uint32_t v = byte + carry;
            int digit = (int)v;
            carry = 0;
            if (v > 128) {
                digit = (int)v - 256;
                carry = 1;
            }
            if (digit == 0) continue;
            const uint32_t mag = (uint32_t)(digit < 0 ? -digit : digit);
            lists[mag - 1].push_back(e | (d << 8) | ((digit < 0 ? 1u : 0u) << 16));
        }
    }
    items.clear();
    offsets.assign(129, 0);
    for (uint32_t k = 0; k < 128; k++) {
        offsets[k] = (uint32_t)items.size();
        items.insert(items.end(), lists[k].begin(), lists[k].end());
    }
    offsets[128] = (uint32_t)items.size();
    perm.resize(128);
    for (uint32_t k = 0; k < 128; k++) perm[k] = k;
    std::stable_sort(perm.begin(), perm.end(),
                     [&](uint32_t a, uint32_t b) { return lists[a].size() > lists[b].size(); });
}

bool write_all(FILE *f, const void *p, size_t n) { return std::fwrite(p, 1, n, f) == n; }

}  // namespace

}  // namespace vkzg_amd

using namespace vkzg_amd;

extern "C" vkzg_result vkzg_amd_write_v1_tables(vkzg_prover *p, const char *path) {
    if (!p || !path) return VKZG_ERR_BADARGS;
    std::lock_guard<std::mutex> lock(p->mu);
    DeviceGuard guard(p->device);
    DeviceArena tmp;
    std::vector<uint32_t> positions((size_t)8192 * kV1Positions * 24);
    std::vector<Fr> roots(8192), roots_inv(8192), kpm(128);
    try {
        hipStream_t s = p->stream;
        G1J *pwj = tmp.dalloc<G1J>((size_t)kV1Positions * 8192);
        Fp *prefix = tmp.dalloc<Fp>((size_t)kV1Positions * 8192);
        G1A *pw = tmp.dalloc<G1A>((size_t)kV1Positions * 8192);
        uint32_t *words = tmp.dalloc<uint32_t>(positions.size());
        Fr *d_kpm = tmp.dalloc<Fr>(128);
        static_assert(kV1Positions == 32, "2^(8d) positions d < 32: the (8, 0) window plan");
        ck(launch_table_chains(pwj, p->xext, window_plan(8, 0), s));
        ck(launch_jac_to_affine(pw, pwj, prefix, (uint64_t)kV1Positions * 8192, 16, s));
        ck(launch_v1_positions(words, pw, s));
        ck(launch_phase_b_scalars(d_kpm, p->roots, s));
        ck(hipMemcpyAsync(positions.data(), words, positions.size() * 4, hipMemcpyDeviceToHost, s));
        ck(hipMemcpyAsync(roots.data(), p->roots, 8192 * sizeof(Fr), hipMemcpyDeviceToHost, s));
        ck(hipMemcpyAsync(roots_inv.data(), p->roots_inv, 8192 * sizeof(Fr), hipMemcpyDeviceToHost, s));
        ck(hipMemcpyAsync(kpm.data(), d_kpm, 128 * sizeof(Fr), hipMemcpyDeviceToHost, s));
        ck(hipStreamSynchronize(s));
    } catch (HipFail &) {
        tmp.release();
        return VKZG_ERR_GPU;
    }
    tmp.release();
    std::vector<uint32_t> items_p, items_m, off_p, off_m, perm_p, perm_m;
    v1_items(kpm.data(), items_p, off_p, perm_p);
    v1_items(kpm.data() + 64, items_m, off_m, perm_m);
    const uint32_t header[4] = {kV1Magic, kV1Version, (uint32_t)items_p.size(), (uint32_t)items_m.size()};
    FILE *f = std::fopen(path, "wb");
    if (!f) return VKZG_ERR_IO;
    bool ok = write_all(f, header, sizeof(header)) && write_all(f, positions.data(), positions.size() * 4) &&
              write_all(f, roots.data(), roots.size() * sizeof(Fr)) &&
              write_all(f, roots_inv.data(), roots_inv.size() * sizeof(Fr)) &&
              write_all(f, items_p.data(), items_p.size() * 4) && write_all(f, items_m.data(), items_m.size() * 4) &&
              write_all(f, off_p.data(), off_p.size() * 4) && write_all(f, off_m.data(), off_m.size() * 4) &&
              write_all(f, perm_p.data(), perm_p.size() * 4) && write_all(f, perm_m.data(), perm_m.size() * 4) &&
              write_all(f, cfr::INV_4096_M, sizeof(cfr::INV_4096_M));
    ok = (std::fclose(f) == 0) && ok;
    return ok ? VKZG_OK : VKZG_ERR_IO;
}
