// SHA-256 (FIPS 180-4) on the host, for the Fiat-Shamir challenge of batch
// cell-proof verification.  Uses the x86 SHA extensions when the CPU has
// them, else a portable implementation; both give the same digests.
#pragma once

#include <cstddef>
#include <cstdint>

namespace vkzg_amd {
namespace sha256 {

struct Ctx {
    uint32_t h[8];
    uint8_t buf[64];
    size_t buflen;
    uint64_t bytes;
    bool hw;
};

bool hw_available();
void init(Ctx &c, bool allow_hw = true);
void update(Ctx &c, const void *data, size_t len);
void finish(Ctx &c, uint8_t out[32]);

}  // namespace sha256
}  // namespace vkzg_amd
