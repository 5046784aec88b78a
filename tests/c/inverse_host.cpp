// Host build of vkzg_amd/csrc/inverse.h for tests/test_inverse.py: reads one
// hexadecimal integer below 2^382 per line, prints "ok l0 l1 ... l13" with the
// limbs of the signed inverse (limb 13 signed), decimal.
#define __host__
#define __device__
#include <cstdio>
#include <cstring>
#include <string>

#include "../../vkzg_amd/csrc/inverse.h"

int main() {
    char line[256];
    while (std::fgets(line, sizeof line, stdin)) {
        std::string hex(line);
        while (!hex.empty() && (hex.back() == '\n' || hex.back() == '\r')) hex.pop_back();
        if (hex.empty()) continue;
        uint32_t y[14] = {};
        // big-endian hex -> 28-bit limbs (7 hex digits per limb)
        int limb = 0;
        for (int end = (int)hex.size(); end > 0 && limb < 14; end -= 7, limb++) {
            const int start = end - 7 > 0 ? end - 7 : 0;
            y[limb] = (uint32_t)std::strtoul(hex.substr(start, end - start).c_str(), nullptr, 16);
        }
        int32_t v[14];
        const bool ok = vkzg_amd::bgcd::inverse(v, y);
        std::printf("%d", ok ? 1 : 0);
        for (int i = 0; i < 14; i++) std::printf(" %d", v[i]);
        std::printf("\n");
    }
    return 0;
}
