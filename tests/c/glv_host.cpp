// Host build of vkzg_amd/csrc/glv.h for tests/test_verify_host.py: reads one
// hexadecimal scalar below r per line, prints "k1 k2" in hexadecimal.
#define __host__
#define __device__
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../vkzg_amd/csrc/glv.h"

int main() {
    char line[256];
    while (std::fgets(line, sizeof line, stdin)) {
        std::string hex(line);
        while (!hex.empty() && (hex.back() == '\n' || hex.back() == '\r')) hex.pop_back();
        if (hex.empty()) continue;
        uint32_t k[8] = {};
        int w = 0;
        for (int end = (int)hex.size(); end > 0 && w < 8; end -= 8, w++) {
            const int start = end - 8 > 0 ? end - 8 : 0;
            k[w] = (uint32_t)std::strtoul(hex.substr(start, end - start).c_str(), nullptr, 16);
        }
        uint32_t k1[4], k2[4];
        vkzg_amd::glv::split(k1, k2, k);
        std::printf("%08x%08x%08x%08x %08x%08x%08x%08x\n", k1[3], k1[2], k1[1], k1[0], k2[3], k2[2], k2[1], k2[0]);
    }
    return 0;
}
