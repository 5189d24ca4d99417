// CPU check of opf::FastDiv against the hardware divide (every divisor up to
// 4096 over edge and random dividends, plus large and power-of-two divisors)
// and of the XCD block mapping (a bijection with contiguous per-XCD ranges).
#include <cstdio>
#include <cstdint>
#include <random>
#include <vector>

#include "fastdiv.hpp"

int main() {
    std::mt19937_64 rng(7);
    long long checks = 0, bad = 0;
    auto test = [&](uint32_t d, uint32_t x) {
        opf::FastDiv f(d);
        ++checks;
        if (f.div(x) != x / d) {
            if (bad < 10) printf("FAIL d=%u x=%u got %u want %u\n", d, x, f.div(x), x / d);
            ++bad;
        }
    };
    const uint32_t edge[] = {0u, 1u, 2u, 3u, 0x7fffffffu, 0x80000000u, 0xfffffffeu, 0xffffffffu};
    for (uint32_t d = 1; d <= 4096; ++d) {
        for (uint32_t e : edge) test(d, e);
        for (uint32_t q = 0; q < 4; ++q) {  // around multiples of d
            uint64_t base = (uint64_t)(rng() % (0xffffffffu / d)) * d;
            for (int o = -2; o <= 2; ++o) {
                int64_t x = (int64_t)base + o;
                if (x >= 0 && x <= 0xffffffffll) test(d, (uint32_t)x);
            }
        }
        for (int r = 0; r < 64; ++r) test(d, (uint32_t)rng());
    }
    for (int k = 0; k < 32; ++k) {
        uint32_t d = 1u << k;
        for (int r = 0; r < 256; ++r) test(d, (uint32_t)rng());
        for (uint32_t e : edge) test(d, e);
    }
    for (int r = 0; r < 200000; ++r) {
        uint32_t d = (uint32_t)rng();
        if (d == 0) continue;
        test(d, (uint32_t)rng());
        test(d, d);
        test(d, d - 1);
    }
    // xcd_block is a bijection of [0, G) whose XCD ranges are contiguous
    for (uint32_t G = 1; G <= 5000; ++G) {
        std::vector<int> seen(G, 0);
        for (uint32_t b = 0; b < G; ++b) {
            uint32_t l = opf::xcd_block(b, G);
            ++checks;
            if (l >= G || seen[l]++) {
                if (bad < 10) printf("FAIL xcd_block G=%u b=%u -> %u\n", G, b, l);
                ++bad;
            }
            // consecutive hardware blocks of one XCD get consecutive logical blocks
            if (b >= 8 && opf::xcd_block(b, G) != opf::xcd_block(b - 8, G) + 1) {
                if (bad < 10) printf("FAIL xcd_block contiguity G=%u b=%u\n", G, b);
                ++bad;
            }
        }
    }
    printf("%lld checks, %lld failures\n", checks, bad);
    if (bad) return 1;
    printf("passed\n");
    return 0;
}
