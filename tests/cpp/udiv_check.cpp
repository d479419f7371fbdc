// Checks golap_fmt::udiv32 against the hardware division: every divisor 1..100000 with edge and
// random numerators, and random divisors over the whole 32-bit range. Exit status 0 = pass.
#include <cstdio>
#include <random>

#include "../../golap_public_amd/csrc/udiv.hpp"

int main() {
    std::mt19937_64 rng(12345);
    uint64_t checks = 0;
    auto check = [&](uint32_t n, uint32_t d, golap_fmt::UDiv32 u) {
        ++checks;
        if (golap_fmt::udiv32(n, u) != n / d) {
            std::printf("FAIL n=%u d=%u got %u want %u\n", n, d, golap_fmt::udiv32(n, u), n / d);
            return false;
        }
        return true;
    };
    auto all = [&](uint32_t d) {
        const golap_fmt::UDiv32 u = golap_fmt::udiv32_magic(d);
        const uint32_t edges[] = {0u, 1u, d - 1, d, d + 1u, 2 * d - 1, 2 * d, 0xffffffffu, 0xfffffffeu,
                                  0xffffffffu / d * d, 0xffffffffu / d * d - 1, 0x80000000u, 0x7fffffffu};
        for (uint32_t n : edges)
            if (!check(n, d, u)) return false;
        for (int i = 0; i < 64; ++i)
            if (!check((uint32_t)rng(), d, u)) return false;
        return true;
    };
    for (uint32_t d = 1; d <= 100000; ++d)
        if (!all(d)) return 1;
    for (int i = 0; i < 200000; ++i)
        if (!all((uint32_t)rng() | 1u) || !all(1u + (uint32_t)(rng() >> 33))) return 1;
    const uint32_t special[] = {3600u, 86400u, 0x80000000u, 0xffffffffu, 0xfffffffeu, 0x80000001u};
    for (uint32_t d : special)
        if (!all(d)) return 1;
    std::printf("udiv32: %llu checks passed\n", (unsigned long long)checks);
    return 0;
}
