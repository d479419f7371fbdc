// Exact unsigned 32-bit division by a divisor known only at run time, as a multiply-high and
// two shifts (T. Granlund, P. Montgomery, "Division by Invariant Integers using
// Multiplication", PLDI 1994, section 4, figure 4.1): for 1 <= d < 2^32 with l = ceil(log2 d),
//   m = floor(2^32 * (2^l - d) / d) + 1,  t = mulhi(m, n),  q = (t + ((n - t) >> s1)) >> s2,
// s1 = min(l, 1), s2 = max(l - 1, 0), gives q = floor(n / d) for every 32-bit n.
#pragma once
#include <cstdint>

#ifndef __HIPCC__
#define GOLAP_HD
#else
#include <hip/hip_runtime.h>
#define GOLAP_HD __host__ __device__
#endif

namespace golap_fmt {

struct UDiv32 {
    uint32_t m;
    uint32_t s1, s2;
};

inline UDiv32 udiv32_magic(uint32_t d) {  // d >= 1
    uint32_t l = 0;
    while (l < 32 && (1ull << l) < d) ++l;
    const unsigned __int128 num = ((unsigned __int128)1 << 32) * (((unsigned __int128)1 << l) - d);
    UDiv32 u;
    u.m = (uint32_t)(num / d + 1);
    u.s1 = l < 1 ? l : 1;
    u.s2 = l > 1 ? l - 1 : 0;
    return u;
}

GOLAP_HD inline uint32_t udiv32(uint32_t n, UDiv32 u) {
    const uint32_t t = (uint32_t)(((uint64_t)u.m * n) >> 32);
    return (t + ((n - t) >> u.s1)) >> u.s2;
}

}  // namespace golap_fmt
