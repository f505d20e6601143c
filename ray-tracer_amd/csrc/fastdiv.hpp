// fastdiv.hpp -- unsigned 32-bit division by a divisor that is fixed for a
// whole launch (image width, light count), without the generic divide
// sequence in the kernel: the launcher computes a multiplier once, and each
// quotient is one high multiply, a subtract, an add and two shifts.
//
// Method (Granlund and Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, the round-up variant with a 33-bit multiplier):
// for d >= 2 let s = ceil(log2 d) and m = floor(2^32 (2^s - d) / d) + 1, so
// that 2^32 + m = ceil(2^(32+s) / d); then for every n < 2^32
//     floor(n / d) = (t + ((n - t) >> 1)) >> (s - 1),  t = floor(n m / 2^32).
// d = 1 returns n.  tests/test_invariant_division.py compiles this header on the host
// and checks it against plain division.
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#define RSH_FD_HD __host__ __device__
#else
#define RSH_FD_HD
#endif

namespace rsh {

struct FastDiv {
    uint32_t d, m, s;
};

inline FastDiv fast_div(uint32_t d) {
    FastDiv f{d, 0u, 0u};
    if (d <= 1u) return f;
    uint32_t s = 0;
    while ((1ull << s) < d) s++;
    f.m = (uint32_t)(((1ull << 32) * ((1ull << s) - d)) / d + 1u);
    f.s = s;
    return f;
}

RSH_FD_HD inline uint32_t fdiv(uint32_t n, const FastDiv& f) {
    if (f.d == 1u) return n;
    const uint32_t t = (uint32_t)(((uint64_t)n * f.m) >> 32);
    return (t + ((n - t) >> 1)) >> (f.s - 1u);
}

}  // namespace rsh
