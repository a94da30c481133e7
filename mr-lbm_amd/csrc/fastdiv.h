// Division of 32-bit unsigned integers by a divisor fixed at launch: one
// multiply-high and three shifts/adds, exact for every 32-bit dividend
// (round-up multiplier method, Granlund and Montgomery, "Division by
// invariant integers using multiplication", PLDI 1994, fig. 4.1). The step
// kernels use it for tile coordinates: with uniform operands it compiles to
// scalar instructions, where a division by a runtime value would take a
// float reciprocal and conversions on the vector unit in every wave.
#pragma once

#include <stdint.h>

#ifndef __HIPCC__
#define MRLBM_HD
#else
#define MRLBM_HD __host__ __device__
#endif

namespace mrlbm {

struct FastDiv {
    uint32_t d, m, s1, s2;
    MRLBM_HD uint32_t div(uint32_t n) const {
        const uint32_t t = (uint32_t)(((uint64_t)m * n) >> 32);
        return (t + ((n - t) >> s1)) >> s2;
    }
};

// l = ceil(log2 d); m = floor(2^32 (2^l - d) / d) + 1; s1 = min(l, 1);
// s2 = max(l - 1, 0). d >= 1.
inline FastDiv make_fastdiv(uint32_t d) {
    uint32_t l = 0;
    while (l < 32 && ((uint64_t)1 << l) < d) l++;
    FastDiv f;
    f.d = d;
    f.m = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << l) - d)) / d + 1);
    f.s1 = l < 1 ? l : 1;
    f.s2 = l > 0 ? l - 1 : 0;
    return f;
}

}  // namespace mrlbm
