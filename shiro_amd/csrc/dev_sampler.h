// QMC sampler for the kernels: Shiro's fallback OpenQmcSampler
// (reference src/render/OpenQmcSampler.cpp:20-71, 110-135), bit-exact with
// the CPU.  Dimensions 0 and 1 are an R2 sequence shared by all pixels;
// higher dimensions are a radical inverse in base prime[dim % 24] rotated
// (Cranley-Patterson) by a per-(pixel, dim) hash.  Branch b of Next2D uses
// dimensions 2b and 2b+1; Next1D(b) is the first of them.
//
// Bit-exactness: the radical inverse accumulates fp32 products in the same
// order as the CPU (digit * weight, weight *= 1/base) with contraction off;
// 1/base comes from a table of correctly rounded constants; integer division
// by a compile-time base lets the compiler use multiply-high sequences.
#pragma once

#include "dev_math.h"

namespace shiro_hip {

DEV uint32_t mix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return h;
}

DEV float hash_to_unit(uint32_t key) { return static_cast<float>(mix32(key) & 0x00ffffffu) / 16777216.0f; }

template <uint32_t B>
DEV float radical_inverse_b(uint32_t n) {
    constexpr float inv = 1.0f / static_cast<float>(B);
    float w = inv;
    float acc = 0.0f;
    while (n > 0u) {
        const uint32_t q = n / B;
        acc += static_cast<float>(n - q * B) * w;
        n = q;
        w *= inv;
    }
    return acc;
}

// Base chosen by dim % 24 (primes 2..89).
DEV float radical_inverse_slot(uint32_t slot, uint32_t n) {
    switch (slot) {
        case 0: return radical_inverse_b<2>(n);
        case 1: return radical_inverse_b<3>(n);
        case 2: return radical_inverse_b<5>(n);
        case 3: return radical_inverse_b<7>(n);
        case 4: return radical_inverse_b<11>(n);
        case 5: return radical_inverse_b<13>(n);
        case 6: return radical_inverse_b<17>(n);
        case 7: return radical_inverse_b<19>(n);
        case 8: return radical_inverse_b<23>(n);
        case 9: return radical_inverse_b<29>(n);
        case 10: return radical_inverse_b<31>(n);
        case 11: return radical_inverse_b<37>(n);
        case 12: return radical_inverse_b<41>(n);
        case 13: return radical_inverse_b<43>(n);
        case 14: return radical_inverse_b<47>(n);
        case 15: return radical_inverse_b<53>(n);
        case 16: return radical_inverse_b<59>(n);
        case 17: return radical_inverse_b<61>(n);
        case 18: return radical_inverse_b<67>(n);
        case 19: return radical_inverse_b<71>(n);
        case 20: return radical_inverse_b<73>(n);
        case 21: return radical_inverse_b<79>(n);
        case 22: return radical_inverse_b<83>(n);
        default: return radical_inverse_b<89>(n);
    }
}

DEV float qmc_value(uint32_t px, uint32_t py, uint32_t s, uint32_t dim) {
    if (dim < 2u) {
        const float step = dim == 0u ? 0.7548776662466927f : 0.5698402909980532f;
        const float v = 0.5f + step * static_cast<float>(s);
        return v - floorf(v);
    }
    const float seq = radical_inverse_slot(dim % 24u, s + 1u);
    const uint32_t key = (px * 0x8da6b343u) ^ (py * 0xd8163841u) ^ (dim * 0xcb1ab31fu) ^ 0x9e3779b9u;
    const float v = seq + hash_to_unit(key);
    return v - floorf(v);
}

struct Sampler {
    uint32_t px, py, s;
    DEV V2 next2(uint32_t b) const { return V2{qmc_value(px, py, s, 2u * b), qmc_value(px, py, s, 2u * b + 1u)}; }
    DEV float next1(uint32_t b) const { return qmc_value(px, py, s, 2u * b); }
};

}  // namespace shiro_hip
