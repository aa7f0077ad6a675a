// Counter-based random numbers for the sweep kernels.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): a 4x32-bit counter and a 2x32-bit key map to
// four 32-bit words through ten rounds of two 32x32->64 multiplies.
// Each single-site update draws one block keyed by the context seed with
// counter (site, replica, pass index lo, pass index hi), so the numbers a
// site sees do not depend on launch shape, colour order or thread mapping.
//
// The reference draws from one sequential Xoroshiro128++ stream per runner
// (MC.hpp:95-100); a per-site counter RNG is what makes the colour-parallel
// sweep reproducible, at the price of a different (but equally distributed)
// random sequence.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pyro {

struct Words4 {
    uint32_t x, y, z, w;
};

__host__ __device__ inline Words4 philox4x32_10(Words4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        // each 32x32 -> 64-bit product as one multiply whose halves are both
        // used (one v_mad_u64_u32 on the GPU instead of a low and a high
        // multiply, all three quarter rate)
        const uint64_t p0 = uint64_t(0xD2511F53u) * c.x;
        const uint64_t p1 = uint64_t(0xCD9E8D57u) * c.z;
        const uint32_t lo0 = uint32_t(p0), hi0 = uint32_t(p0 >> 32);
        const uint32_t lo1 = uint32_t(p1), hi1 = uint32_t(p1 >> 32);
        c = Words4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// Five 24-bit uniforms in (0,1) from one 128-bit block: the high 24 bits of
// each word, plus the four low bytes' first three packed into a fifth.
// u = (m + 1/2) * 2^-24 is exact in float and double, so the FP32 and FP64
// paths consume identical uniforms.
struct Uniform5 {
    uint32_t m[5];
};

__host__ __device__ inline Uniform5 split_uniform5(Words4 w) {
    Uniform5 u;
    u.m[0] = w.x >> 8;
    u.m[1] = w.y >> 8;
    u.m[2] = w.z >> 8;
    u.m[3] = w.w >> 8;
    u.m[4] = (w.x & 0xFFu) | ((w.y & 0xFFu) << 8) | ((w.z & 0xFFu) << 16);
    return u;
}

// Component c (0..4) of u for a c known only at run time.
__host__ __device__ inline uint32_t uniform_at(const Uniform5& u, int c) {
    return c == 0 ? u.m[0] : c == 1 ? u.m[1] : c == 2 ? u.m[2] : c == 3 ? u.m[3] : u.m[4];
}

template <typename Real>
__host__ __device__ inline Real u24(uint32_t m) {
    return (Real(m) + Real(0.5)) * Real(5.9604644775390625e-08);  // 2^-24
}

// 53-bit uniform in (0,1) from two words (used by the replica exchange).
__host__ __device__ inline double u53(uint32_t a, uint32_t b) {
    const uint64_t m = (uint64_t(a) << 21) ^ uint64_t(b >> 11);
    return (double(m) + 0.5) * 1.1102230246251565e-16;  // 2^-53
}

// Stream tags: distinct keys for distinct uses of the same seed.
constexpr uint32_t kTagSweep = 0x00000000u;
constexpr uint32_t kTagInit = 0x5A17C0DEu;
constexpr uint32_t kTagSwap = 0x7E3B5EEDu;

}  // namespace pyro
