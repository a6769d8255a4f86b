// Device RNG helpers: splitmix64 seeding, jump by GF(2) matrix powers, and
// the reference's uniform / normal draws.
#pragma once

#include "common.hpp"

namespace madi {

// 128x128 GF(2) matrix, column-major: column j is the image of state bit j
// (bits 0..63 = s0, 64..127 = s1) as (lo = s0 part, hi = s1 part).
struct Gf2Col {
    uint64_t lo, hi;
};
constexpr int kJumpPowers = 48;  // J^(2^b) for b < 48: walker indices < 2^48

__device__ __forceinline__ Rng splitmix_state(uint64_t seed)
{
    uint64_t z = seed + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return Rng{z, z};
}

__device__ __forceinline__ Rng gf2_apply(const Gf2Col* __restrict__ M, Rng s)
{
    uint64_t lo = 0, hi = 0;
#pragma unroll 8
    for (int j = 0; j < 64; ++j) {
        const uint64_t m = 0ull - ((s.s0 >> j) & 1ull);
        lo ^= M[j].lo & m;
        hi ^= M[j].hi & m;
    }
#pragma unroll 8
    for (int j = 0; j < 64; ++j) {
        const uint64_t m = 0ull - ((s.s1 >> j) & 1ull);
        lo ^= M[64 + j].lo & m;
        hi ^= M[64 + j].hi & m;
    }
    return Rng{lo, hi};
}

// State of walker `index` for `seed`: jump^index(splitmix(seed)).
// jump_pow holds kJumpPowers matrices J^(2^b), 128 columns each.
__device__ __forceinline__ Rng rng_for_walker(uint64_t seed, uint64_t index, const Gf2Col* __restrict__ jump_pow)
{
    Rng s = splitmix_state(seed);
    for (int b = 0; b < kJumpPowers && (index >> b) != 0; ++b)
        if ((index >> b) & 1ull) s = gf2_apply(jump_pow + 128 * b, s);
    return s;
}

__device__ __forceinline__ double uniform_f64(Rng& r) { return u53_to_f64(rng_next(r)); }

// numba normal_float64 (numba-src:199-221): two float32 uniforms, Box-Muller
// first output, FP64 arithmetic; the float32 logarithm is taken as the
// correctly rounded value float(log(double(u1))).  normal_f64_of takes the
// two draws, normal_f64 makes them.
__device__ __forceinline__ double normal_f64_of(uint64_t a1, uint64_t a2)
{
    const float u1 = u53_to_f32(a1);
    const float u2 = u53_to_f32(a2);
    const float lg = (float)log((double)u1);
    return sqrt(-2.0 * (double)lg) * cos(6.283185307179586 * (double)u2);
}

__device__ __forceinline__ double normal_f64(Rng& r)
{
    const uint64_t a1 = rng_next(r);
    const uint64_t a2 = rng_next(r);
    return normal_f64_of(a1, a2);
}

}  // namespace madi
