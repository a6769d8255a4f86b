// Shared device-side definitions for the MADI walk kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/madi_hip.h"

namespace madi {

// ------------------------------------------------------------------ RNG
// xoroshiro128+ (rotations 55/14/36) with splitmix64 seeding; walker i of a
// seed owns the state advanced by i jumps of 2^64 steps.  This is the stream
// numba.cuda.random defines (numba-src:45-240), restated from the published
// algorithm.

struct Rng {
    uint64_t s0, s1;
};

__device__ __forceinline__ uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

__device__ __forceinline__ uint64_t rng_next(Rng& r)
{
    const uint64_t a = r.s0;
    uint64_t b = r.s1;
    const uint64_t out = a + b;
    b ^= a;
    r.s0 = rotl64(a, 55) ^ b ^ (b << 14);
    r.s1 = rotl64(b, 36);
    return out;
}

__device__ __forceinline__ double u53_to_f64(uint64_t x) { return (double)(x >> 11) * 0x1.0p-53; }

// float32(uniform53): the 53-bit integer rounded to the nearest float, then
// scaled by 2^-53 (exact).  Equals float((x>>11) * 2^-53) of the reference.
__device__ __forceinline__ float u53_to_f32(uint64_t x)
{
    const uint64_t k = x >> 11;
    const double d = (double)k;  // exact: k < 2^53
    return (float)d * 0x1.0p-53f;
}

// ------------------------------------------------------------- geometry

// KD node of the FP64 exact classifier.  Nodes are stored in breadth-first
// order so the top levels form a contiguous prefix that kernels stage in LDS.
struct KdNodeF {
    double x, y, z, pad;
};
struct KdNodeI {
    int32_t left, right, parent, sid_axis;  // sid_axis = sid << 2 | axis
};

struct GeomDev {
    // exact KD classifier (FP64)
    const KdNodeF* kd_f;
    const KdNodeI* kd_i;
    const double* kd_annulus;   // per KD point
    const int32_t* kd_label;    // KD point -> cell label (periodic images map to their cell)
    const double* kd_seed_xyz;  // per KD point, xyz (for seed-to-seed separation)
    int32_t n_kd;
    int32_t n_kd_lds;           // number of leading BFS nodes staged in LDS
    // Delaunay-graph classifier (FP32 with FP64 certification)
    const int32_t* nb_off;      // n_cells + 1
    const float4* nb_rec;       // (ê.x, ê.y, ê.z, h): unit edge vector and half length
    const int32_t* nb_id;       // neighbour cell of each edge
    const double* nb_e64;       // 3 per edge: exact edge vector c_k + shift - c_a
    const float* cell_alpha;    // per cell
    const double* cell_alpha64;
    const float* cell_hmax;     // max half edge length per cell
    const double* cell_xyz;     // per cell world coordinates
    const int32_t* hint;        // hint_n^3 cells: a seed near each hint-grid box centre
    int32_t n_cells;
    int32_t hint_n;
    double hint_h;
    float alpha_min;
    float alpha_max;
    // domain
    double L;
    double source_lo, source_hi;
    int32_t periodic;
    int32_t free_water;
};

constexpr int kMaxGraphIters = 256;

}  // namespace madi
