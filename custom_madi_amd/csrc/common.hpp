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

// Rotate right by k (0 < k < 32) as two funnel shifts of the 32-bit halves.
__device__ __forceinline__ uint64_t rotr64_lt32(uint64_t x, uint32_t k)
{
    const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    const uint32_t nlo = __builtin_amdgcn_alignbit(hi, lo, k);
    const uint32_t nhi = __builtin_amdgcn_alignbit(lo, hi, k);
    return ((uint64_t)nhi << 32) | nlo;
}

// x ^ y ^ z in one instruction: gfx950's three-input bitwise operation with
// truth table 0x96 (the compiler does not form it from two XORs).
__device__ __forceinline__ uint32_t xor3_u32(uint32_t x, uint32_t y, uint32_t z)
{
    uint32_t r;
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(r) : "v"(x), "v"(y), "v"(z));
    return r;
}

__device__ __forceinline__ uint64_t rng_next(Rng& r)
{
    const uint64_t a = r.s0;
    uint64_t b = r.s1;
    const uint64_t out = a + b;
    b ^= a;
    const uint64_t ra = rotr64_lt32(a, 9);  // rotl 55
    const uint64_t bs = b << 14;
    r.s0 = ((uint64_t)xor3_u32((uint32_t)(ra >> 32), (uint32_t)(b >> 32), (uint32_t)(bs >> 32)) << 32) |
           xor3_u32((uint32_t)ra, (uint32_t)b, (uint32_t)bs);
    r.s1 = rotr64_lt32(b, 28);  // rotl 36
    return out;
}

// (x >> 11) * 2^-53 as hi * 2^-32 + (lo >> 11) * 2^-53: both terms and the
// sum are exact in FP64 (53 bits), so this is the same value.
__device__ __forceinline__ double u53_to_f64(uint64_t x)
{
    const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    return __builtin_fma((double)hi, 0x1.0p-32, (double)(lo >> 11) * 0x1.0p-53);
}

// float32(uniform53): the exact FP64 uniform rounded once to float, which is
// float((x>>11) * 2^-53) of the reference.
__device__ __forceinline__ float u53_to_f32(uint64_t x) { return (float)u53_to_f64(x); }

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
    const int4* cell_rec;       // per cell: (list offset, list end, hmax bits, alpha bits), one load
    const int4* nb_krec;        // per edge slot: neighbour k and k's (list offset, list end, hmax bits)
    const float* nb_kalpha;     // per edge slot: alpha of the neighbour
    const double* cell_xyz;     // per cell world coordinates
    const double4* nb_u64;      // per edge slot (ê, h) in FP64 (with a KD tree only; sentinels h = +inf)
    const int8_t* cell_flags;   // per cell: 1 = the FP64 graph labels defer to the KD classifier
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

// Load through a global-address-space view of p.  Pointers read from device
// memory (a batch's work table) are generic, and generic loads take the flat
// path, whose waits also cover outstanding LDS operations; every pointer
// this is used on points to device global memory.
template <typename T>
__device__ __forceinline__ T gload(const T* p)
{
    return *(const __attribute__((address_space(1))) T*)p;
}

// The same for float4 records (one 16-byte load of a native vector).
__device__ __forceinline__ float4 gload(const float4* p)
{
    typedef float v4f __attribute__((ext_vector_type(4)));
    const v4f v = *(const __attribute__((address_space(1))) v4f*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ int4 gload(const int4* p)
{
    typedef int v4i __attribute__((ext_vector_type(4)));
    const v4i v = *(const __attribute__((address_space(1))) v4i*)p;
    return make_int4(v.x, v.y, v.z, v.w);
}

constexpr int kMaxGraphIters = 256;

// Each cell's edge list is padded to a multiple of kEdgePad with sentinel
// edges (unit vector 0, half length +inf, neighbour = the cell itself, FP64
// vector of length 1e30): their margin is +inf in FP32 and their bisector is
// never the nearer one in FP64, so scans may read whole groups of kEdgePad.
constexpr int kEdgePad = 8;
// Largest padded edge list of one cell: the split walk's requests carry a
// list length in 24 bits.
constexpr int kMaxCellEdges = (1 << 24) - kEdgePad;

}  // namespace madi
