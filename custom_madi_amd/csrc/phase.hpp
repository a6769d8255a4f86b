// Fused PGSE phase reduction shared by the walk kernels (device code).
//
// Each walker ends its walk holding dM(pair) = Y(delta) + Y(Delta) -
// Y(Delta + delta) for every (delta, Delta) pair of the plan (the finite-lobe
// moment of madi/signal.py:10-13).  For column (pair, b, dir) the phase is
// pc(pair, b) * (dir . dM(pair)) and the walk kernels need the sums of its
// cos and sin over walkers.  Sums are formed in FP64 in a fixed order: a
// wave butterfly, then the block's waves in index order into the tile's slot;
// madi_tiles_finalize adds tile slots in a fixed chunked order.  Results are therefore
// bitwise reproducible whichever workgroup runs a tile.
#pragma once

#include "common.hpp"

namespace madi {

constexpr int kBlock = 256;          // threads per walk workgroup (4 waves)
constexpr int kWaves = kBlock / 64;
constexpr int kColChunk = 64;        // columns staged per LDS round

// Device copy of madi_phase_plan plus the derived event table.
struct PlanDev {
    int32_t n_pairs, n_b, n_dirs, n_cols;
    int32_t j_delta[MADI_MAX_PAIRS], j_Delta[MADI_MAX_PAIRS], j_sum[MADI_MAX_PAIRS];
    double pc[MADI_MAX_PAIRS * MADI_MAX_B];
    double dirs[MADI_MAX_DIRS * 3];
    float pc32[MADI_MAX_PAIRS * MADI_MAX_B];
    float dirs32[MADI_MAX_DIRS * 3];
};

__device__ __forceinline__ double wave_sum_f64(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ long long wave_popc(bool pred)
{
    return (long long)__popcll(__ballot(pred));
}

// cos/sin of a phase in FP32: reduce to revolutions in [-0.5, 0.5] and use
// the hardware sin/cos (which take revolutions).
__device__ __forceinline__ void sincos_f32(float ph, float& s, float& c)
{
    float r = ph * 0.15915494309189535f;
    r = r - rintf(r);
    s = __builtin_amdgcn_sinf(r);
    c = __builtin_amdgcn_cosf(r);
}

// FP32 phase of column (p, b, d) for the moment (mx, my, mz): pc(p, b) times
// the moment's projection on direction d.
__device__ __forceinline__ float column_phase_f32(const PlanDev& plan, int p, int b, int d, float mx, float my, float mz)
{
    const float proj = fmaf(plan.dirs32[3 * d], mx, fmaf(plan.dirs32[3 * d + 1], my, plan.dirs32[3 * d + 2] * mz));
    return plan.pc32[p * plan.n_b + b] * proj;
}

// Accumulate one column value pair into the wave's LDS slot and flush full
// chunks to the tile slot.  lds: kWaves x kColChunk x 2 doubles.
__device__ __forceinline__ void column_put(double cs, double sn, int col, int n_cols, double* lds, double* tile_slot)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    cs = wave_sum_f64(cs);
    sn = wave_sum_f64(sn);
    const int k = col % kColChunk;
    if (lane == 0) {
        lds[(wave * kColChunk + k) * 2 + 0] = cs;
        lds[(wave * kColChunk + k) * 2 + 1] = sn;
    }
    if (k == kColChunk - 1 || col == n_cols - 1) {
        __syncthreads();
        const int base = col - k;
        const int t = threadIdx.x;
        if (t <= k) {
            double a = 0.0, b = 0.0;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                a += lds[(w * kColChunk + t) * 2 + 0];
                b += lds[(w * kColChunk + t) * 2 + 1];
            }
            tile_slot[(base + t) * 2 + 0] = a;
            tile_slot[(base + t) * 2 + 1] = b;
        }
        __syncthreads();
    }
}

// Emit every column of the plan for this thread's walker.  dM is the
// walker's moment per pair (unused pairs ignored); active=false contributes
// zeros (threads past the end of the walker range, escaped walkers).
template <typename T, int NP = MADI_MAX_PAIRS>
__device__ void emit_columns(const PlanDev& plan, const T (&dM)[NP][3], bool active, double* lds, double* tile_slot)
{
    const int n_b = plan.n_b, n_dirs = plan.n_dirs, n_cols = plan.n_cols;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        if (p < plan.n_pairs) {
            const T mx = dM[p][0], my = dM[p][1], mz = dM[p][2];
            for (int b = 0; b < n_b; ++b) {
                for (int d = 0; d < n_dirs; ++d) {
                    const int col = (p * n_b + b) * n_dirs + d;
                    double cs = 0.0, sn = 0.0;
                    if constexpr (sizeof(T) == 8) {
                        const double proj = plan.dirs[3 * d] * mx + plan.dirs[3 * d + 1] * my + plan.dirs[3 * d + 2] * mz;
                        const double ph = plan.pc[p * n_b + b] * proj;
                        if (active) {
                            cs = cos(ph);
                            sn = sin(ph);
                        }
                    } else {
                        float s, c;
                        sincos_f32(column_phase_f32(plan, p, b, d, mx, my, mz), s, c);
                        if (active) {
                            cs = (double)c;
                            sn = (double)s;
                        }
                    }
                    column_put(cs, sn, col, n_cols, lds, tile_slot);
                }
            }
        }
    }
}

// Y-grid event table: at grid index j, pair p gains +Y (delta, Delta) or
// -Y (Delta + delta).
struct Events {
    int32_t n;
    int32_t n_pairs;
    int32_t j[3 * MADI_MAX_PAIRS];
    int32_t pair[3 * MADI_MAX_PAIRS];
    int32_t sign[3 * MADI_MAX_PAIRS];
};

}  // namespace madi
