// MAP library matching: for every voxel row m of a measured matrix, the
// library row r minimising the residual of madi/library.py's matchers, in
// FP64 with the order of the reference's per-voxel kernels
// (madi/fitters_gpu.py:96-154): each dot product is accumulated over the
// features in order with fused multiply-adds, and the scan keeps the first
// row of smallest score (strict <, so ties go to the lowest index).
//
//   fixed S0 : score_j = |r_j|^2 - 2 m.r_j,  residual = |m|^2 + min score
//   free S0  : s0_j = m.r_j / rr_j, residual_j = |m|^2 - (m.r_j)^2 / rr_j,
//              rr_j = max(|r_j|^2, 1e-30), residual_j = +inf where s0_j <= 0
//
// Layout: both matrices row-major FP64 (voxel / library row, then feature).
// A workgroup of 256 threads owns 128 voxels and walks the whole library in
// tiles of 64 rows; each thread accumulates an 8 x 4 block of dot products
// (voxels tv + 16 i, rows tl + 16 l) in registers.  Up to 24 columns
// (k_map_match_res) the voxel tile stays in LDS for all columns and the
// library tiles are double-buffered with a register prefetch, one barrier per
// tile; wider acquisitions (k_map_match) stage both tiles 8 columns at a
// time.  Each
// thread keeps its own best row per voxel in increasing row order; the 16
// threads of a voxel then combine (score, row) lexicographically, which
// returns the lowest row among the equal minima exactly as the sequential
// scan does.  The kernel is bound by FP64 FMA issue (2 flops per voxel,
// row, feature).  Compiled without FP contraction: only the explicit fma()
// calls fuse.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "launch.hpp"

namespace madi {
namespace {

constexpr int kMatchThreads = 256;
constexpr int kTV = 128;  // voxels per workgroup
constexpr int kTL = 64;   // library rows per tile
constexpr int kKC = 8;    // features per LDS stage
constexpr int kRV = kTV / 16;
constexpr int kRL = kTL / 16;

// out[r] = sum_k x[r,k]^2 in feature order (fused), then max(., floor) with
// NaN kept (np.maximum semantics) when floor > 0.
__global__ __launch_bounds__(256) void k_row_sumsq(const double* __restrict__ x, int64_t n, int32_t nf, double floor_v,
                                                   double* __restrict__ out)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const double* row = x + r * (int64_t)nf;
    double s = 0.0;
    for (int k = 0; k < nf; ++k) s = fma(row[k], row[k], s);
    if (floor_v > 0.0 && s < floor_v) s = floor_v;
    out[r] = s;
}

__device__ __forceinline__ void take_better(double& s, int& j, double& z, double os, int oj, double oz)
{
    if (os < s || (os == s && oj < j)) {
        s = os;
        j = oj;
        z = oz;
    }
}

// One thread's best rows per voxel (rows in increasing order within the thread).
struct Best {
    double s[kRV], z[kRV], mn[kRV];
    int j[kRV];
};

// Score the thread's kRV x kRL dot products of library tile j0 and keep the
// first strict minimum per voxel.  Fixed S0: fma(-2, d, rn) equals the
// reference's rn - 2 d exactly (doubling is exact).
template <bool FREE>
__device__ __forceinline__ void score_tile(const double (&acc)[kRV][kRL], int64_t j0, int tl, int64_t n_lib,
                                           const double* __restrict__ rnorm, Best& B)
{
#pragma unroll
    for (int l = 0; l < kRL; ++l) {
        const int64_t j = j0 + tl + 16 * l;
        if (j < n_lib) {
            const double rn = rnorm[j];
#pragma unroll
            for (int i = 0; i < kRV; ++i) {
                const double d = acc[i][l];
                if (FREE) {
                    const double s0 = d / rn;
                    double res = B.mn[i] - (d * d) / rn;
                    if (s0 <= 0.0) res = INFINITY;
                    if (res < B.s[i]) {
                        B.s[i] = res;
                        B.j[i] = (int)j;
                        B.z[i] = s0;
                    }
                } else {
                    // scores are never -0 (rn >= +0), and a NaN score never
                    // wins: fmin keeps the running minimum as the strict scan does
                    const double score = fma(-2.0, d, rn);
                    B.j[i] = score < B.s[i] ? (int)j : B.j[i];
                    B.s[i] = fmin(B.s[i], score);
                }
            }
        }
    }
}

// Combine the 16 threads of each voxel (lanes tv, tv+16, tv+32, tv+48 of a
// wave, then the four waves) and write the block's voxels.  Ws/Wj may alias
// LDS the caller no longer reads (the caller synchronises before the call).
template <bool FREE>
__device__ __forceinline__ void finish(Best& B, int64_t v0, int64_t n_vox, const double* __restrict__ mnorm,
                                       int64_t* __restrict__ best_idx, double* __restrict__ best_resid,
                                       double* __restrict__ best_s0, double (*Ws)[4][kTV], int (*Wj)[kTV])
{
    const int t = threadIdx.x, tv = t & 15, w = t >> 6;
#pragma unroll
    for (int i = 0; i < kRV; ++i) {
#pragma unroll
        for (int m = 16; m < 64; m <<= 1) {
            const double os = __shfl_xor(B.s[i], m, 64);
            const int oj = __shfl_xor(B.j[i], m, 64);
            const double oz = FREE ? __shfl_xor(B.z[i], m, 64) : 0.0;
            take_better(B.s[i], B.j[i], B.z[i], os, oj, oz);
        }
    }
    if ((t & 63) < 16) {
#pragma unroll
        for (int i = 0; i < kRV; ++i) {
            const int v = tv + 16 * i;
            Ws[0][w][v] = B.s[i];
            Ws[1][w][v] = B.z[i];
            Wj[w][v] = B.j[i];
        }
    }
    __syncthreads();
    if (t < kTV) {
        double s = Ws[0][0][t], z = Ws[1][0][t];
        int j = Wj[0][t];
        for (int q = 1; q < 4; ++q) take_better(s, j, z, Ws[0][q][t], Wj[q][t], Ws[1][q][t]);
        const int64_t v = v0 + t;
        if (v < n_vox) {
            best_idx[v] = j;
            if (FREE) {
                best_resid[v] = s;
                best_s0[v] = z;
            } else {
                best_resid[v] = mnorm[v] + s;
            }
        }
    }
}

__device__ __forceinline__ void best_init(Best& B, int64_t v0, int64_t n_vox, const double* __restrict__ mnorm)
{
    const int tv = threadIdx.x & 15;
#pragma unroll
    for (int i = 0; i < kRV; ++i) {
        B.s[i] = INFINITY;
        B.z[i] = 0.0;
        B.j[i] = 0;
        const int64_t v = v0 + tv + 16 * i;
        B.mn[i] = v < n_vox ? mnorm[v] : 0.0;
    }
}

// Up to kResF columns: the block's voxel tile stays in LDS for every column,
// library tiles are double-buffered (the next tile is fetched into registers
// while the current one is used), one barrier per library tile.
constexpr int kResF = 24;
constexpr int kPre = (kTL * kResF + kMatchThreads - 1) / kMatchThreads;

template <bool FREE>
__global__ __launch_bounds__(kMatchThreads) __attribute__((amdgpu_waves_per_eu(FREE ? 2 : 3))) void k_map_match_res(const double* __restrict__ M, int64_t n_vox,
                                                                 const double* __restrict__ R, int64_t n_lib,
                                                                 int32_t nf, const double* __restrict__ mnorm,
                                                                 const double* __restrict__ rnorm,
                                                                 int64_t* __restrict__ best_idx,
                                                                 double* __restrict__ best_resid,
                                                                 double* __restrict__ best_s0)
{
    __shared__ double As[kResF * kTV];      // [k][v]; reused for the final combine
    __shared__ double Bs[2][kResF * kTL];   // [k][l]
    const int t = threadIdx.x, tv = t & 15, tl = t >> 4;
    const int64_t v0 = (int64_t)blockIdx.x * kTV;
    Best B;
    best_init(B, v0, n_vox, mnorm);

    // the voxel tile, all columns (row-major reads, consecutive threads)
    for (int e = t; e < kTV * nf; e += kMatchThreads) {
        const int v = e / nf, k = e - v * nf;
        const int64_t gv = v0 + v;
        As[k * kTV + v] = gv < n_vox ? M[gv * nf + k] : 0.0;
    }
    // this thread's slots of a library tile, (row l << 16) | column k per
    // prefetch register, -1 when unused
    int slot[kPre];
#pragma unroll
    for (int q = 0; q < kPre; ++q) {
        const int e = t + kMatchThreads * q;
        const int l = e / nf;
        slot[q] = e < kTL * nf ? (l << 16) | (e - l * nf) : -1;
    }
    double pre[kPre];
#pragma unroll
    for (int q = 0; q < kPre; ++q) {
        const int l = slot[q] >> 16, k = slot[q] & 0xffff;
        pre[q] = (slot[q] >= 0 && l < n_lib) ? R[(int64_t)l * nf + k] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < kPre; ++q)
        if (slot[q] >= 0) Bs[0][(slot[q] & 0xffff) * kTL + (slot[q] >> 16)] = pre[q];
    __syncthreads();

    int cur = 0;
    for (int64_t j0 = 0; j0 < n_lib; j0 += kTL) {
        const int64_t jn = j0 + kTL;
        if (jn < n_lib) {
#pragma unroll
            for (int q = 0; q < kPre; ++q) {
                const int l = slot[q] >> 16, k = slot[q] & 0xffff;
                pre[q] = (slot[q] >= 0 && jn + l < n_lib) ? R[(jn + l) * nf + k] : 0.0;
            }
        }
        // the first column's products start the sums: a*b equals fma(a, b, +0)
        // but for the sign of a zero, which no score or S0 test can see
        double acc[kRV][kRL];
        const double* Bc = Bs[cur];
        if (nf > 0) {
#pragma unroll
            for (int i = 0; i < kRV; ++i)
#pragma unroll
                for (int l = 0; l < kRL; ++l) acc[i][l] = As[tv + 16 * i] * Bc[tl + 16 * l];
        } else {
#pragma unroll
            for (int i = 0; i < kRV; ++i)
#pragma unroll
                for (int l = 0; l < kRL; ++l) acc[i][l] = 0.0;
        }
        for (int k = 1; k < nf; ++k) {
            double a[kRV], b[kRL];
#pragma unroll
            for (int i = 0; i < kRV; ++i) a[i] = As[k * kTV + tv + 16 * i];
#pragma unroll
            for (int l = 0; l < kRL; ++l) b[l] = Bc[k * kTL + tl + 16 * l];
#pragma unroll
            for (int i = 0; i < kRV; ++i)
#pragma unroll
                for (int l = 0; l < kRL; ++l) acc[i][l] = fma(a[i], b[l], acc[i][l]);
        }
        score_tile<FREE>(acc, j0, tl, n_lib, rnorm, B);
        if (jn < n_lib) {
            // the other buffer was last read before the previous barrier
#pragma unroll
            for (int q = 0; q < kPre; ++q)
                if (slot[q] >= 0) Bs[cur ^ 1][(slot[q] & 0xffff) * kTL + (slot[q] >> 16)] = pre[q];
        }
        __syncthreads();
        cur ^= 1;
    }
    finish<FREE>(B, v0, n_vox, mnorm, best_idx, best_resid, best_s0, reinterpret_cast<double (*)[4][kTV]>(As),
                 reinterpret_cast<int (*)[kTV]>(As + 2 * 4 * kTV));
}

// Any column count: features staged 8 at a time for both tiles.
template <bool FREE>
__global__ __launch_bounds__(kMatchThreads) void k_map_match(const double* __restrict__ M, int64_t n_vox,
                                                             const double* __restrict__ R, int64_t n_lib, int32_t nf,
                                                             const double* __restrict__ mnorm,
                                                             const double* __restrict__ rnorm,
                                                             int64_t* __restrict__ best_idx,
                                                             double* __restrict__ best_resid,
                                                             double* __restrict__ best_s0)
{
    __shared__ double As[kKC][kTV];
    __shared__ double Bs[kKC][kTL];
    __shared__ double Ws[2][4][kTV];
    __shared__ int Wj[4][kTV];

    const int t = threadIdx.x;
    const int tv = t & 15, tl = t >> 4;
    const int64_t v0 = (int64_t)blockIdx.x * kTV;
    Best B;
    best_init(B, v0, n_vox, mnorm);

    for (int64_t j0 = 0; j0 < n_lib; j0 += kTL) {
        double acc[kRV][kRL];
#pragma unroll
        for (int i = 0; i < kRV; ++i)
#pragma unroll
            for (int l = 0; l < kRL; ++l) acc[i][l] = 0.0;

        for (int f0 = 0; f0 < nf; f0 += kKC) {
            const int kc = nf - f0 < kKC ? nf - f0 : kKC;
            for (int e = t; e < kTV * kKC; e += kMatchThreads) {
                const int v = e / kKC, k = e % kKC;
                const int64_t gv = v0 + v;
                As[k][v] = (gv < n_vox && k < kc) ? M[gv * nf + f0 + k] : 0.0;
            }
            for (int e = t; e < kTL * kKC; e += kMatchThreads) {
                const int l = e / kKC, k = e % kKC;
                const int64_t gj = j0 + l;
                Bs[k][l] = (gj < n_lib && k < kc) ? R[gj * nf + f0 + k] : 0.0;
            }
            __syncthreads();
            for (int k = 0; k < kc; ++k) {
                double a[kRV], b[kRL];
#pragma unroll
                for (int i = 0; i < kRV; ++i) a[i] = As[k][tv + 16 * i];
#pragma unroll
                for (int l = 0; l < kRL; ++l) b[l] = Bs[k][tl + 16 * l];
#pragma unroll
                for (int i = 0; i < kRV; ++i)
#pragma unroll
                    for (int l = 0; l < kRL; ++l) acc[i][l] = fma(a[i], b[l], acc[i][l]);
            }
            __syncthreads();
        }
        score_tile<FREE>(acc, j0, tl, n_lib, rnorm, B);
    }
    finish<FREE>(B, v0, n_vox, mnorm, best_idx, best_resid, best_s0, Ws, Wj);
}

int64_t map_workspace_doubles(int64_t n_vox, int64_t n_lib) { return n_vox + n_lib; }

}  // namespace
}  // namespace madi

extern "C" {

int madi_map_workspace(int64_t n_vox, int64_t n_lib, int64_t* bytes)
{
    if (!bytes || n_vox < 0 || n_lib < 0) {
        madi::set_error("madi_map_workspace: invalid arguments");
        return MADI_ERR_ARG;
    }
    const int64_t d = madi::map_workspace_doubles(n_vox, n_lib);
    *bytes = (d > 0 ? d : 1) * (int64_t)sizeof(double);
    return MADI_OK;
}

int madi_map_match(const double* d_measured, int64_t n_vox, const double* d_lib, int64_t n_lib, int32_t n_feat,
                   int32_t fit_s0, void* d_workspace, int64_t workspace_bytes, int64_t* d_best_idx,
                   double* d_best_resid, double* d_best_s0, void* stream)
{
    using namespace madi;
    if (n_vox < 0 || n_lib < 1 || n_feat < 0 || (fit_s0 != 0 && fit_s0 != 1) || !d_workspace ||
        ((uintptr_t)d_workspace & 7) != 0 || (n_vox > 0 && (!d_best_idx || !d_best_resid)) ||
        (n_vox > 0 && fit_s0 && !d_best_s0) || (n_feat > 0 && (!d_lib || (n_vox > 0 && !d_measured)))) {
        set_error("madi_map_match: invalid arguments");
        return MADI_ERR_ARG;
    }
    int64_t need = 0;
    int rc = madi_map_workspace(n_vox, n_lib, &need);
    if (rc) return rc;
    if (workspace_bytes < need) {
        set_error("madi_map_match: workspace of %lld bytes is smaller than the %lld required",
                  (long long)workspace_bytes, (long long)need);
        return MADI_ERR_ARG;
    }
    const int64_t blocks = (n_vox + kTV - 1) / kTV;
    if (blocks > 0x7fffffffLL || n_lib > 0x7fffffffLL - kTL) {
        set_error("madi_map_match: %lld voxels or %lld library rows exceed one launch", (long long)n_vox,
                  (long long)n_lib);
        return MADI_ERR_ARG;
    }
    if (n_vox == 0) return MADI_OK;
    hipStream_t s = (hipStream_t)stream;
    double* mnorm = (double*)d_workspace;
    double* rnorm = mnorm + n_vox;
    hipLaunchKernelGGL(k_row_sumsq, dim3((unsigned)((n_vox + 255) / 256)), dim3(256), 0, s, d_measured, n_vox, n_feat,
                       0.0, mnorm);
    hipLaunchKernelGGL(k_row_sumsq, dim3((unsigned)((n_lib + 255) / 256)), dim3(256), 0, s, d_lib, n_lib, n_feat,
                       fit_s0 ? 1e-30 : 0.0, rnorm);
    if ((rc = check_launch("k_row_sumsq"))) return rc;
    const dim3 grid((unsigned)blocks), block(kMatchThreads);
    const bool res = n_feat <= kResF;
    if (fit_s0 && res)
        hipLaunchKernelGGL(k_map_match_res<true>, grid, block, 0, s, d_measured, n_vox, d_lib, n_lib, n_feat, mnorm,
                           rnorm, d_best_idx, d_best_resid, d_best_s0);
    else if (res)
        hipLaunchKernelGGL(k_map_match_res<false>, grid, block, 0, s, d_measured, n_vox, d_lib, n_lib, n_feat, mnorm,
                           rnorm, d_best_idx, d_best_resid, d_best_s0);
    else if (fit_s0)
        hipLaunchKernelGGL(k_map_match<true>, grid, block, 0, s, d_measured, n_vox, d_lib, n_lib, n_feat, mnorm, rnorm,
                           d_best_idx, d_best_resid, d_best_s0);
    else
        hipLaunchKernelGGL(k_map_match<false>, grid, block, 0, s, d_measured, n_vox, d_lib, n_lib, n_feat, mnorm, rnorm,
                           d_best_idx, d_best_resid, d_best_s0);
    return check_launch("k_map_match");
}

}  // extern "C"
