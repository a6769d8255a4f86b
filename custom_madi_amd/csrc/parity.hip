// FP64 reference-semantics kernels: exact classifier, walk, column reduce,
// RNG states.  Built with -ffp-contract=off (no FMA contraction), so the
// arithmetic order matches the reference GPU kernels and the CPU oracle.
#include "common.hpp"
#include "kd_exact.hpp"
#include "phase.hpp"
#include "rng.hpp"
#include "launch.hpp"

namespace madi {

constexpr int kLdsNodes = 511;  // top 9 levels of the KD tree

// ------------------------------------------------------------ classify

__global__ __launch_bounds__(kBlock) void k_classify_f64(GeomDev g, const double* __restrict__ pts, int64_t n,
                                                        int32_t* __restrict__ cell, int8_t* __restrict__ inside,
                                                        int8_t* __restrict__ ok, unsigned long long* counters)
{
    __shared__ KdNodeF sf[kLdsNodes];
    __shared__ KdNodeI si[kLdsNodes];
    const KdView v = kd_stage(g, sf, si, kLdsNodes);
    long long nodes = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int c = 0, in = 0;
        const int k = classify_point_f64(g, v, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], c, in, nodes);
        cell[i] = c;
        inside[i] = (int8_t)in;
        ok[i] = (int8_t)k;
    }
    if (counters) {
        const long long tot = (long long)wave_sum_f64((double)nodes);
        if ((threadIdx.x & 63) == 0) atomicAdd(&counters[MADI_CNT_KD_NODES], (unsigned long long)tot);
    }
}

// ----------------------------------------------------------------- RNG

__global__ __launch_bounds__(kBlock) void k_rng_init(uint64_t seed, uint64_t first, int64_t n,
                                                    uint64_t* __restrict__ out, const Gf2Col* __restrict__ jump_pow)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Rng s = rng_for_walker(seed, first + (uint64_t)i, jump_pow);
    out[2 * i] = s.s0;
    out[2 * i + 1] = s.s1;
}

// ---------------------------------------------------------------- walk

// One walker of the reference step loop (madi/walker_gpu.py:531-656).
// Mode kY writes Y / inside samples; mode kFused accumulates the plan's
// moments instead and reduces phases at the end.
enum WalkMode { kY = 0, kFused = 1 };

template <int MODE>
__global__ __launch_bounds__(kBlock) void k_walk_f64(GeomDev g, WalkDev p, const Gf2Col* __restrict__ jump_pow,
                                                    const double* __restrict__ gi, const double* __restrict__ gd,
                                                    const double* __restrict__ gu, double* __restrict__ Y,
                                                    int8_t* __restrict__ itr, int8_t* __restrict__ esc_out,
                                                    int8_t* __restrict__ err_out, FusedOut fo,
                                                    unsigned long long* counters)
{
    __shared__ KdNodeF sf[kLdsNodes];
    __shared__ KdNodeI si[kLdsNodes];
    __shared__ double lds_cols[kWaves * kColChunk * 2];
    const KdView v = kd_stage(g, sf, si, kLdsNodes);

    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = w < p.n_walkers;
    const int G = p.n_grid;
    long long nodes = 0, crossings = 0, rejects = 0;

    double px = 0.0, py = 0.0, pz = 0.0;
    Rng r{0, 0};
    if (live) {
        if (p.use_golden) {
            px = gi[3 * w];
            py = gi[3 * w + 1];
            pz = gi[3 * w + 2];
        } else {
            r = rng_for_walker(p.seed, p.first_walker + (uint64_t)w, jump_pow);
            const double span = g.source_hi - g.source_lo;
            px = g.source_lo + uniform_f64(r) * span;
            py = g.source_lo + uniform_f64(r) * span;
            pz = g.source_lo + uniform_f64(r) * span;
        }
    }
    int cell = 0, in = 0;
    int err = 0, escaped = 0;
    if (live && !classify_point_f64(g, v, px, py, pz, cell, in, nodes)) {
        err = 1;
        escaped = 1;
    }
    const double half = 0.5 * g.L;
    double xp = px - half, yp = py - half, zp = pz - half;
    double Yx = 0.0, Yy = 0.0, Yz = 0.0;
    double dM[MADI_MAX_PAIRS][3];
#pragma unroll
    for (int q = 0; q < MADI_MAX_PAIRS; ++q) dM[q][0] = dM[q][1] = dM[q][2] = 0.0;
    if (MODE == kY && live) {
        double* Yw = Y + (size_t)w * G * 3;
        Yw[0] = Yw[1] = Yw[2] = 0.0;
        itr[(size_t)w * G] = (int8_t)(err ? 0 : in);
    }
    if (MODE == kFused) {
        if (live && !err && fo.occupancy) { /* sample j = 0 counted below */ }
        const long long c0 = wave_popc(live && !err && in);
        if ((threadIdx.x & 63) == 0 && c0) atomicAdd((unsigned long long*)&fo.occupancy[0], (unsigned long long)c0);
    }
    int next_ev = 0;
    for (int32_t step = 0; step < p.n_steps; ++step) {
        if (live && !escaped) {
            double dx, dy, dz, u;
            if (p.use_golden) {
                const size_t k = (size_t)step * p.n_walkers + w;
                dx = gd[3 * k];
                dy = gd[3 * k + 1];
                dz = gd[3 * k + 2];
                u = gu[k];
            } else {
                dx = normal_f64(r) * p.sigma;
                dy = normal_f64(r) * p.sigma;
                dz = normal_f64(r) * p.sigma;
                u = uniform_f64(r);
            }
            double nx = px + dx, ny = py + dy, nz = pz + dz;
            if (!g.periodic && (nx < 0.0 || nx >= g.L || ny < 0.0 || ny >= g.L || nz < 0.0 || nz >= g.L)) {
                escaped = 1;
            } else {
                int ncell = 0, nin = 0;
                if (!classify_point_f64(g, v, nx, ny, nz, ncell, nin, nodes)) {
                    err = 1;
                    escaped = 1;
                } else {
                    int m = 0;
                    if (in != nin) m = 1;
                    else if (in && nin && cell != ncell) m = 2;
                    if (m > 0) ++crossings;
                    if (m > 0 && p.pp < 1.0) {
                        const double thr = m == 1 ? p.pp : p.pp * p.pp;
                        if (u >= thr) {
                            nx = px;
                            ny = py;
                            nz = pz;
                            ncell = cell;
                            nin = in;
                            ++rejects;
                        }
                    }
                    px = nx;
                    py = ny;
                    pz = nz;
                    cell = ncell;
                    in = nin;
                    const double xn = px - half, yn = py - half, zn = pz - half;
                    Yx += 0.5 * (xp + xn) * p.ts;
                    Yy += 0.5 * (yp + yn) * p.ts;
                    Yz += 0.5 * (zp + zn) * p.ts;
                    xp = xn;
                    yp = yn;
                    zp = zn;
                }
            }
        }
        const int32_t s1 = step + 1;
        if (s1 % p.steps_per_h == 0) {
            const int j = s1 / p.steps_per_h;
            if (MODE == kY) {
                if (live) {
                    double* Yw = Y + ((size_t)w * G + j) * 3;
                    Yw[0] = Yx;
                    Yw[1] = Yy;
                    Yw[2] = Yz;
                    itr[(size_t)w * G + j] = (int8_t)in;
                }
            } else {
                const long long c = wave_popc(live && !err && in);
                if ((threadIdx.x & 63) == 0 && c) atomicAdd((unsigned long long*)&fo.occupancy[j], (unsigned long long)c);
                while (next_ev < fo.ev.n && fo.ev.j[next_ev] == j) {
                    const int q = fo.ev.pair[next_ev];
                    const double sg = (double)fo.ev.sign[next_ev];
#pragma unroll
                    for (int t = 0; t < MADI_MAX_PAIRS; ++t)
                        if (t == q) {
                            dM[t][0] += sg * Yx;
                            dM[t][1] += sg * Yy;
                            dM[t][2] += sg * Yz;
                        }
                    ++next_ev;
                }
            }
        }
    }
    if (MODE == kY) {
        if (live) {
            esc_out[w] = (int8_t)escaped;
            err_out[w] = (int8_t)err;
        }
    } else {
        const long long ne = wave_popc(live && escaped);
        const long long nr = wave_popc(live && err);
        if ((threadIdx.x & 63) == 0) {
            if (ne) atomicAdd((unsigned long long*)&fo.flags[MADI_FLAG_ESCAPED], (unsigned long long)ne);
            if (nr) atomicAdd((unsigned long long*)&fo.flags[MADI_FLAG_CLASSIFIER_ERR], (unsigned long long)nr);
        }
        double* slot = fo.partial + (size_t)(fo.tile_base + blockIdx.x) * fo.plan->n_cols * 2;
        emit_columns<double>(*fo.plan, dM, live && !escaped, lds_cols, slot);
    }
    if (counters) {
        const double a = wave_sum_f64((double)nodes);
        const double b = wave_sum_f64((double)crossings);
        const double c = wave_sum_f64((double)rejects);
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&counters[MADI_CNT_KD_NODES], (unsigned long long)a);
            atomicAdd(&counters[MADI_CNT_CROSSINGS], (unsigned long long)b);
            atomicAdd(&counters[MADI_CNT_REJECTS], (unsigned long long)c);
            // Every live walker classifies its start and each unescaped step.
        }
    }
}

// -------------------------------------------------------------- reduce

// Columns x chunks grid; each block sums a contiguous range of (walker, axis)
// samples of one column; chunk partials are then added in chunk order.
__global__ __launch_bounds__(kBlock) void k_reduce_f64(const double* __restrict__ Y, int64_t n, int32_t G,
                                                      const int32_t* __restrict__ jd, const int32_t* __restrict__ jD,
                                                      const int32_t* __restrict__ js, const double* __restrict__ pc,
                                                      int32_t n_cols, int32_t n_chunks, double* __restrict__ part)
{
    __shared__ double sc[kBlock], ss[kBlock];
    const int col = blockIdx.y;
    const int chunk = blockIdx.x;
    const int64_t total = n * 3;
    const int64_t per = (total + n_chunks - 1) / n_chunks;
    const int64_t lo = chunk * per;
    const int64_t hi = lo + per < total ? lo + per : total;
    const int a = jd[col], b = jD[col], c = js[col];
    const double k = pc[col];
    double cs = 0.0, sn = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const int64_t w = i / 3;
        const int ax = (int)(i - 3 * w);
        const double* Yw = Y + (size_t)w * G * 3;
        const double dMv = Yw[3 * a + ax] + Yw[3 * b + ax] - Yw[3 * c + ax];
        const double ph = k * dMv;
        cs += cos(ph);
        sn += sin(ph);
    }
    sc[threadIdx.x] = cs;
    ss[threadIdx.x] = sn;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sc[threadIdx.x] += sc[threadIdx.x + s];
            ss[threadIdx.x] += ss[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[((size_t)col * n_chunks + chunk) * 2 + 0] = sc[0];
        part[((size_t)col * n_chunks + chunk) * 2 + 1] = ss[0];
    }
}

__global__ void k_reduce_finish(const double* __restrict__ part, int32_t n_cols, int32_t n_chunks,
                                double* __restrict__ cos_out, double* __restrict__ sin_out)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= n_cols) return;
    double a = 0.0, b = 0.0;
    for (int k = 0; k < n_chunks; ++k) {
        a += part[((size_t)col * n_chunks + k) * 2 + 0];
        b += part[((size_t)col * n_chunks + k) * 2 + 1];
    }
    cos_out[col] = a;
    sin_out[col] = b;
}

// ------------------------------------------------------------ launches

int launch_classify_f64(const GeomDev& g, const double* pts, int64_t n, int32_t* cell, int8_t* inside, int8_t* ok,
                        unsigned long long* counters, hipStream_t s)
{
    if (n == 0) return MADI_OK;
    int64_t blocks = (n + kBlock - 1) / kBlock;
    if (blocks > 65535) blocks = 65535;
    hipLaunchKernelGGL(k_classify_f64, dim3((unsigned)blocks), dim3(kBlock), 0, s, g, pts, n, cell, inside, ok,
                       counters);
    return check_launch("k_classify_f64");
}

int launch_rng_init(uint64_t seed, uint64_t first, int64_t n, uint64_t* out, const Gf2Col* jump_pow, hipStream_t s)
{
    if (n == 0) return MADI_OK;
    const int64_t blocks = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_rng_init, dim3((unsigned)blocks), dim3(kBlock), 0, s, seed, first, n, out, jump_pow);
    return check_launch("k_rng_init");
}

int launch_walk_f64(const GeomDev& g, const WalkDev& p, const Gf2Col* jump_pow, const double* gi, const double* gd,
                    const double* gu, double* Y, int8_t* itr, int8_t* esc, int8_t* err, unsigned long long* counters,
                    hipStream_t s)
{
    if (p.n_walkers == 0) return MADI_OK;
    const int64_t blocks = (p.n_walkers + kBlock - 1) / kBlock;
    FusedOut fo{};
    hipLaunchKernelGGL(k_walk_f64<kY>, dim3((unsigned)blocks), dim3(kBlock), 0, s, g, p, jump_pow, gi, gd, gu, Y,
                       itr, esc, err, fo, counters);
    return check_launch("k_walk_f64");
}

int launch_walk_f64_fused(const GeomDev& g, const WalkDev& p, const Gf2Col* jump_pow, const FusedOut& fo,
                          unsigned long long* counters, hipStream_t s)
{
    if (p.n_walkers == 0) return MADI_OK;
    const int64_t blocks = (p.n_walkers + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_walk_f64<kFused>, dim3((unsigned)blocks), dim3(kBlock), 0, s, g, p, jump_pow, nullptr,
                       nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, fo, counters);
    return check_launch("k_walk_f64_fused");
}

int launch_reduce_f64(const double* Y, int64_t n, int32_t G, const int32_t* jd, const int32_t* jD, const int32_t* js,
                      const double* pc, int32_t n_cols, double* part, int32_t n_chunks, double* cos_out,
                      double* sin_out, hipStream_t s)
{
    if (n_cols == 0) return MADI_OK;
    hipLaunchKernelGGL(k_reduce_f64, dim3(n_chunks, n_cols), dim3(kBlock), 0, s, Y, n, G, jd, jD, js, pc, n_cols,
                       n_chunks, part);
    int rc = check_launch("k_reduce_f64");
    if (rc) return rc;
    hipLaunchKernelGGL(k_reduce_finish, dim3((n_cols + 127) / 128), dim3(128), 0, s, part, n_cols, n_chunks, cos_out,
                       sin_out);
    return check_launch("k_reduce_finish");
}

}  // namespace madi
