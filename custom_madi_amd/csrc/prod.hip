// Production walk: FP32 seed-relative state, Delaunay-graph classifier with
// FP64 certification, certified safe ball, fused PGSE phase reduction.
//
// Classification.  A point's label needs its nearest seed s1 and the test of
// every shifted facet of s1 (SI Eq. S2).  The facets that can bind are those
// of s1's Delaunay neighbours (the contracted cell is the Voronoi cell eroded
// by alpha, and the Voronoi cell is the intersection of the half-spaces of
// the Delaunay neighbours only), and a seed that is nearer to the point than
// all its Delaunay neighbours is the nearest seed.  So from an anchor seed a
// the classifier evaluates, for every Delaunay edge a->k with unit vector u
// and half length h, the signed distance of the point p (relative to a) to
// the bisector plane:  m = h - p.u.  m < 0 means k is nearer (greedy move to
// the nearest such k); at the nearest seed, inside <=> min m >= alpha_a.
// This is the reference's decision (madi/walker_gpu.py:272-335) restated on
// the Delaunay graph; any FP32 decision closer than its rounding bound to a
// threshold is re-decided in FP64 for the same represented point.
//
// Safe ball.  For an inside point the contracted polytope contains the ball
// of radius min m - alpha_a; for an outside point no contracted cell is
// nearer than min(alpha_a - min m, min m + alpha_min).  Later endpoints
// inside that ball keep the label without evaluating any facet.
#include "common.hpp"
#include "phase.hpp"
#include "rng.hpp"
#include "launch.hpp"

namespace madi {

struct Locate {
    int ok;
    int inside;
    int uncertain_nearest;  // a neighbour is equidistant within FP32 error
    float minm;
    float eps;
};

// FP64 decision for the point anchor + p (p relative to the anchor).  Moves
// the anchor to the FP64 nearest seed; returns 0 if the greedy walk does not
// settle or two seeds coincide.
__device__ int locate_f64(const GeomDev& g, int& a, float& px, float& py, float& pz, float& ox, float& oy, float& oz,
                          int& inside, long long& facets)
{
    double qx = px, qy = py, qz = pz;
    double sx = ox, sy = oy, sz = oz;
    for (int it = 0; it < kMaxGraphIters; ++it) {
        const int lo = g.nb_off[a], hi = g.nb_off[a + 1];
        const double da2 = qx * qx + qy * qy + qz * qz;
        double best = da2;
        int bk = -1;
        for (int k = lo; k < hi; ++k) {
            const double ex = g.nb_e64[3 * k], ey = g.nb_e64[3 * k + 1], ez = g.nb_e64[3 * k + 2];
            const double dx = qx - ex, dy = qy - ey, dz = qz - ez;
            const double dk2 = dx * dx + dy * dy + dz * dz;
            if (dk2 < best) {
                best = dk2;
                bk = k;
            }
        }
        facets += hi - lo;
        if (bk >= 0) {
            const double ex = g.nb_e64[3 * bk], ey = g.nb_e64[3 * bk + 1], ez = g.nb_e64[3 * bk + 2];
            qx -= ex;
            qy -= ey;
            qz -= ez;
            sx += ex;
            sy += ey;
            sz += ez;
            a = g.nb_id[bk];
            continue;
        }
        const double a1 = g.cell_alpha64[a];
        int in = 1;
        for (int k = lo; k < hi; ++k) {
            const double ex = g.nb_e64[3 * k], ey = g.nb_e64[3 * k + 1], ez = g.nb_e64[3 * k + 2];
            const double dx = qx - ex, dy = qy - ey, dz = qz - ez;
            const double dk2 = dx * dx + dy * dy + dz * dz;
            const double sep = sqrt(ex * ex + ey * ey + ez * ez);
            if (sep <= 0.0) return 0;
            const double margin = (dk2 - da2) / (2.0 * sep);
            if (margin < a1) in = 0;
        }
        inside = in;
        px = (float)qx;
        py = (float)qy;
        pz = (float)qz;
        ox = (float)sx;
        oy = (float)sy;
        oz = (float)sz;
        return 1;
    }
    return 0;
}

// FP32 greedy location from anchor a; p and o are updated on every move.
__device__ __forceinline__ Locate locate_f32(const GeomDev& g, int& a, float& px, float& py, float& pz, float& ox,
                                             float& oy, float& oz, long long& iters, long long& facets)
{
    Locate r{0, 0, 0, 0.f, 0.f};
    for (int it = 0; it < kMaxGraphIters; ++it) {
        const int lo = g.nb_off[a], hi = g.nb_off[a + 1];
        const float eps = 4.8e-7f * (fabsf(px) + fabsf(py) + fabsf(pz) + g.cell_hmax[a]);
        float best = 0.f, minm = __builtin_inff(), mabs = __builtin_inff();
        int bk = -1;
        for (int k = lo; k < hi; ++k) {
            const float4 e = g.nb_rec[k];
            const float m = e.w - (px * e.x + py * e.y + pz * e.z);
            const float q = m * e.w;
            if (m < -eps && q < best) {
                best = q;
                bk = k;
            }
            minm = fminf(minm, m);
            mabs = fminf(mabs, fabsf(m));
        }
        facets += hi - lo;
        if (bk >= 0) {
            const double ex = g.nb_e64[3 * bk], ey = g.nb_e64[3 * bk + 1], ez = g.nb_e64[3 * bk + 2];
            px = (float)((double)px - ex);
            py = (float)((double)py - ey);
            pz = (float)((double)pz - ez);
            ox = (float)((double)ox + ex);
            oy = (float)((double)oy + ey);
            oz = (float)((double)oz + ez);
            a = g.nb_id[bk];
            ++iters;
            continue;
        }
        r.ok = 1;
        r.minm = minm;
        r.eps = eps;
        r.uncertain_nearest = mabs <= eps;
        r.inside = minm >= g.cell_alpha[a];
        return r;
    }
    return r;
}

// Hint-grid anchor for a world point.
__device__ __forceinline__ int hint_anchor(const GeomDev& g, double x, double y, double z)
{
    const int n = g.hint_n;
    int ix = (int)floor(x / g.hint_h), iy = (int)floor(y / g.hint_h), iz = (int)floor(z / g.hint_h);
    ix = ix < 0 ? 0 : (ix >= n ? n - 1 : ix);
    iy = iy < 0 ? 0 : (iy >= n ? n - 1 : iy);
    iz = iz < 0 ? 0 : (iz >= n ? n - 1 : iz);
    return g.hint[(ix * n + iy) * n + iz];
}

__device__ __forceinline__ double wrap_f64(double x, double L)
{
    double w = x - L * floor(x / L);
    return w >= L ? w - L : w;
}

// Certified label of a world point; returns ok.  anchor/p/o describe the
// located point on return.
__device__ int classify_world(const GeomDev& g, double x, double y, double z, bool resolve_cell, int& a, float& px,
                              float& py, float& pz, float& ox, float& oy, float& oz, int& inside, Locate& L,
                              long long& iters, long long& facets, long long& rechecks)
{
    if (g.periodic) {
        x = wrap_f64(x, g.L);
        y = wrap_f64(y, g.L);
        z = wrap_f64(z, g.L);
    }
    a = hint_anchor(g, x, y, z);
    const double* c = g.cell_xyz + 3 * a;
    px = (float)(x - c[0]);
    py = (float)(y - c[1]);
    pz = (float)(z - c[2]);
    ox = -px;
    oy = -py;
    oz = -pz;
    L = locate_f32(g, a, px, py, pz, ox, oy, oz, iters, facets);
    if (!L.ok) return 0;
    inside = L.inside;
    const float a1 = g.cell_alpha[a];
    if (fabsf(L.minm - a1) <= L.eps || (resolve_cell && L.uncertain_nearest)) {
        ++rechecks;
        if (!locate_f64(g, a, px, py, pz, ox, oy, oz, inside, facets)) return 0;
        L.uncertain_nearest = 0;
        L.minm = 0.f;  // no safe ball after an FP64 decision
    }
    return 1;
}

__global__ __launch_bounds__(kBlock) void k_classify_graph(GeomDev g, const double* __restrict__ pts, int64_t n,
                                                          int32_t* __restrict__ cell, int8_t* __restrict__ inside,
                                                          int8_t* __restrict__ ok, unsigned long long* counters)
{
    long long iters = 0, facets = 0, rechecks = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (g.free_water) {
            cell[i] = 0;
            inside[i] = 0;
            ok[i] = 1;
            continue;
        }
        int a, in = 0;
        float px, py, pz, ox, oy, oz;
        Locate L;
        const int k = classify_world(g, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], true, a, px, py, pz, ox, oy, oz,
                                     in, L, iters, facets, rechecks);
        cell[i] = k ? a : 0;
        inside[i] = (int8_t)(k ? in : 0);
        ok[i] = (int8_t)k;
    }
    if (counters) {
        const double a = wave_sum_f64((double)iters), b = wave_sum_f64((double)facets),
                     c = wave_sum_f64((double)rechecks);
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&counters[MADI_CNT_GRAPH_ITERS], (unsigned long long)a);
            atomicAdd(&counters[MADI_CNT_FACETS], (unsigned long long)b);
            atomicAdd(&counters[MADI_CNT_FP64_RECHECK], (unsigned long long)c);
        }
    }
}

// Box-Muller first output in FP32 from the same two float32 uniforms the
// FP64 walk uses; cos(2 pi u2) via the hardware cosine in revolutions.
__device__ __forceinline__ float normal_f32(Rng& r)
{
    const float u1 = u53_to_f32(rng_next(r));
    const float u2 = u53_to_f32(rng_next(r));
    return __builtin_sqrtf(-2.0f * __logf(u1)) * __builtin_amdgcn_cosf(u2);
}

__device__ __forceinline__ void safe_ball(const GeomDev& g, int a, const Locate& L, int inside, float& r2)
{
    float r;
    if (L.minm == 0.f) r = 0.f;  // FP64-decided point
    else if (inside) r = L.minm - g.cell_alpha[a];
    else if (L.uncertain_nearest) r = 0.f;
    else r = fminf(g.cell_alpha[a] - L.minm, L.minm + g.alpha_min);
    r -= 2.f * L.eps;
    r2 = r > 0.f ? r * r : 0.f;
}

__global__ __launch_bounds__(kBlock) void k_walk_f32(GeomDev g, WalkDev p, const Gf2Col* __restrict__ jump_pow,
                                                    FusedOut fo, unsigned long long* counters)
{
    __shared__ double lds_cols[kWaves * kColChunk * 2];
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = w < p.n_walkers;
    long long n_class = 0, n_safe = 0, iters = 0, facets = 0, rechecks = 0, crossings = 0, rejects = 0;

    Rng rng{0, 0};
    double x0 = 0.0, y0 = 0.0, z0 = 0.0;
    if (live) {
        rng = rng_for_walker(p.seed, p.first_walker + (uint64_t)w, jump_pow);
        const double span = g.source_hi - g.source_lo;
        x0 = g.source_lo + uniform_f64(rng) * span;
        y0 = g.source_lo + uniform_f64(rng) * span;
        z0 = g.source_lo + uniform_f64(rng) * span;
    }
    int a = 0, in = 0, cell = 0, err = 0, escaped = 0;
    float px = 0.f, py = 0.f, pz = 0.f, ox = 0.f, oy = 0.f, oz = 0.f;
    float bx = 0.f, by = 0.f, bz = 0.f, r2 = 0.f;
    if (live && !g.free_water) {
        Locate L;
        ++n_class;
        if (!classify_world(g, x0, y0, z0, false, a, px, py, pz, ox, oy, oz, in, L, iters, facets, rechecks)) {
            err = 1;
            escaped = 1;
        } else {
            cell = a;
            ox = -px;
            oy = -py;
            oz = -pz;
            bx = px;
            by = py;
            bz = pz;
            safe_ball(g, a, L, in, r2);
        }
    } else if (live) {
        ox = oy = oz = 0.f;
        r2 = __builtin_inff();  // free water: every endpoint is outside
    }
    // escape bounds on the displacement D = o + p (world = start + D)
    const float lox = (float)(-x0), loy = (float)(-y0), loz = (float)(-z0);
    const float hix = (float)(g.L - x0), hiy = (float)(g.L - y0), hiz = (float)(g.L - z0);

    float Sx = 0.f, Sy = 0.f, Sz = 0.f, Cx = 0.f, Cy = 0.f, Cz = 0.f;
    float Dx = ox + px, Dy = oy + py, Dz = oz + pz;
    float dM[MADI_MAX_PAIRS][3];
#pragma unroll
    for (int q = 0; q < MADI_MAX_PAIRS; ++q) dM[q][0] = dM[q][1] = dM[q][2] = 0.f;
    {
        const long long c0 = wave_popc(live && !err && in);
        if ((threadIdx.x & 63) == 0 && c0) atomicAdd((unsigned long long*)&fo.occupancy[0], (unsigned long long)c0);
    }
    const float sig = (float)p.sigma;
    const float ts = (float)p.ts;
    int next_ev = 0;
    for (int32_t step = 0; step < p.n_steps; ++step) {
        if (live && !escaped) {
            const float dx = normal_f32(rng) * sig;
            const float dy = normal_f32(rng) * sig;
            const float dz = normal_f32(rng) * sig;
            const uint64_t ku = rng_next(rng) >> 11;
            float qx = px + dx, qy = py + dy, qz = pz + dz;
            const float nDx = ox + qx, nDy = oy + qy, nDz = oz + qz;
            if (!g.periodic && (nDx < lox || nDx >= hix || nDy < loy || nDy >= hiy || nDz < loz || nDz >= hiz)) {
                escaped = 1;
            } else {
                int na = a, nin = in, ncell = cell;
                float nox = ox, noy = oy, noz = oz;
                float nbx = bx, nby = by, nbz = bz, nr2 = r2;
                const float tx = qx - bx, ty = qy - by, tz = qz - bz;
                if (tx * tx + ty * ty + tz * tz < r2) {
                    ++n_safe;
                } else {
                    ++n_class;
                    Locate L = locate_f32(g, na, qx, qy, qz, nox, noy, noz, iters, facets);
                    if (!L.ok) {
                        err = 1;
                        escaped = 1;
                    } else {
                        nin = L.inside;
                        const float a1 = g.cell_alpha[na];
                        if (fabsf(L.minm - a1) <= L.eps) {
                            ++rechecks;
                            if (!locate_f64(g, na, qx, qy, qz, nox, noy, noz, nin, facets)) {
                                err = 1;
                                escaped = 1;
                            }
                            L.minm = 0.f;
                        }
                        ncell = na;
                        nbx = qx;
                        nby = qy;
                        nbz = qz;
                        safe_ball(g, na, L, nin, nr2);
                    }
                }
                if (!escaped) {
                    int m = 0;
                    if (in != nin) m = 1;
                    else if (in && nin && cell != ncell) m = 2;
                    bool accept = true;
                    if (m > 0) {
                        ++crossings;
                        if (ku >= (m == 1 ? p.rej1 : p.rej2)) {
                            accept = false;
                            ++rejects;
                        }
                    }
                    if (accept) {
                        a = na;
                        in = nin;
                        cell = ncell;
                        px = qx;
                        py = qy;
                        pz = qz;
                        ox = nox;
                        oy = noy;
                        oz = noz;
                        bx = nbx;
                        by = nby;
                        bz = nbz;
                        r2 = nr2;
                    }
                }
            }
            Dx = ox + px;
            Dy = oy + py;
            Dz = oz + pz;
            // Kahan-compensated S += D
            {
                const float yx = Dx - Cx, tx = Sx + yx;
                Cx = (tx - Sx) - yx;
                Sx = tx;
                const float yy = Dy - Cy, ty = Sy + yy;
                Cy = (ty - Sy) - yy;
                Sy = ty;
                const float yz = Dz - Cz, tz = Sz + yz;
                Cz = (tz - Sz) - yz;
                Sz = tz;
            }
        }
        const int32_t s1 = step + 1;
        if (s1 % p.steps_per_h == 0) {
            const int j = s1 / p.steps_per_h;
            const long long c = wave_popc(live && !err && in);
            if ((threadIdx.x & 63) == 0 && c) atomicAdd((unsigned long long*)&fo.occupancy[j], (unsigned long long)c);
            if (next_ev < fo.ev.n && fo.ev.j[next_ev] == j) {
                // Y(t_j) = ts * (sum_{k<=n} D_k - D_n / 2), relative to the start point
                const float Yx = ts * ((Sx - Cx) - 0.5f * Dx);
                const float Yy = ts * ((Sy - Cy) - 0.5f * Dy);
                const float Yz = ts * ((Sz - Cz) - 0.5f * Dz);
                while (next_ev < fo.ev.n && fo.ev.j[next_ev] == j) {
                    const int q = fo.ev.pair[next_ev];
                    const float sg = (float)fo.ev.sign[next_ev];
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
    {
        const long long ne = wave_popc(live && escaped);
        const long long nr = wave_popc(live && err);
        if ((threadIdx.x & 63) == 0) {
            if (ne) atomicAdd((unsigned long long*)&fo.flags[MADI_FLAG_ESCAPED], (unsigned long long)ne);
            if (nr) atomicAdd((unsigned long long*)&fo.flags[MADI_FLAG_CLASSIFIER_ERR], (unsigned long long)nr);
        }
    }
    double* slot = fo.partial + (size_t)(fo.tile_base + blockIdx.x) * fo.plan->n_cols * 2;
    emit_columns<float>(*fo.plan, dM, live && !escaped, lds_cols, slot);
    if (counters) {
        const double v0 = wave_sum_f64((double)n_class), v1 = wave_sum_f64((double)n_safe),
                     v2 = wave_sum_f64((double)iters), v3 = wave_sum_f64((double)rechecks),
                     v4 = wave_sum_f64((double)facets), v5 = wave_sum_f64((double)crossings),
                     v6 = wave_sum_f64((double)rejects);
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&counters[MADI_CNT_CLASSIFY], (unsigned long long)v0);
            atomicAdd(&counters[MADI_CNT_SAFE_HIT], (unsigned long long)v1);
            atomicAdd(&counters[MADI_CNT_GRAPH_ITERS], (unsigned long long)v2);
            atomicAdd(&counters[MADI_CNT_FP64_RECHECK], (unsigned long long)v3);
            atomicAdd(&counters[MADI_CNT_FACETS], (unsigned long long)v4);
            atomicAdd(&counters[MADI_CNT_CROSSINGS], (unsigned long long)v5);
            atomicAdd(&counters[MADI_CNT_REJECTS], (unsigned long long)v6);
        }
    }
}

// Sum tile slots [tile_lo, tile_lo + n_tiles) per column in tile order.
__global__ void k_finalize_tiles(const double* __restrict__ partial, int64_t tile_lo, int64_t n_tiles, int32_t n_cols,
                                 double* __restrict__ cos_out, double* __restrict__ sin_out)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= n_cols) return;
    double a = 0.0, b = 0.0;
    for (int64_t t = 0; t < n_tiles; ++t) {
        const double* s = partial + ((size_t)(tile_lo + t) * n_cols + col) * 2;
        a += s[0];
        b += s[1];
    }
    cos_out[col] = a;
    sin_out[col] = b;
}

int launch_classify_graph(const GeomDev& g, const double* pts, int64_t n, int32_t* cell, int8_t* inside, int8_t* ok,
                          unsigned long long* counters, hipStream_t s)
{
    if (n == 0) return MADI_OK;
    int64_t blocks = (n + kBlock - 1) / kBlock;
    if (blocks > 65535) blocks = 65535;
    hipLaunchKernelGGL(k_classify_graph, dim3((unsigned)blocks), dim3(kBlock), 0, s, g, pts, n, cell, inside, ok,
                       counters);
    return check_launch("k_classify_graph");
}

int launch_walk_f32_fused(const GeomDev& g, const WalkDev& p, const Gf2Col* jump_pow, const FusedOut& fo,
                          unsigned long long* counters, hipStream_t s)
{
    if (p.n_walkers == 0) return MADI_OK;
    const int64_t blocks = (p.n_walkers + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_walk_f32, dim3((unsigned)blocks), dim3(kBlock), 0, s, g, p, jump_pow, fo, counters);
    return check_launch("k_walk_f32");
}

int launch_finalize_tiles(const double* partial, int64_t tile_lo, int64_t n_tiles, int32_t n_cols, double* cos_out,
                          double* sin_out, hipStream_t s)
{
    if (n_cols == 0) return MADI_OK;
    hipLaunchKernelGGL(k_finalize_tiles, dim3((n_cols + 127) / 128), dim3(128), 0, s, partial, tile_lo, n_tiles, n_cols,
                       cos_out, sin_out);
    return check_launch("k_finalize_tiles");
}

}  // namespace madi
