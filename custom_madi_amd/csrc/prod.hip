// Production walk: FP32 seed-relative state, Delaunay-graph classifier with
// FP64 certification, a certified plane tier, fused PGSE phase reduction.
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
// the most negative such k); at the nearest seed, inside <=> min m >= alpha_a.
// This is the reference's decision (madi/walker_gpu.py:272-335) restated on
// the Delaunay graph; an FP32 decision closer than its rounding bound to the
// alpha threshold is re-decided in FP64 for the same represented point.
//
// Each cell's edges are sorted by h.  Since m_j >= h_j - |p|, the edge loop
// stops as soon as h_j - |p| reaches the second-smallest margin seen so far
// (or a cap): no unvisited edge can change the decision or the bounds below.
//
// Plane tier.  A full classification at x records its binding plane (the
// edge a->k of smallest margin), the second-smallest margin m2 over a's
// other edges and n2 over k's other edges.  For a later point q within
// R_a = m2 - alpha_a of x on a's side of the plane (resp. R_k = n2 - alpha_k
// on k's side) every other plane keeps its margin above alpha, so the label
// follows from the one plane: s = h - q.u;  s >= 0 -> cell a, inside iff
// s >= alpha_a;  s < 0 -> cell k, inside iff -s >= alpha_k.  Membrane
// bounces and gap walkers crossing a bisector are answered by this tier.
#include "common.hpp"
#include "phase.hpp"
#include "rng.hpp"
#include "launch.hpp"

namespace madi {

constexpr float kUlp = 4.8e-7f;  // 8 * 2^-24: FP32 margin error per unit of |p|_1 + h

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

// Result of a full FP32 classification.
struct Refresh {
    int ok;
    int inside;
    int near_tie;   // smallest margin within rounding of 0: nearest seed ambiguous
    int uncertain;  // smallest margin within rounding of alpha: inside ambiguous
    int c;          // edge of the smallest margin (-1: none below the cap)
    float m1, m2;   // smallest and (capped) second-smallest margin
    float eps;
};

// Greedy location from anchor a plus the margins of a's edges.  p and o are
// moved with the anchor.
__device__ __forceinline__ Refresh locate_f32(const GeomDev& g, int& a, float& px, float& py, float& pz, float& ox,
                                             float& oy, float& oz, float cap, long long& iters, long long& facets)
{
    Refresh r{0, 0, 0, 0, -1, 0.f, 0.f, 0.f};
    for (int it = 0; it < kMaxGraphIters; ++it) {
        const int lo = g.nb_off[a], hi = g.nb_off[a + 1];
        const float pn = __builtin_sqrtf(px * px + py * py + pz * pz);
        const float eps = kUlp * (fabsf(px) + fabsf(py) + fabsf(pz) + g.cell_hmax[a]);
        float m1 = __builtin_inff(), m2 = __builtin_inff();
        int c = -1;
        for (int k = lo; k < hi; ++k) {
            const float4 e = g.nb_rec[k];
            if (e.w - pn >= fminf(m2, cap)) break;
            const float m = e.w - (px * e.x + py * e.y + pz * e.z);
            ++facets;
            if (m < m1) {
                m2 = m1;
                m1 = m;
                c = k;
            } else {
                m2 = fminf(m2, m);
            }
        }
        if (m1 < -eps) {
            const double ex = g.nb_e64[3 * c], ey = g.nb_e64[3 * c + 1], ez = g.nb_e64[3 * c + 2];
            px = (float)((double)px - ex);
            py = (float)((double)py - ey);
            pz = (float)((double)pz - ez);
            ox = (float)((double)ox + ex);
            oy = (float)((double)oy + ey);
            oz = (float)((double)oz + ez);
            a = g.nb_id[c];
            ++iters;
            continue;
        }
        const float a1 = g.cell_alpha[a];
        r.ok = 1;
        r.c = c;
        r.m1 = m1;
        r.m2 = fminf(m2, cap);
        r.eps = eps;
        r.near_tie = m1 <= eps;
        r.inside = m1 >= a1;
        r.uncertain = fabsf(m1 - a1) <= eps;
        return r;
    }
    return r;
}

// Plane-tier state of one walker (registers).
struct Tier {
    float ux, uy, uz, h;  // plane of edge a->k, unit normal towards k
    float cx, cy, cz;     // centre, relative to the anchor
    float ra2, rk2;       // squared validity radii on a's and on k's side (0 = none)
    float alpha_a, alpha_k;
    float eps;            // rounding bound of s over the validity ball
    int k;
};

__device__ __forceinline__ void tier_clear(Tier& t)
{
    t.ux = t.uy = t.uz = 0.f;
    t.h = 0.f;
    t.cx = t.cy = t.cz = 0.f;
    t.ra2 = t.rk2 = 0.f;
    t.alpha_a = t.alpha_k = 0.f;
    t.eps = 0.f;
    t.k = -1;
}

// Build the tier at p (anchor a, located by r); k-side radius from k's edges.
__device__ __forceinline__ void tier_build(const GeomDev& g, int a, float px, float py, float pz, const Refresh& r,
                                           float cap, Tier& t, long long& facets)
{
    const float a1 = g.cell_alpha[a];
    t.cx = px;
    t.cy = py;
    t.cz = pz;
    t.alpha_a = a1;
    if (r.c < 0) {
        // no edge below the cap: every margin is at least cap
        t.ux = t.uy = t.uz = 0.f;
        t.h = cap;
        t.k = -1;
        t.alpha_k = 0.f;
        const float ra = cap - a1 - 2.f * r.eps;
        t.ra2 = ra > 0.f ? ra * ra : 0.f;
        t.rk2 = 0.f;
        t.eps = r.eps + kUlp * cap;
        return;
    }
    const float4 e = g.nb_rec[r.c];
    t.ux = e.x;
    t.uy = e.y;
    t.uz = e.z;
    t.h = e.w;
    const int k = g.nb_id[r.c];
    t.k = k;
    const float ak = g.cell_alpha[k];
    t.alpha_k = ak;
    const float ra = r.m2 - fmaxf(a1, 0.f) - 2.f * r.eps;
    // k-side: margins of p with respect to k's edges, excluding the shared plane
    const float qx = (float)((double)px - g.nb_e64[3 * r.c]);
    const float qy = (float)((double)py - g.nb_e64[3 * r.c + 1]);
    const float qz = (float)((double)pz - g.nb_e64[3 * r.c + 2]);
    const float qn = __builtin_sqrtf(qx * qx + qy * qy + qz * qz);
    const float epsk = kUlp * (fabsf(qx) + fabsf(qy) + fabsf(qz) + g.cell_hmax[k]);
    float n2 = cap;
    const int lo = g.nb_off[k], hi = g.nb_off[k + 1];
    for (int j = lo; j < hi; ++j) {
        const float4 f = g.nb_rec[j];
        if (f.w - qn >= n2) break;
        if (f.w == e.w && f.x == -e.x && f.y == -e.y && f.z == -e.z && g.nb_id[j] == a) continue;
        const float n = f.w - (qx * f.x + qy * f.y + qz * f.z);
        ++facets;
        n2 = fminf(n2, n);
    }
    const float rk = n2 - fmaxf(ak, 0.f) - 2.f * fmaxf(r.eps, epsk);
    t.ra2 = ra > 0.f ? ra * ra : 0.f;
    t.rk2 = rk > 0.f ? rk * rk : 0.f;
    const float rmax = fmaxf(ra, rk);
    t.eps = fmaxf(r.eps, epsk) + kUlp * (1.8f * fmaxf(rmax, 0.f) + e.w);
}

// Tier answer for q (relative to the anchor); returns 1 when certified.
__device__ __forceinline__ int tier_query(const Tier& t, int a, float qx, float qy, float qz, int& cell, int& inside)
{
    const float dx = qx - t.cx, dy = qy - t.cy, dz = qz - t.cz;
    const float d2 = dx * dx + dy * dy + dz * dz;
    const float s = t.h - (qx * t.ux + qy * t.uy + qz * t.uz);
    if (s >= 0.f) {
        if (!(d2 < t.ra2) || fabsf(s - t.alpha_a) <= t.eps) return 0;
        cell = a;
        inside = s >= t.alpha_a;
        return 1;
    }
    if (!(d2 < t.rk2) || fabsf(-s - t.alpha_k) <= t.eps || -s <= t.eps) return 0;
    cell = t.k;
    inside = -s >= t.alpha_k;
    return 1;
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

// Anchor a world point: hint seed and the point relative to it.
__device__ __forceinline__ void anchor_world(const GeomDev& g, double x, double y, double z, int& a, float& px,
                                             float& py, float& pz)
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
}

__global__ __launch_bounds__(kBlock) void k_classify_graph(GeomDev g, const double* __restrict__ pts, int64_t n,
                                                          int32_t* __restrict__ cell, int8_t* __restrict__ inside,
                                                          int8_t* __restrict__ ok, unsigned long long* counters)
{
    long long iters = 0, facets = 0, rechecks = 0;
    const float cap = __builtin_inff();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (g.free_water) {
            cell[i] = 0;
            inside[i] = 0;
            ok[i] = 1;
            continue;
        }
        int a, in = 0;
        float px, py, pz;
        anchor_world(g, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], a, px, py, pz);
        float ox = -px, oy = -py, oz = -pz;
        Refresh r = locate_f32(g, a, px, py, pz, ox, oy, oz, cap, iters, facets);
        int k = r.ok;
        in = r.inside;
        if (k && (r.uncertain || r.near_tie)) {
            ++rechecks;
            k = locate_f64(g, a, px, py, pz, ox, oy, oz, in, facets);
        }
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

// Full classification of q from anchor a, certified, with a new tier.
// Returns 0 on a classifier failure.
__device__ __forceinline__ int classify_step(const GeomDev& g, int& a, float& qx, float& qy, float& qz, float& ox,
                                             float& oy, float& oz, float cap, int& cell, int& inside, Tier& t,
                                             long long& iters, long long& facets, long long& rechecks)
{
    Refresh r = locate_f32(g, a, qx, qy, qz, ox, oy, oz, cap, iters, facets);
    if (!r.ok) return 0;
    cell = a;
    inside = r.inside;
    if (r.uncertain) {
        ++rechecks;
        if (!locate_f64(g, a, qx, qy, qz, ox, oy, oz, inside, facets)) return 0;
        cell = a;
        tier_clear(t);
        return 1;
    }
    tier_build(g, a, qx, qy, qz, r, cap, t, facets);
    return 1;
}

struct TileOut {
    double* slot;          // n_cols x 2 doubles of this tile
    long long* occupancy;  // n_grid
    long long* flags;      // MADI_N_FLAGS
};

constexpr int kMaxLdsGrid = 256;  // occupancy counters kept in LDS up to this many Y-grid times

enum LanePhase : int { kStep = 0, kScanA = 1, kScanK = 2, kDone = 3 };

__device__ __forceinline__ void occ_add(bool in_lds, int* lds_occ, long long* occ, int j)
{
    if (in_lds) atomicAdd(&lds_occ[j], 1);
    else atomicAdd((unsigned long long*)&occ[j], 1ull);
}

// Tier from a finished two-sided scan: plane c of anchor a, a-side bound m2a,
// k-side bound n2 (both capped), rounding bounds epsa / epsk, centre q.
__device__ __forceinline__ void tier_from_scans(const GeomDev& g, int a, int c, float qx, float qy, float qz,
                                                float m2a, float epsa, float n2, float epsk, Tier& t)
{
    const float4 e = g.nb_rec[c];
    const int k = g.nb_id[c];
    const float aa = g.cell_alpha[a], ak = g.cell_alpha[k];
    t.ux = e.x;
    t.uy = e.y;
    t.uz = e.z;
    t.h = e.w;
    t.cx = qx;
    t.cy = qy;
    t.cz = qz;
    t.k = k;
    t.alpha_a = aa;
    t.alpha_k = ak;
    const float ep = fmaxf(epsa, epsk);
    const float ra = m2a - fmaxf(aa, 0.f) - 2.f * epsa;
    const float rk = n2 - fmaxf(ak, 0.f) - 2.f * ep;
    t.ra2 = ra > 0.f ? ra * ra : 0.f;
    t.rk2 = rk > 0.f ? rk * rk : 0.f;
    t.eps = ep + kUlp * (1.8f * fmaxf(fmaxf(ra, rk), 0.f) + e.w);
}

// Tier when no edge lies below the cap: every margin is at least cap.
__device__ __forceinline__ void tier_no_plane(const GeomDev& g, int a, float qx, float qy, float qz, float cap,
                                              float epsa, Tier& t)
{
    const float aa = g.cell_alpha[a];
    t.ux = t.uy = t.uz = 0.f;
    t.h = cap;
    t.cx = qx;
    t.cy = qy;
    t.cz = qz;
    t.k = -1;
    t.alpha_a = aa;
    t.alpha_k = 0.f;
    const float ra = cap - aa - 2.f * epsa;
    t.ra2 = ra > 0.f ? ra * ra : 0.f;
    t.rk2 = 0.f;
    t.eps = epsa + kUlp * cap;
}

// One tile of kBlock walkers [w_lo, w_lo + kBlock) of an entry: the whole
// walk, then the phase columns into the tile's slot.  All threads of the
// block must call it (it synchronises on LDS).  NP >= plan.n_pairs bounds
// the moment registers.
//
// Each lane runs its walker as a small state machine, so lanes of a wave need
// not be at the same step: per iteration a lane either takes its next step
// (draws, proposal, plane tier) or advances its pending classification by E
// edges (kScanA: greedy location and a-side margins; kScanK: k-side margins
// for the new tier).  A step completes when its label is known.  The wave's
// cost per iteration is bounded by one step plus E edge evaluations, instead
// of the longest classification among its 64 lanes.  Every walker still
// consumes its own random stream in order, so results do not depend on E.
template <int NP, int E>
__device__ __forceinline__ void walk_tile_f32(const GeomDev& g, const WalkDev& p, const Gf2Col* __restrict__ jump_pow,
                                              int64_t w_lo, const PlanDev& plan, const Events& ev, const TileOut& out,
                                              float cap_width, unsigned long long* counters, double* lds_cols,
                                              int* lds_occ)
{
    const int64_t w = w_lo + threadIdx.x;
    const bool live = w < p.n_walkers;
    long long n_class = 0, n_tier = 0, iters = 0, facets = 0, rechecks = 0, crossings = 0, rejects = 0;
    const float cap = g.alpha_max + cap_width;
    const bool occ_in_lds = p.n_grid <= kMaxLdsGrid;
    if (occ_in_lds)
        for (int j = threadIdx.x; j < p.n_grid; j += blockDim.x) lds_occ[j] = 0;
    __syncthreads();

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
    Tier t;
    tier_clear(t);
    if (live && !g.free_water) {
        anchor_world(g, x0, y0, z0, a, px, py, pz);
        ox = -px;
        oy = -py;
        oz = -pz;
        ++n_class;
        if (!classify_step(g, a, px, py, pz, ox, oy, oz, cap, cell, in, t, iters, facets, rechecks)) {
            err = 1;
            escaped = 1;
        }
        ox = -px;
        oy = -py;
        oz = -pz;
    } else if (live) {
        // free water: a tier that answers "outside" everywhere
        t.h = 1.f;
        t.alpha_a = 2.f;
        t.ra2 = __builtin_inff();
        t.eps = 0.f;
    }
    const float lox = (float)(-x0), loy = (float)(-y0), loz = (float)(-z0);
    const float hix = (float)(g.L - x0), hiy = (float)(g.L - y0), hiz = (float)(g.L - z0);

    float Sx = 0.f, Sy = 0.f, Sz = 0.f, Cx = 0.f, Cy = 0.f, Cz = 0.f;
    float dM[NP][3];
#pragma unroll
    for (int q = 0; q < NP; ++q) dM[q][0] = dM[q][1] = dM[q][2] = 0.f;
    if (live && !err && in) occ_add(occ_in_lds, lds_occ, out.occupancy, 0);

    const float sig = (float)p.sigma;
    const float ts = (float)p.ts;
    int phase = (live && !escaped && p.n_steps > 0) ? kStep : kDone;
    int step = 0, until_sample = p.steps_per_h, j_sample = 0, next_ev = 0;
    // pending step: proposal (relative to the scan anchor), offset, reject bits
    float qx = 0.f, qy = 0.f, qz = 0.f, nox = 0.f, noy = 0.f, noz = 0.f;
    int na = 0, rj = 0;
    int pin = 0, pcell = 0;  // label of the pending proposal, once decided
    // scan state
    int cur = 0, end = 0, c = -1;
    float m1 = 0.f, m2 = 0.f, pn = 0.f, eps = 0.f;
    float kx = 0.f, ky = 0.f, kz = 0.f;
    int sa_c = -1;
    float sa_m2 = 0.f, sa_eps = 0.f;

    while (__any(phase != kDone)) {
        int complete = 0, tier_mode = 0;  // tier on accept: 0 keep, 1 two-sided, 2 clear, 3 no plane
        if (phase == kStep) {
            const float dx = normal_f32(rng) * sig;
            const float dy = normal_f32(rng) * sig;
            const float dz = normal_f32(rng) * sig;
            const uint64_t ku = rng_next(rng) >> 11;
            rj = (ku >= p.rej1 ? 1 : 0) | (ku >= p.rej2 ? 2 : 0);
            qx = px + dx;
            qy = py + dy;
            qz = pz + dz;
            nox = ox;
            noy = oy;
            noz = oz;
            na = a;
            const float nDx = ox + qx, nDy = oy + qy, nDz = oz + qz;
            if (!g.periodic && (nDx < lox || nDx >= hix || nDy < loy || nDy >= hiy || nDz < loz || nDz >= hiz)) {
                escaped = 1;
                phase = kDone;
                // samples after an escape keep the frozen label
                for (int j = j_sample + 1; j < p.n_grid; ++j)
                    if (in) occ_add(occ_in_lds, lds_occ, out.occupancy, j);
            } else if (tier_query(t, a, qx, qy, qz, pcell, pin)) {
                ++n_tier;
                complete = 1;
            } else {
                ++n_class;
                cur = g.nb_off[na];
                end = g.nb_off[na + 1];
                m1 = m2 = __builtin_inff();
                c = -1;
                pn = __builtin_sqrtf(qx * qx + qy * qy + qz * qz);
                eps = kUlp * (fabsf(qx) + fabsf(qy) + fabsf(qz) + g.cell_hmax[na]);
                phase = kScanA;
            }
        }
        if (phase == kScanA || phase == kScanK) {
            const bool sa = phase == kScanA;
            const float sx = sa ? qx : kx, sy = sa ? qy : ky, sz = sa ? qz : kz;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                if (cur < end) {
                    const float4 r = g.nb_rec[cur];
                    if (r.w - pn >= fminf(m2, cap)) {
                        cur = end;
                    } else {
                        const float m = r.w - (sx * r.x + sy * r.y + sz * r.z);
                        ++facets;
                        if (m < m1) {
                            m2 = m1;
                            m1 = m;
                            c = cur;
                        } else {
                            m2 = fminf(m2, m);
                        }
                        ++cur;
                    }
                }
            }
            if (cur >= end) {
                if (sa) {
                    if (m1 < -eps) {
                        // a neighbour is nearer: move the anchor and rescan
                        const double ex = g.nb_e64[3 * c], ey = g.nb_e64[3 * c + 1], ez = g.nb_e64[3 * c + 2];
                        qx = (float)((double)qx - ex);
                        qy = (float)((double)qy - ey);
                        qz = (float)((double)qz - ez);
                        nox = (float)((double)nox + ex);
                        noy = (float)((double)noy + ey);
                        noz = (float)((double)noz + ez);
                        na = g.nb_id[c];
                        ++iters;
                        cur = g.nb_off[na];
                        end = g.nb_off[na + 1];
                        m1 = m2 = __builtin_inff();
                        c = -1;
                        pn = __builtin_sqrtf(qx * qx + qy * qy + qz * qz);
                        eps = kUlp * (fabsf(qx) + fabsf(qy) + fabsf(qz) + g.cell_hmax[na]);
                    } else {
                        const float a1 = g.cell_alpha[na];
                        pin = m1 >= a1;
                        pcell = na;
                        if (fabsf(m1 - a1) <= eps) {
                            ++rechecks;
                            if (!locate_f64(g, na, qx, qy, qz, nox, noy, noz, pin, facets)) {
                                err = 1;
                                escaped = 1;
                                phase = kDone;
                            } else {
                                pcell = na;
                                complete = 1;
                                tier_mode = 2;
                            }
                        } else if (c < 0) {
                            sa_eps = eps;
                            complete = 1;
                            tier_mode = 3;
                        } else {
                            sa_c = c;
                            sa_m2 = fminf(m2, cap);
                            sa_eps = eps;
                            const int k = g.nb_id[c];
                            kx = (float)((double)qx - g.nb_e64[3 * c]);
                            ky = (float)((double)qy - g.nb_e64[3 * c + 1]);
                            kz = (float)((double)qz - g.nb_e64[3 * c + 2]);
                            cur = g.nb_off[k];
                            end = g.nb_off[k + 1];
                            m1 = m2 = __builtin_inff();
                            pn = __builtin_sqrtf(kx * kx + ky * ky + kz * kz);
                            eps = kUlp * (fabsf(kx) + fabsf(ky) + fabsf(kz) + g.cell_hmax[k]);
                            phase = kScanK;
                        }
                    }
                } else {
                    // k-side scan done: its smallest margin is at most the shared
                    // plane's, so its second smallest bounds every other k plane
                    complete = 1;
                    tier_mode = 1;
                }
            }
        }
        if (complete) {
            int m = 0;
            if (in != pin) m = 1;
            else if (in && pin && cell != pcell) m = 2;
            bool accept = true;
            if (m > 0) {
                ++crossings;
                if ((rj >> (m - 1)) & 1) {
                    accept = false;
                    ++rejects;
                }
            }
            if (accept) {
                if (tier_mode == 1) tier_from_scans(g, na, sa_c, qx, qy, qz, sa_m2, sa_eps, fminf(m2, cap), eps, t);
                else if (tier_mode == 2) tier_clear(t);
                else if (tier_mode == 3) tier_no_plane(g, na, qx, qy, qz, cap, sa_eps, t);
                a = na;
                in = pin;
                cell = pcell;
                px = qx;
                py = qy;
                pz = qz;
                ox = nox;
                oy = noy;
                oz = noz;
            }
            const float Dx = ox + px, Dy = oy + py, Dz = oz + pz;
            const float yx = Dx - Cx, tx = Sx + yx;
            Cx = (tx - Sx) - yx;
            Sx = tx;
            const float yy = Dy - Cy, ty = Sy + yy;
            Cy = (ty - Sy) - yy;
            Sy = ty;
            const float yz = Dz - Cz, tz = Sz + yz;
            Cz = (tz - Sz) - yz;
            Sz = tz;
            ++step;
            if (--until_sample == 0) {
                until_sample = p.steps_per_h;
                const int j = ++j_sample;
                if (in) occ_add(occ_in_lds, lds_occ, out.occupancy, j);
                if (next_ev < ev.n && ev.j[next_ev] == j) {
                    // Y(t_j) = ts * (sum_{k<=n} D_k - D_n / 2), relative to the start point
                    const float Yx = ts * ((Sx - Cx) - 0.5f * Dx);
                    const float Yy = ts * ((Sy - Cy) - 0.5f * Dy);
                    const float Yz = ts * ((Sz - Cz) - 0.5f * Dz);
                    while (next_ev < ev.n && ev.j[next_ev] == j) {
                        const int q = ev.pair[next_ev];
                        const float sg = (float)ev.sign[next_ev];
#pragma unroll
                        for (int u = 0; u < NP; ++u)
                            if (u == q) {
                                dM[u][0] += sg * Yx;
                                dM[u][1] += sg * Yy;
                                dM[u][2] += sg * Yz;
                            }
                        ++next_ev;
                    }
                }
            }
            phase = step >= p.n_steps ? kDone : kStep;
        }
    }
    {
        const long long ne = wave_popc(live && escaped);
        const long long nr = wave_popc(live && err);
        if ((threadIdx.x & 63) == 0) {
            if (ne) atomicAdd((unsigned long long*)&out.flags[MADI_FLAG_ESCAPED], (unsigned long long)ne);
            if (nr) atomicAdd((unsigned long long*)&out.flags[MADI_FLAG_CLASSIFIER_ERR], (unsigned long long)nr);
        }
    }
    __syncthreads();
    if (occ_in_lds)
        for (int j = threadIdx.x; j < p.n_grid; j += blockDim.x)
            if (lds_occ[j]) atomicAdd((unsigned long long*)&out.occupancy[j], (unsigned long long)lds_occ[j]);
    emit_columns<float, NP>(plan, dM, live && !escaped, lds_cols, out.slot);
    if (counters) {
        const double v0 = wave_sum_f64((double)n_class), v1 = wave_sum_f64((double)n_tier),
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

// One entry: block b runs tile b.
template <int NP, int E>
__global__ __launch_bounds__(kBlock) void k_walk_f32(GeomDev g, WalkDev p, const Gf2Col* __restrict__ jump_pow,
                                                    FusedOut fo, float cap_width, unsigned long long* counters)
{
    __shared__ double lds_cols[kWaves * kColChunk * 2];
    __shared__ int lds_occ[kMaxLdsGrid];
    TileOut out;
    out.slot = fo.partial + (size_t)(fo.tile_base + blockIdx.x) * fo.plan->n_cols * 2;
    out.occupancy = fo.occupancy;
    out.flags = fo.flags;
    walk_tile_f32<NP, E>(g, p, jump_pow, (int64_t)blockIdx.x * kBlock, *fo.plan, fo.ev, out, cap_width, counters,
                         lds_cols, lds_occ);
}

// Persistent work-stealing launch over many entries: each workgroup takes
// the next global tile index from an atomic counter, finds its work item by
// binary search on work_begin, and runs that tile.  Every workgroup exits
// once the counter passes the total, so the grid drains.
template <int NP, int E>
__global__ __launch_bounds__(kBlock) void k_batch_f32(const WorkDev* __restrict__ work, int32_t n_work, int64_t total,
                                                     unsigned long long* queue, const PlanDev* __restrict__ plan,
                                                     Events ev, const Gf2Col* __restrict__ jump_pow, float cap_width,
                                                     unsigned long long* counters)
{
    __shared__ double lds_cols[kWaves * kColChunk * 2];
    __shared__ int lds_occ[kMaxLdsGrid];
    __shared__ long long s_tile;
    __shared__ int s_item;
    for (;;) {
        if (threadIdx.x == 0) {
            const long long t = (long long)atomicAdd(queue, 1ull);
            int lo = 0, hi = n_work - 1;
            if (t < total) {
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (work[mid].work_begin <= t) lo = mid;
                    else hi = mid - 1;
                }
            }
            s_tile = t;
            s_item = lo;
        }
        __syncthreads();
        const long long t = s_tile;
        const int e = s_item;
        __syncthreads();
        if (t >= total) break;
        const WorkDev& W = work[e];
        const int64_t tile = W.tile_lo + (t - W.work_begin);
        TileOut out;
        out.slot = W.partial + (size_t)tile * plan->n_cols * 2;
        out.occupancy = W.occupancy;
        out.flags = W.flags;
        walk_tile_f32<NP, E>(W.g, W.w, jump_pow, tile * kBlock, *plan, ev, out, cap_width, counters, lds_cols,
                             lds_occ);
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

#define MADI_DISPATCH_NP_E(NP_VAL, E_VAL, LAUNCH)                  \
    do {                                                           \
        const int np_ = (NP_VAL), e_ = (E_VAL);                    \
        if (np_ <= 2) {                                            \
            if (e_ <= 4) LAUNCH(2, 4);                             \
            else if (e_ <= 8) LAUNCH(2, 8);                        \
            else LAUNCH(2, 16);                                    \
        } else if (np_ <= 4) {                                     \
            if (e_ <= 4) LAUNCH(4, 4);                             \
            else if (e_ <= 8) LAUNCH(4, 8);                        \
            else LAUNCH(4, 16);                                    \
        } else {                                                   \
            if (e_ <= 4) LAUNCH(MADI_MAX_PAIRS, 4);                \
            else if (e_ <= 8) LAUNCH(MADI_MAX_PAIRS, 8);           \
            else LAUNCH(MADI_MAX_PAIRS, 16);                       \
        }                                                          \
    } while (0)

int launch_walk_f32_fused(const GeomDev& g, const WalkDev& p, const Gf2Col* jump_pow, const FusedOut& fo,
                          float cap_width, int scan_chunk, unsigned long long* counters, hipStream_t s)
{
    if (p.n_walkers == 0) return MADI_OK;
    const int64_t blocks = (p.n_walkers + kBlock - 1) / kBlock;
#define MADI_LAUNCH_WALK(NPV, EV)                                                                                   \
    hipLaunchKernelGGL((k_walk_f32<NPV, EV>), dim3((unsigned)blocks), dim3(kBlock), 0, s, g, p, jump_pow, fo,     \
                       cap_width, counters)
    MADI_DISPATCH_NP_E(fo.ev.n_pairs, scan_chunk, MADI_LAUNCH_WALK);
#undef MADI_LAUNCH_WALK
    return check_launch("k_walk_f32");
}

int launch_batch_f32(const WorkDev* work, int32_t n_work, int64_t total, unsigned long long* queue, const PlanDev* plan,
                     const Events& ev, const Gf2Col* jump_pow, float cap_width, int scan_chunk, int32_t grid,
                     unsigned long long* counters, hipStream_t s)
{
    if (total == 0) return MADI_OK;
#define MADI_LAUNCH_BATCH(NPV, EV)                                                                                  \
    hipLaunchKernelGGL((k_batch_f32<NPV, EV>), dim3(grid), dim3(kBlock), 0, s, work, n_work, total, queue, plan, ev, \
                       jump_pow, cap_width, counters)
    MADI_DISPATCH_NP_E(ev.n_pairs, scan_chunk, MADI_LAUNCH_BATCH);
#undef MADI_LAUNCH_BATCH
    return check_launch("k_batch_f32");
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
