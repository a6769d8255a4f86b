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
// stops as soon as h_j - |p| reaches the third-smallest margin seen so far
// (k's side: the second; or a cap): no unvisited edge can change the
// decision or the bounds below (scan_chunk).
//
// Plane tier.  A full classification at x records a's three planes of
// smallest margin (the binding plane a->k first), a lower bound r of a's
// other margins and, for an outside point, n2 over k's other edges.  Near x
// the label follows from those planes: s = h - q.u on the binding plane;
// s >= 0 (and the other two planes non-negative) -> cell a, inside iff every
// plane's margin is at least alpha_a; s < 0 -> cell k, inside iff
// -s >= alpha_k; with validity radii r (outside answers) and r - alpha_a
// (inside answers) on a's side, n2 and n2 - alpha_k on k's side
// (tier_query).  Membrane bounces and gap walkers crossing a bisector are
// answered by this tier.
#include "common.hpp"
#include "phase.hpp"
#include "rng.hpp"
#include "launch.hpp"

// The exact FP64 KD classifier, for the verification build of the split
// walk: compiled without FP contraction, so its arithmetic keeps the
// operation order of madi_classify's (parity.hip is built that way too).
#pragma clang fp contract(off)
#include "kd_exact.hpp"
#pragma clang fp contract(fast)

namespace madi {

constexpr float kUlp = 4.8e-7f;  // 8 * 2^-24: FP32 margin error per unit of |p|_1 + h

// FP64 decision for the point anchor + p (p relative to the anchor).  Moves
// the anchor to the FP64 nearest seed; returns 0 if the greedy walk does not
// settle or two seeds coincide.
template <typename Count>
__device__ int locate_f64(const GeomDev& g, int& a, float& px, float& py, float& pz, float& ox, float& oy, float& oz,
                          int& inside, Count& facets)
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

// Plane-tier state of one walker (registers).  The centre is the point where
// the tier was built, relative to the anchor.
struct Tier {
    float ux, uy, uz, h;           // plane of edge a->k, unit normal towards k
    float ux2, uy2, uz2, h2;       // a's planes of second- and third-smallest margin
    float ux3, uy3, uz3, h3;       // (absent: 0, 0, 0, h >= cap, never crossed)
    float cx, cy, cz;              // centre
    float ra_in2, ra_out2;         // squared radii on a's side: inside / outside answers
    float rk_in2, rk_out2;         // the same on k's side (0 = no k-side answers)
    float rout2;                   // squared radius of a ball that is outside every contracted cell
    float alpha_a, alpha_k;
    float eps;                     // rounding bound of s over the validity ball
    int k;
    int a_off, a_end;              // the anchor's edge list (a_off < 0: not known)
    float a_hmax;                  // and its largest half length
};

__device__ __forceinline__ void tier_clear(Tier& t)
{
    t.ux = t.uy = t.uz = 0.f;
    t.h = 0.f;
    t.ux2 = t.uy2 = t.uz2 = 0.f;
    t.h2 = 0.f;
    t.ux3 = t.uy3 = t.uz3 = 0.f;
    t.h3 = 0.f;
    t.cx = t.cy = t.cz = 0.f;
    t.ra_in2 = t.ra_out2 = t.rk_in2 = t.rk_out2 = t.rout2 = 0.f;
    t.alpha_a = t.alpha_k = 0.f;
    t.eps = 0.f;
    t.k = -1;
    t.a_off = -1;
    t.a_end = 0;
    t.a_hmax = 0.f;
}

// Free water: every point is outside.
__device__ __forceinline__ void tier_free_water(Tier& t)
{
    tier_clear(t);
    t.h = 1.f;
    t.alpha_a = 2.f;
    t.ra_out2 = __builtin_inff();
}

// |q|_1 bound below which tier_query may certify an exit of an inside
// walker's cell, 0 for none.  In a periodic box 0.49 L keeps the anchor's
// seed nearer to q than any periodic image of it.  Exits past the tier's
// ball need a step longer than the annulus, rare once every annulus is 2.5
// step deviations or more: the certificate is then not worth its test.
__device__ __forceinline__ float exit_bound(const GeomDev& g, const WalkDev& p)
{
    if (!exit_certificate_pays(g.alpha_max, p.sigma)) return 0.f;
    return g.periodic ? (float)(0.49 * g.L) : __builtin_inff();
}

// Tier answer for q (relative to the anchor a); returns 1 when certified.
//
// Let x be the centre, m1 <= m2 <= m3 <= m4 the smallest margins of a's
// planes at x (planes 1 = a->k, 2 and 3 kept in the tier; the radii use a
// lower bound r of m4, of every margin but the three smallest) and n2 the
// second-smallest margin of k's planes at x (all capped).  Margins change by
// at most |q - x|, so for |q - x| < r every plane of a other than planes 1-3
// keeps a positive margin: q lies in a's Voronoi cell iff s_i = h_i - q.u_i
// >= 0 for i = 1, 2, 3, and is then outside the contracted cell iff
// min(s1, s2, s3) < alpha_a; inside needs every other plane at alpha_a or
// more as well, i.e. |q - x| < r - alpha_a.  Across plane 1 the same holds
// on k's side with n2 and alpha_k (q is in Vor(k) once it is on k's side of
// the a-k bisector and within n2 of x, whatever a's other planes say); across
// plane 2 or 3 alone there is no answer.  An outside answer has the larger
// radius r (n2): walkers in the extracellular gap, whose next plane is often
// closer than alpha + a step, keep their tier much longer.
//
// The outside ball (rout2): every contracted cell j is the Voronoi cell of j
// eroded by alpha_j, so it lies at least alpha_j + dist(x, Vor(j)) from the
// centre x.  The neighbour k's is thus at least alpha_k + m1 away and a's
// own at least alpha_a - m1; the ball of radius min(m2, n2) around x lies in
// Vor(a) u Vor(k) (no other plane of a or k is crossed within it), so every
// third cell is at least alpha_min + min(m2, n2) away (tier_make).  A point
// within the smallest of these distances of x is outside every cell,
// whichever seed is nearest; near Voronoi edges, where m2 is small, this
// ball is the larger certificate.
//
// Rounding: a decision closer than eps to 0 or alpha is left to a full
// classification, except that both sides of a bisector are outside there
// (alpha > eps on each side), which needs no decision.
__device__ __forceinline__ int tier_query(const Tier& t, int a, float qx, float qy, float qz, int& cell, int& inside)
{
    const float dx = qx - t.cx, dy = qy - t.cy, dz = qz - t.cz;
    const float d2 = dx * dx + dy * dy + dz * dz;
    if (d2 < t.rout2) {
        cell = a;  // the cell of an outside label is not used
        inside = 0;
        return 1;
    }
    const float s = t.h - (qx * t.ux + qy * t.uy + qz * t.uz);
    if (s >= 0.f) {
        const float s2 = t.h2 - (qx * t.ux2 + qy * t.uy2 + qz * t.uz2);
        const float s3 = t.h3 - (qx * t.ux3 + qy * t.uy3 + qz * t.uz3);
        const float s23 = fminf(s2, s3);
        const float sm = fminf(s, s23);
        const int ins = sm >= t.alpha_a;
        if (s23 < 0.f || fabsf(sm - t.alpha_a) <= t.eps || !(d2 < (ins ? t.ra_in2 : t.ra_out2))) return 0;
        cell = a;
        inside = ins;
        return 1;
    }
    const float r = -s;
    const int ins = r >= t.alpha_k;
    if (r <= t.eps || fabsf(r - t.alpha_k) <= t.eps || !(d2 < (ins ? t.rk_in2 : t.rk_out2))) return 0;
    cell = t.k;
    inside = ins;
    return 1;
}

// Exit certificate for a proposal q the tier does not label, from a walker
// inside its anchor cell a whose draw rejects both crossing kinds (rj == 3):
// q lies outside a's contracted cell when plane 1 (on a's side: plane 1, 2
// or 3) has a margin below alpha_a by more than the rounding bound at q (the
// centre's plus kUlp * |q - x|_1).  The planes are edges of a when k != a:
// a missing second or third edge leaves its plane at (0, h >= cap), which
// never passes.  |q|_1 < leave (exit_bound) keeps q nearer to a's seed than to any
// periodic image of it, so q lies in another cell or outside a's contracted
// one: the walker crosses one membrane or two, and rj == 3 rejects either.
// The caller then completes the step as a rejected crossing, the outcome
// its true label gives.
__device__ __forceinline__ bool tier_exit(const Tier& t, int a, float qx, float qy, float qz, float leave)
{
    const float s = t.h - (qx * t.ux + qy * t.uy + qz * t.uz);
    const float s2 = t.h2 - (qx * t.ux2 + qy * t.uy2 + qz * t.uz2);
    const float s3 = t.h3 - (qx * t.ux3 + qy * t.uy3 + qz * t.uz3);
    const float sm = s >= 0.f ? fminf(s, fminf(s2, s3)) : s;
    const float lim =
        t.alpha_a - fmaf(kUlp * 1.125f, fabsf(qx - t.cx) + fabsf(qy - t.cy) + fabsf(qz - t.cz), t.eps);
    return t.k != a && sm < lim && fabsf(qx) + fabsf(qy) + fabsf(qz) < leave;
}

// Hint-grid anchor for a world point.
__device__ __forceinline__ int hint_anchor(const GeomDev& g, double x, double y, double z)
{
    const int n = g.hint_n;
    int ix = (int)floor(x / g.hint_h), iy = (int)floor(y / g.hint_h), iz = (int)floor(z / g.hint_h);
    ix = ix < 0 ? 0 : (ix >= n ? n - 1 : ix);
    iy = iy < 0 ? 0 : (iy >= n ? n - 1 : iy);
    iz = iz < 0 ? 0 : (iz >= n ? n - 1 : iz);
    return gload(g.hint + (ix * n + iy) * n + iz);
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
    px = (float)(x - gload(c));
    py = (float)(y - gload(c + 1));
    pz = (float)(z - gload(c + 2));
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

// Box-Muller first output in FP32, scaled by sigma, from the top 32 bits of
// the two draws the FP64 walk turns into its uniforms (u = hi * 2^-32, within
// 2^-24 of them; u1 = 0 is replaced by 2^-33):
// sqrt(c * log2(u1)) * cos(2 pi u2) with c = -2 sigma^2 ln 2, log2(u1) taken
// as log2(hi) - 32, and the hardware cosine in revolutions.
// Round-to-nearest u32 -> f32 as the one conversion instruction (the
// compiler widens a conversion of the high half of a 64-bit value into a
// 64-bit-to-float sequence).
__device__ __forceinline__ float u32_to_f32(uint32_t v)
{
    float f;
    asm("v_cvt_f32_u32 %0, %1" : "=v"(f) : "v"(v));
    return f;
}

// max(v, 0.5) as one instruction for a converted integer v (never NaN; the
// compiler's fmaxf would first canonicalise v with a second v_max).
__device__ __forceinline__ float max_half(float v)
{
    float m;
    asm("v_max_f32 %0, 0.5, %1" : "=v"(m) : "v"(v));
    return m;
}

__device__ __forceinline__ float normal_scaled_f32(Rng& r, float c)
{
    const float h1 = max_half(u32_to_f32((uint32_t)(rng_next(r) >> 32)));
    const float u2 = u32_to_f32((uint32_t)(rng_next(r) >> 32)) * 0x1.0p-32f;
    return __builtin_amdgcn_sqrtf(fmaf(c, __builtin_amdgcn_logf(h1), -32.f * c)) * __builtin_amdgcn_cosf(u2);
}

// Running state of one facet scan: margins of the point s (relative to seed
// `cell`) over the cell's edges in order of half length.  cur < 0: the scan
// of (cell, s) is set up by scan_begin on its first group.
struct ScanState {
    int cur, end, c, c2, c3, cell;       // c, c2, c3: edges of the three smallest margins (-1: none)
    float m1, m2, m3, m4, xb, pe, eps;  // xb: lower bound of every unvisited edge's margin
    float alpha;       // the cell's annulus
    float nx, ny, nz;  // -s
};

// Bounds of the scan of S.cell at s = -(nx, ny, nz), given the cell's edge
// list [off, end), largest half length hmax and annulus.  eps bounds the FP32
// rounding of every margin of the cell (three fused roundings of terms at
// most h + |s|_1, plus the FP32 unit vectors and half lengths: below
// 4 * 2^-24 * (|s|_1 + h_max)); pe is an upper bound of |s| (the hardware
// square root, 1 ulp, of the rounded sum of squares, raised by 2^-20
// relative) plus eps, so h - pe is a lower bound of the computed margin of
// any edge of half length h.
__device__ __forceinline__ void scan_setup(ScanState& S, int off, int end, float hmax, float alpha)
{
    S.cur = off;
    S.end = end;
    S.alpha = alpha;
    S.c = S.c2 = S.c3 = -1;
    S.m1 = S.m2 = S.m3 = S.m4 = S.xb = __builtin_inff();
    S.eps = kUlp * (fabsf(S.nx) + fabsf(S.ny) + fabsf(S.nz) + hmax);
    const float n2 = fmaf(S.nx, S.nx, fmaf(S.ny, S.ny, S.nz * S.nz));
    S.pe = fmaf(__builtin_amdgcn_sqrtf(n2), 1.f + 0x1.0p-20f, S.eps);
}

// The same with the cell's record read from the geometry (one load).
__device__ __forceinline__ void scan_begin(const GeomDev& g, ScanState& S)
{
    const int4 r = gload(g.cell_rec + S.cell);
    scan_setup(S, r.x, r.y, __int_as_float(r.z), __int_as_float(r.w));
}

__device__ __forceinline__ void scan_start(const GeomDev& g, int cell, float sx, float sy, float sz, ScanState& S)
{
    S.cell = cell;
    S.nx = -sx;
    S.ny = -sy;
    S.nz = -sz;
    scan_begin(g, S);
}

// One group of kEdgePad edges (requires S.cur < S.end; lists are padded with
// sentinels whose margin is +inf).  All records are fetched before any is
// used, so a group costs one memory round trip.  The scan ends when the
// group's last half length minus pe reaches min(m3, cap) (an anchor scan,
// third) or min(m2, cap) (k's side): edges are sorted by h, so every later
// computed margin is at least that bound (xb), the scan's m1, c, m2, c2
// (and for an anchor scan m3, c3) are those of the whole list up to the
// cap, and min(m4, xb) bounds every margin but the three smallest from
// below.  Evaluating the rest of the last group cannot change those values:
// the group needs no per-edge branch.
// Evaluates the group r (the records at S.cur) into S; see scan_chunk.
__device__ __forceinline__ void scan_group(const float4 (&r)[kEdgePad], float cap, ScanState& S, uint32_t& facets,
                                           bool third)
{
    // chunk-local codes: cl -1 = the previous c; cl2 -2 = the previous c2,
    // -1 = the previous c; cl3 -3 = the previous c3, -2 / -1 as for cl2
    int cl = -1, cl2 = -2, cl3 = -3;
#pragma unroll
    for (int u = 0; u < kEdgePad; ++u) {
        const float m = fmaf(S.nx, r[u].x, fmaf(S.ny, r[u].y, fmaf(S.nz, r[u].z, r[u].w)));
        const bool lt1 = m < S.m1, lt2 = m < S.m2, lt3 = m < S.m3;
        cl3 = lt2 ? cl2 : (lt3 ? u : cl3);
        cl2 = lt1 ? cl : (lt2 ? u : cl2);
        cl = lt1 ? u : cl;
        // m1 <= m2 <= m3 <= m4 throughout, so each update is a median
        S.m4 = __builtin_amdgcn_fmed3f(S.m3, m, S.m4);
        S.m3 = __builtin_amdgcn_fmed3f(S.m2, m, S.m3);
        S.m2 = __builtin_amdgcn_fmed3f(S.m1, m, S.m2);
        S.m1 = fminf(S.m1, m);
    }
    const int c0 = S.c, c20 = S.c2;
    S.c3 = cl3 == -3 ? S.c3 : (cl3 == -2 ? c20 : (cl3 == -1 ? c0 : S.cur + cl3));
    S.c2 = cl2 == -2 ? c20 : (cl2 == -1 ? c0 : S.cur + cl2);
    if (cl >= 0) S.c = S.cur + cl;
    facets += kEdgePad;
    S.cur += kEdgePad;
    // an anchor scan (third) runs until the third margin is exact, k's side
    // until the second
    const float bound = r[kEdgePad - 1].w - S.pe;
    if (bound >= fminf(third ? S.m3 : S.m2, cap)) {
        S.xb = bound;
        S.cur = S.end;
    }
}

__device__ __forceinline__ void scan_chunk(const GeomDev& g, float cap, ScanState& S, uint32_t& facets, bool third)
{
    const float4* __restrict__ rec = g.nb_rec + S.cur;
    float4 r[kEdgePad];
#pragma unroll
    for (int u = 0; u < kEdgePad; ++u) r[u] = gload(rec + u);
    scan_group(r, cap, S, facets, third);
}

// Two groups in one memory round trip: the second group's records are
// fetched with the first when the list has them, and evaluated only if the
// scan goes on after the first, so the scan state is that of two
// scan_chunk calls.
__device__ __forceinline__ void scan_chunk2(const GeomDev& g, float cap, ScanState& S, uint32_t& facets, bool third)
{
    const float4* __restrict__ rec = g.nb_rec + S.cur;
    const bool has2 = S.cur + kEdgePad < S.end;
    float4 r[kEdgePad], r2[kEdgePad];
#pragma unroll
    for (int u = 0; u < kEdgePad; ++u) r[u] = gload(rec + u);
    if (has2) {
#pragma unroll
        for (int u = 0; u < kEdgePad; ++u) r2[u] = gload(rec + kEdgePad + u);
    }
    scan_group(r, cap, S, facets, third);
    if (has2 && S.cur < S.end) scan_group(r2, cap, S, facets, third);
}

struct Scan {
    float m1, m2, m3, r, eps;  // r: a lower bound of every margin but the three smallest
    int c, c2, c3;             // edges of m1, m2 and m3 (-1: none)
};

// Whole scan of one cell.
__device__ __forceinline__ Scan scan_cell(const GeomDev& g, int cell, float px, float py, float pz, float cap,
                                          uint32_t& facets, bool third)
{
    ScanState S;
    scan_start(g, cell, px, py, pz, S);
    while (S.cur < S.end) scan_chunk(g, cap, S, facets, third);
    return Scan{S.m1, S.m2, S.m3, fminf(S.m4, S.xb), S.eps, S.c, S.c2, S.c3};
}

// Tier from the scans at q: edges c, c2, c3 of anchor a with the three
// smallest margins (m1 the smallest, m2a the second), a lower bound ra of
// every other margin of a; k-side bound n2 when the k side was scanned (an
// outside point).  c < 0 means no edge below the cap (every margin is at
// least cap: pass m2a = ra = cap; the tier then answers "inside a" up to
// cap - alpha_a); m2a = ra = -inf gives a tier that answers nothing.  A
// missing second or third edge leaves that plane at (0, max(ra, 0)) with
// ra >= cap then (the scan ran past the cap).  One code path for every case,
// so the walk's update is branch-free.
__device__ __forceinline__ void tier_make(const GeomDev& g, int a, int c, int c2, int c3, float qx, float qy, float qz,
                                          float m1, float m2a, float ra, float epsa, bool kside, float n2, float epsk,
                                          Tier& t)
{
    const bool plane = c >= 0, plane2 = plane && c2 >= 0, plane3 = plane2 && c3 >= 0;
    const int ci = plane ? c : 0, ci2 = plane2 ? c2 : 0, ci3 = plane3 ? c3 : 0;
    const float4 e0 = gload(g.nb_rec + ci), e20 = gload(g.nb_rec + ci2), e30 = gload(g.nb_rec + ci3);
    const int k = plane ? gload(g.nb_id + ci) : a;
    const float4 e = plane ? e0 : make_float4(0.f, 0.f, 0.f, m2a);
    const float hr = fmaxf(ra, 0.f);
    const float4 e2 = plane2 ? e20 : make_float4(0.f, 0.f, 0.f, hr);
    const float4 e3 = plane3 ? e30 : make_float4(0.f, 0.f, 0.f, hr);
    const int4 ar = gload(g.cell_rec + a);
    // k's annulus from the edge's own copy (nb_kalpha[c] = cell_alpha[nb_id[c]]),
    // so every load here depends on the scan's edges only: one round trip
    const float ka = gload(g.nb_kalpha + ci);
    const float aa = __int_as_float(ar.w), ak = plane ? ka : aa;
    t.a_off = ar.x;
    t.a_end = ar.y;
    t.a_hmax = __int_as_float(ar.z);
    kside = kside && plane;
    const float ep = kside ? fmaxf(epsa, epsk) : epsa;
    const float ra_in = ra - fmaxf(aa, 0.f) - 2.f * ep, ra_out = ra - 2.f * ep;
    const float rk_in = kside ? n2 - fmaxf(ak, 0.f) - 2.f * ep : 0.f;
    const float rk_out = kside ? n2 - 2.f * ep : 0.f;
    t.ux = e.x;
    t.uy = e.y;
    t.uz = e.z;
    t.h = e.w;
    t.ux2 = e2.x;
    t.uy2 = e2.y;
    t.uz2 = e2.z;
    t.h2 = e2.w;
    t.ux3 = e3.x;
    t.uy3 = e3.y;
    t.uz3 = e3.z;
    t.h3 = e3.w;
    t.cx = qx;
    t.cy = qy;
    t.cz = qz;
    t.k = k;
    t.alpha_a = aa;
    t.alpha_k = ak;
    t.ra_in2 = ra_in > 0.f ? ra_in * ra_in : 0.f;
    t.ra_out2 = ra_out > 0.f ? ra_out * ra_out : 0.f;
    t.rk_in2 = rk_in > 0.f ? rk_in * rk_in : 0.f;
    t.rk_out2 = rk_out > 0.f ? rk_out * rk_out : 0.f;
    // outside point (kside): distance to the nearest contracted cell is at
    // least min(alpha_a - m1, alpha_k + m1, alpha_min + max(min(m2, n2), 0))
    const float rout =
        kside ? fminf(fminf(aa - m1, ak + m1), g.alpha_min + fmaxf(fminf(m2a, n2), 0.f)) - 2.f * ep : 0.f;
    t.rout2 = rout > 0.f ? rout * rout : 0.f;
    t.eps = ep + kUlp * (1.8f * fmaxf(fmaxf(ra_out, rk_out), 0.f) +
                         fmaxf(fmaxf(fabsf(e.w), fabsf(e2.w)), fabsf(e3.w)));
}

// Full classification of q from anchor a: greedy location, label, FP64
// re-decision near the alpha threshold, and a new tier.  q and o move with
// the anchor.  The k side is scanned only for outside points (an inside
// walker reaches k's side only by crossing the whole gap).  Returns 0 on a
// classifier failure.
__device__ int classify_full(const GeomDev& g, int& a, float& qx, float& qy, float& qz, float& ox, float& oy,
                             float& oz, float cap, int& inside, Tier& t, uint32_t& iters, uint32_t& facets,
                             uint32_t& rechecks)
{
    for (int it = 0; it <= kMaxGraphIters; ++it) {
        const Scan s = scan_cell(g, a, qx, qy, qz, cap, facets, true);
        if (s.m1 < -s.eps) {
            // a neighbour is nearer: move the anchor and rescan
            const double ex = g.nb_e64[3 * s.c], ey = g.nb_e64[3 * s.c + 1], ez = g.nb_e64[3 * s.c + 2];
            qx = (float)((double)qx - ex);
            qy = (float)((double)qy - ey);
            qz = (float)((double)qz - ez);
            ox = (float)((double)ox + ex);
            oy = (float)((double)oy + ey);
            oz = (float)((double)oz + ez);
            a = g.nb_id[s.c];
            ++iters;
            continue;
        }
        const float a1 = g.cell_alpha[a];
        inside = s.m1 >= a1;
        if (fabsf(s.m1 - a1) <= s.eps) {
            ++rechecks;
            if (!locate_f64(g, a, qx, qy, qz, ox, oy, oz, inside, facets)) return 0;
            tier_clear(t);
            return 1;
        }
        if (s.c < 0 || s.m1 >= cap) {
            tier_make(g, a, -1, -1, -1, qx, qy, qz, 0.f, cap, cap, s.eps, false, 0.f, 0.f, t);
            return 1;
        }
        const bool kside = !inside;
        float n2 = 0.f, epsk = 0.f;
        if (kside) {
            // k's margins at q; the smallest is at most the shared plane's
            // (-m1 <= eps), so the second smallest bounds every other k plane
            // whenever it is positive
            const float kx = (float)((double)qx - g.nb_e64[3 * s.c]);
            const float ky = (float)((double)qy - g.nb_e64[3 * s.c + 1]);
            const float kz = (float)((double)qz - g.nb_e64[3 * s.c + 2]);
            const Scan sk = scan_cell(g, g.nb_id[s.c], kx, ky, kz, cap, facets, false);
            n2 = fminf(sk.m2, cap);
            epsk = sk.eps;
        }
        tier_make(g, a, s.c, s.c2, s.c3, qx, qy, qz, s.m1, fminf(s.m2, cap), fminf(s.r, cap), s.eps, kside, n2, epsk,
                  t);
        return 1;
    }
    return 0;
}

struct TileOut {
    double* slot;          // n_cols x 2 doubles of this tile
    long long* occupancy;  // n_grid
    long long* flags;      // MADI_N_FLAGS
};


// LDS a tile walk needs.
template <int NP>
struct TileLds {
    double cols[kWaves * kColChunk * 2];
    int occ[kMaxLdsGrid];
    float evc[kMaxLdsGrid][NP];  // per grid time: summed event signs per (delta, Delta) pair
};

// A proposal under classification: point relative to anchor na, the
// anchor's offset, and the a-side results kept for the tier.
struct Pending {
    float qx, qy, qz, ox, oy, oz;
    int na, pin, moves, sa_c, sa_c2, sa_c3;
    float sa_m1, sa_m2, sa_r, sa_eps;  // sa_r: the tier radius bound (Scan::r, capped)
};

constexpr int kScanA = 1, kScanK = 2;  // scan phases of a pending classification

__device__ __forceinline__ void pending_start(Pending& P, ScanState& S, int& phase)
{
    P.moves = 0;
    P.sa_c = P.sa_c2 = P.sa_c3 = -1;
    P.sa_m1 = P.sa_m2 = P.sa_r = P.sa_eps = 0.f;
    S.cell = P.na;
    S.nx = -P.qx;
    S.ny = -P.qy;
    S.nz = -P.qz;
    S.cur = -1;
    phase = kScanA;
}

// Advances a pending classification by one group of edges (TWO: by up to
// two, fetched in one round trip) and, when a scan ends, by its end-of-scan
// step: greedy move to a nearer neighbour, the
// label with FP64 re-decision near alpha, the start of k's side.  Returns 1
// when the label is known (tmode: 1 two-sided tier, 2 none, 3 no plane
// below the cap, 4 a side only), -1 on a classifier failure, 0 otherwise.
template <bool TWO = false>
__device__ __forceinline__ int classify_round(const GeomDev& g, float cap, int& phase, ScanState& S, Pending& P,
                                              int& tmode, uint32_t& nfac, uint32_t& iters, uint32_t& facets,
                                              uint32_t& rechecks)
{
    if (S.cur < 0) scan_begin(g, S);
    if (TWO) scan_chunk2(g, cap, S, nfac, phase == kScanA);
    else scan_chunk(g, cap, S, nfac, phase == kScanA);
    if (S.cur < S.end) return 0;
    int next;  // 1: greedy move to the nearer neighbour, 2: scan k's side
    if (phase == kScanK) {
        // k's smallest margin is at most the shared plane's (-m1 <= eps),
        // so its second smallest bounds every other k plane when positive
        tmode = 1;
        return 1;
    }
    if (S.m1 < -S.eps) {
        if (++P.moves > kMaxGraphIters) return -1;  // the greedy walk does not settle
        next = 1;
    } else {
        const float a1 = S.alpha;  // the annulus of P.na, the cell just scanned
        P.pin = S.m1 >= a1;
        P.sa_c = S.c;
        P.sa_c2 = S.c2;
        P.sa_c3 = S.c3;
        P.sa_m1 = S.m1;
        P.sa_m2 = fminf(S.m2, cap);
        P.sa_r = fminf(fminf(S.m4, S.xb), cap);
        P.sa_eps = S.eps;
        if (fabsf(S.m1 - a1) <= S.eps) {
            ++rechecks;
            if (!locate_f64(g, P.na, P.qx, P.qy, P.qz, P.ox, P.oy, P.oz, P.pin, facets)) return -1;
            tmode = 2;
            return 1;
        }
        if (S.c < 0 || S.m1 >= cap) {
            // every margin is at least cap (which edge attains the smallest
            // is then immaterial)
            tmode = 3;
            return 1;
        }
        if (P.pin) {
            // an inside walker reaches k's side only across the whole gap
            tmode = 4;
            return 1;
        }
        next = 2;
    }
    // the point relative to neighbour k of the binding edge, and k's record:
    // one round of loads, after which k's scan starts with its first group
    const int c = S.c;
    const double ex = gload(g.nb_e64 + 3 * c), ey = gload(g.nb_e64 + 3 * c + 1), ez = gload(g.nb_e64 + 3 * c + 2);
    const int4 kr = gload(g.nb_krec + c);
    const float ka = gload(g.nb_kalpha + c);
    const int k = kr.x;
    const float kx = (float)((double)P.qx - ex);
    const float ky = (float)((double)P.qy - ey);
    const float kz = (float)((double)P.qz - ez);
    if (next == 1) {
        P.qx = kx;
        P.qy = ky;
        P.qz = kz;
        P.ox = (float)((double)P.ox + ex);
        P.oy = (float)((double)P.oy + ey);
        P.oz = (float)((double)P.oz + ez);
        P.na = k;
        ++iters;
    }
    S.cell = k;
    S.nx = -kx;
    S.ny = -ky;
    S.nz = -kz;
    scan_setup(S, kr.y, kr.z, __int_as_float(kr.w), ka);
    phase = next == 1 ? kScanA : kScanK;
    return 0;
}

// The tier a completed classification leaves (see tier_make).
__device__ __forceinline__ void tier_of(const GeomDev& g, const Pending& P, int tmode, const ScanState& S, float cap,
                                        Tier& t)
{
    const float none = -__builtin_inff();
    const float m2a = tmode == 2 ? none : (tmode == 3 ? cap : P.sa_m2);
    const float ra = tmode == 2 ? none : (tmode == 3 ? cap : P.sa_r);
    // tmode 2 (FP64 re-decision) may have moved the anchor off the scanned
    // cell, whose edges are then not a's: no planes (the tier answers nothing)
    const bool planes = tmode != 2 && tmode != 3;
    tier_make(g, P.na, planes ? P.sa_c : -1, planes ? P.sa_c2 : -1, planes ? P.sa_c3 : -1, P.qx, P.qy, P.qz, P.sa_m1,
              m2a, ra, P.sa_eps, tmode == 1, fminf(S.m2, cap), S.eps, t);
}

// One tile of kBlock walkers [w_lo, w_lo + kBlock) of an entry: the whole
// walk, then the phase columns into the tile's slot.  All threads of the
// block must call it (it synchronises on LDS).  NP >= plan.n_pairs bounds
// the moment registers.
//
// Each lane runs its walker as a small state machine, so lanes of a wave
// need not be at the same step.  Per iteration a lane either takes its next
// step (draws, proposal, plane tier) or, while its proposal is being
// classified, advances the classification by up to R rounds of (one group of
// kEdgePad edges, then the end-of-scan step: greedy move, label, FP64
// re-decision, start of k's side).  A step completes when its label is
// known.  A wave's iteration thus costs one step plus R groups instead of the
// longest classification among its 64 lanes, and walkers with long
// classifications fall behind by a few iterations rather than stalling the
// wave.  Every walker consumes its own random stream in order, so results do
// not depend on R.
template <int NP, int R>
__device__ __forceinline__ void walk_tile_f32(const GeomDev& g, const WalkDev& p, const Gf2Col* __restrict__ jump_pow,
                                              int64_t w_lo, const PlanDev& plan, const Events& ev, const TileOut& out,
                                              float cap_width, unsigned long long* counters, TileLds<NP>& L)
{
    enum : int { kStep = 0, kDone = 3 };
    const int lane = threadIdx.x & 63;
    const int64_t w = w_lo + threadIdx.x;
    const bool live = w < p.n_walkers;
    uint32_t n_class = 0, n_tier = 0, crossings = 0, rejects = 0, nfac = 0;
    uint32_t iters = 0, facets = 0, rechecks = 0;
    const float cap = g.alpha_max + cap_width;
    // grids up to kMaxLdsGrid keep occupancy counters and the event table in
    // LDS; longer grids count in global memory and walk the event list
    const bool occ_in_lds = p.n_grid <= kMaxLdsGrid;
    if (occ_in_lds) {
        for (int j = threadIdx.x; j < p.n_grid; j += kBlock) {
            L.occ[j] = 0;
#pragma unroll
            for (int u = 0; u < NP; ++u) L.evc[j][u] = 0.f;
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int e = 0; e < ev.n; ++e) L.evc[ev.j[e]][ev.pair[e]] += (float)ev.sign[e];
    }
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
        if (!classify_full(g, a, px, py, pz, ox, oy, oz, cap, in, t, iters, facets, rechecks)) {
            err = 1;
            escaped = 1;
        }
        cell = a;
        ox = -px;
        oy = -py;
        oz = -pz;
    } else if (live) {
        tier_free_water(t);
    }
    const float lox = (float)(-x0), loy = (float)(-y0), loz = (float)(-z0);
    const float hix = (float)(g.L - x0), hiy = (float)(g.L - y0), hiz = (float)(g.L - z0);

    float Sx = 0.f, Sy = 0.f, Sz = 0.f, Cx = 0.f, Cy = 0.f, Cz = 0.f, hx = 0.f, hy = 0.f, hz = 0.f;
    float dM[NP][3];
#pragma unroll
    for (int q = 0; q < NP; ++q) dM[q][0] = dM[q][1] = dM[q][2] = 0.f;
    if (live && !err && in) {
        if (occ_in_lds) atomicAdd(&L.occ[0], 1);
        else atomicAdd((unsigned long long*)&out.occupancy[0], 1ull);
    }

    const float nscale = (float)(-2.0 * p.sigma * p.sigma * 0.6931471805599453);
    const float ts = (float)p.ts;
    int phase = (live && !escaped && p.n_steps > 0) ? kStep : kDone;
    int step = 0, until_sample = p.steps_per_h, j_sample = 0, next_ev = 0;
    // the pending step: proposal, label once known, acceptance bits
    Pending P;
    P.qx = P.qy = P.qz = P.ox = P.oy = P.oz = 0.f;
    P.na = P.pin = P.moves = 0;
    P.sa_c = P.sa_c2 = P.sa_c3 = -1;
    P.sa_m1 = P.sa_m2 = P.sa_r = P.sa_eps = 0.f;
    int pcell = 0, rj = 0;
    ScanState S;
    S.cur = S.end = 0;
    S.c = S.c2 = S.c3 = -1;
    S.cell = 0;
    S.m1 = S.m2 = S.m3 = S.m4 = S.xb = S.pe = S.eps = S.nx = S.ny = S.nz = 0.f;
    uint32_t wave_iters = 0, lane_iters = 0;
    while (__any(phase != kDone)) {
        ++wave_iters;
        lane_iters += phase != kDone;
        int complete = 0, tmode = 0;  // tmode 0: keep the tier; 1-4 see classify_round
        if (phase == kStep) {
            const float dx = normal_scaled_f32(rng, nscale);
            const float dy = normal_scaled_f32(rng, nscale);
            const float dz = normal_scaled_f32(rng, nscale);
            const uint64_t ku = rng_next(rng) >> 11;
            rj = (ku >= p.rej1 ? 1 : 0) | (ku >= p.rej2 ? 2 : 0);
            P.qx = px + dx;
            P.qy = py + dy;
            P.qz = pz + dz;
            P.ox = ox;
            P.oy = oy;
            P.oz = oz;
            P.na = a;
            const float nDx = ox + P.qx, nDy = oy + P.qy, nDz = oz + P.qz;
            if (!g.periodic && (nDx < lox || nDx >= hix || nDy < loy || nDy >= hiy || nDz < loz || nDz >= hiz)) {
                escaped = 1;
                phase = kDone;
                // samples after an escape keep the last label
                if (in)
                    for (int j = j_sample + 1; j < p.n_grid; ++j) {
                        if (occ_in_lds) atomicAdd(&L.occ[j], 1);
                        else atomicAdd((unsigned long long*)&out.occupancy[j], 1ull);
                    }
            } else if (tier_query(t, a, P.qx, P.qy, P.qz, pcell, P.pin)) {
                ++n_tier;
                complete = 1;
            } else {
                ++n_class;
                pending_start(P, S, phase);
                if (t.a_off >= 0) scan_setup(S, t.a_off, t.a_end, t.a_hmax, t.alpha_a);  // the anchor is the tier's
            }
        }
#pragma unroll 1
        for (int r = 0; r < R; ++r) {
            if (phase == kScanA || phase == kScanK) {
                const int res = classify_round(g, cap, phase, S, P, tmode, nfac, iters, facets, rechecks);
                if (res > 0) {
                    complete = 1;
                    pcell = P.na;
                    phase = kStep;  // completed below
                } else if (res < 0) {
                    err = 1;
                    escaped = 1;
                    phase = kDone;
                }
            }
        }
        if (complete) {
            int m = 0;
            if (in != P.pin) m = 1;
            else if (in && P.pin && cell != pcell) m = 2;
            bool accept = true;
            if (m > 0) {
                ++crossings;
                if ((rj >> (m - 1)) & 1) {
                    accept = false;
                    ++rejects;
                }
            }
            if (accept) {
                if (tmode != 0) tier_of(g, P, tmode, S, cap, t);
                a = P.na;
                in = P.pin;
                cell = pcell;
                px = P.qx;
                py = P.qy;
                pz = P.qz;
                ox = P.ox;
                oy = P.oy;
                oz = P.oz;
            }
            const float Dx = ox + px, Dy = oy + py, Dz = oz + pz;
            hx += Dx;
            hy += Dy;
            hz += Dz;
            ++step;
            if (--until_sample == 0) {
                until_sample = p.steps_per_h;
                const int j = ++j_sample;
                // fold the interval's plain sum into the compensated total
                const float yx = hx - Cx, tx = Sx + yx;
                Cx = (tx - Sx) - yx;
                Sx = tx;
                const float yy = hy - Cy, ty = Sy + yy;
                Cy = (ty - Sy) - yy;
                Sy = ty;
                const float yz = hz - Cz, tz = Sz + yz;
                Cz = (tz - Sz) - yz;
                Sz = tz;
                hx = hy = hz = 0.f;
                if (in) {
                    if (occ_in_lds) atomicAdd(&L.occ[j], 1);
                    else atomicAdd((unsigned long long*)&out.occupancy[j], 1ull);
                }
                if (occ_in_lds) {
                    // Y(t_j) = ts * (sum_{k<=n} D_k - D_n / 2), relative to the start point
                    const float Yx = ts * ((Sx - Cx) - 0.5f * Dx);
                    const float Yy = ts * ((Sy - Cy) - 0.5f * Dy);
                    const float Yz = ts * ((Sz - Cz) - 0.5f * Dz);
#pragma unroll
                    for (int u = 0; u < NP; ++u) {
                        const float sg = L.evc[j][u];
                        dM[u][0] = fmaf(sg, Yx, dM[u][0]);
                        dM[u][1] = fmaf(sg, Yy, dM[u][1]);
                        dM[u][2] = fmaf(sg, Yz, dM[u][2]);
                    }
                } else if (next_ev < ev.n && ev.j[next_ev] == j) {
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
    facets += nfac;
    {
        const long long ne = wave_popc(live && escaped);
        const long long nr = wave_popc(live && err);
        if (lane == 0) {
            if (ne) atomicAdd((unsigned long long*)&out.flags[MADI_FLAG_ESCAPED], (unsigned long long)ne);
            if (nr) atomicAdd((unsigned long long*)&out.flags[MADI_FLAG_CLASSIFIER_ERR], (unsigned long long)nr);
        }
    }
    __syncthreads();
    if (occ_in_lds)
        for (int j = threadIdx.x; j < p.n_grid; j += kBlock)
            if (L.occ[j]) atomicAdd((unsigned long long*)&out.occupancy[j], (unsigned long long)L.occ[j]);
    emit_columns<float, NP>(plan, dM, live && !escaped, L.cols, out.slot);
    if (counters) {
        const double v0 = wave_sum_f64((double)n_class), v1 = wave_sum_f64((double)n_tier),
                     v2 = wave_sum_f64((double)iters), v3 = wave_sum_f64((double)rechecks),
                     v4 = wave_sum_f64((double)facets), v5 = wave_sum_f64((double)crossings),
                     v6 = wave_sum_f64((double)rejects), v7 = wave_sum_f64((double)lane_iters);
        if (lane == 0) {
            atomicAdd(&counters[MADI_CNT_CLASSIFY], (unsigned long long)v0);
            atomicAdd(&counters[MADI_CNT_SAFE_HIT], (unsigned long long)v1);
            atomicAdd(&counters[MADI_CNT_GRAPH_ITERS], (unsigned long long)v2);
            atomicAdd(&counters[MADI_CNT_FP64_RECHECK], (unsigned long long)v3);
            atomicAdd(&counters[MADI_CNT_FACETS], (unsigned long long)v4);
            atomicAdd(&counters[MADI_CNT_CROSSINGS], (unsigned long long)v5);
            atomicAdd(&counters[MADI_CNT_REJECTS], (unsigned long long)v6);
            atomicAdd(&counters[MADI_CNT_WAVE_ITERS], (unsigned long long)wave_iters);
            atomicAdd(&counters[MADI_CNT_LANE_ITERS], (unsigned long long)v7);
        }
    }
}

// Orders LDS accesses of different lanes of one wave (a wave's LDS
// instructions complete in order; this keeps the compiler from moving them).
__device__ __forceinline__ void wave_lds_fence()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// ------------------------------------------------------------ streamed walk
//
// Lane refill: the walkers of a batch are handed to lanes one at a time, so a
// lane whose walker ends takes the next one instead of idling until the
// slowest walker of its wave ends.  Each wave claims kClaimWalkers walkers of
// a work item at a time from the item's counter, computes their generator
// states into its LDS slots, and hands them out as lanes free up.  A walker's
// first classification runs through the same per-lane state machine as its
// steps, so starting a walker does not stall the wave.  A finished walker
// stores its moments and active flag; k_emit_tiles then forms every tile's
// column sums from them exactly as the tile variants do, so results are
// bitwise those of the other variants.  A wave starts at the item holding its
// share of the batch's walkers and moves to the next item with walkers left
// once its lanes are idle; it exits when no item has walkers left (every
// wave reaches that point: claims only advance the counters).

constexpr int kClaimWalkers = 64;  // one walker per lane: at the end of a batch no wave holds more than
                                   // one walker per lane that others could have taken

template <int NP>
struct StreamLds {
    float evc[kMaxLdsGrid][NP];     // per grid time: summed event signs per pair (the batch shares the plan)
    int occ[kWaves][kMaxLdsGrid];   // per wave: inside counts of its current item
    Rng states[kWaves][kClaimWalkers];
};

// Number of set bits of mask below this lane.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Claims the next kClaimWalkers walkers of item W for the calling wave (all
// 64 lanes active) and writes their generator states to st; returns the
// claim's first walker relative to the item (n_walk or more: none left).
__device__ __forceinline__ int64_t stream_claim(const WorkDev* __restrict__ W, unsigned long long* next,
                                                const Gf2Col* __restrict__ jump_pow, Rng* st, int lane)
{
    unsigned long long b = 0;
    if (lane == 0) b = atomicAdd(next, (unsigned long long)kClaimWalkers);
    b = __shfl(b, 0, 64);
    const int64_t base = (int64_t)__builtin_amdgcn_readfirstlane((uint32_t)b) |
                         ((int64_t)__builtin_amdgcn_readfirstlane((uint32_t)(b >> 32)) << 32);
    if (base >= W->n_walk) return base;
    st[lane] = rng_for_walker(W->w.seed, W->w.first_walker + (uint64_t)(W->walker_lo + base + lane), jump_pow);
    wave_lds_fence();
    return base;
}

template <int NP, int R, bool EXIT>
__device__ __forceinline__ void stream_wave(const WorkDev* __restrict__ work, unsigned long long* item_next, const BatchHdr* hdr,
                            const PlanDev& plan,
                            const Gf2Col* __restrict__ jump_pow, const Events& ev, float cap_width,
                            unsigned long long* counters, StreamLds<NP>& L)
{
    enum : int { kStep = 0, kDone = 3, kIdle = 4 };
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int32_t n_work = hdr->n_work;
    const int64_t total_w = hdr->total_walkers;
    int* occ = L.occ[wv];
    Rng* st = L.states[wv];
    for (int j = lane; j < kMaxLdsGrid; j += 64) occ[j] = 0;
    wave_lds_fence();
    uint32_t n_class = 0, n_tier = 0, crossings = 0, rejects = 0, nfac = 0, n_exit = 0;
    uint32_t iters = 0, facets = 0, rechecks = 0;
    uint32_t wave_iters = 0, lane_iters = 0;
    const int n_pairs = plan.n_pairs;
    if (n_work > 0 && total_w > 0) {
        // home item: the one holding this wave's share of the walkers
        const int64_t n_waves = (int64_t)gridDim.x * kWaves;
        const int64_t pos = ((int64_t)blockIdx.x * kWaves + wv) * total_w / n_waves;
        int lo = 0, hi = n_work - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (work[mid].walker_begin <= pos) lo = mid;
            else hi = mid - 1;
        }
        int e = __builtin_amdgcn_readfirstlane(lo);
        int dry = 0;  // items found without walkers in a row
        while (dry < n_work) {
            const WorkDev* __restrict__ W = work + e;
            unsigned long long* next = item_next + e;
            int64_t claim = stream_claim(W, next, jump_pow, st, lane);
            if (claim >= W->n_walk) {
                ++dry;
                e = e + 1 == n_work ? 0 : e + 1;
                continue;
            }
            dry = 0;
            const GeomDev& g = W->g;
            const WalkDev& p = W->w;
            const int64_t n_walk = W->n_walk, walker_lo = W->walker_lo;
            const float cap = g.alpha_max + cap_width;
            const bool occ_in_lds = p.n_grid <= kMaxLdsGrid;
            const float nscale = (float)(-2.0 * p.sigma * p.sigma * 0.6931471805599453);
            const float leave = EXIT ? exit_bound(g, p) : 0.f;  // EXIT: the certificate is compiled in
            const float ts = (float)p.ts;
            int64_t hand = claim, hand_end = claim + kClaimWalkers < n_walk ? claim + kClaimWalkers : n_walk;
            bool item_dry = false;

            // per-lane walker state
            int phase = kIdle, init = 0;
            int64_t w = 0;
            Rng rng{0, 0};
            int a = 0, in = 0, cell = 0, err = 0, escaped = 0;
            uint32_t n_esc = 0, n_err = 0;
            float px = 0.f, py = 0.f, pz = 0.f, ox = 0.f, oy = 0.f, oz = 0.f;
            float lox = 0.f, loy = 0.f, loz = 0.f, hix = 0.f, hiy = 0.f, hiz = 0.f;
            Tier t;
            tier_clear(t);
            float Sx = 0.f, Sy = 0.f, Sz = 0.f, Cx = 0.f, Cy = 0.f, Cz = 0.f, hx = 0.f, hy = 0.f, hz = 0.f;
            float dM[NP][3];
#pragma unroll
            for (int q = 0; q < NP; ++q) dM[q][0] = dM[q][1] = dM[q][2] = 0.f;
            int step = 0, until_sample = 0, j_sample = 0, next_ev = 0;
            Pending P;
            P.qx = P.qy = P.qz = P.ox = P.oy = P.oz = 0.f;
            P.na = P.pin = P.moves = 0;
            P.sa_c = P.sa_c2 = P.sa_c3 = -1;
            P.sa_m1 = P.sa_m2 = P.sa_r = P.sa_eps = 0.f;
            int pcell = 0, rj = 0;
            ScanState S;
            S.cur = S.end = 0;
            S.c = S.c2 = S.c3 = -1;
            S.cell = 0;
            S.m1 = S.m2 = S.m3 = S.m4 = S.xb = S.pe = S.eps = S.nx = S.ny = S.nz = 0.f;
            for (;;) {
                // hand walkers to idle lanes
                uint64_t need = __ballot(phase == kIdle);
                if (need != 0 && !item_dry) {
                    bool got = false;
                    while (need != 0) {
                        if (hand >= hand_end) {
                            const int64_t nb = stream_claim(W, next, jump_pow, st, lane);
                            if (nb >= n_walk) {
                                item_dry = true;
                                break;
                            }
                            claim = nb;
                            hand = nb;
                            hand_end = nb + kClaimWalkers < n_walk ? nb + kClaimWalkers : n_walk;
                        }
                        const int64_t rank = (int64_t)lanes_below(need);
                        const bool mine = ((need >> lane) & 1ull) && rank < hand_end - hand;
                        if (mine) {
                            w = hand + rank;
                            rng = st[w - claim];
                            got = true;
                        }
                        const uint64_t taken = __ballot(mine);
                        hand += __popcll(taken);
                        need &= ~taken;
                        wave_lds_fence();  // the slots are read before a claim rewrites them
                    }
                    if (got) {
                        // walker start (walk_tile_f32's set-up, the first classification queued)
                        const double span = g.source_hi - g.source_lo;
                        const double x0 = g.source_lo + uniform_f64(rng) * span;
                        const double y0 = g.source_lo + uniform_f64(rng) * span;
                        const double z0 = g.source_lo + uniform_f64(rng) * span;
                        lox = (float)(-x0);
                        loy = (float)(-y0);
                        loz = (float)(-z0);
                        hix = (float)(g.L - x0);
                        hiy = (float)(g.L - y0);
                        hiz = (float)(g.L - z0);
                        Sx = Sy = Sz = Cx = Cy = Cz = hx = hy = hz = 0.f;
#pragma unroll
                        for (int q = 0; q < NP; ++q) dM[q][0] = dM[q][1] = dM[q][2] = 0.f;
                        step = 0;
                        until_sample = p.steps_per_h;
                        j_sample = 0;
                        next_ev = 0;
                        err = escaped = 0;
                        if (g.free_water) {
                            tier_free_water(t);
                            a = cell = in = 0;
                            px = py = pz = ox = oy = oz = 0.f;
                            init = 0;
                            phase = p.n_steps > 0 ? kStep : kDone;
                        } else {
                            anchor_world(g, x0, y0, z0, P.na, P.qx, P.qy, P.qz);
                            P.ox = -P.qx;
                            P.oy = -P.qy;
                            P.oz = -P.qz;
                            ++n_class;
                            pending_start(P, S, phase);
                            init = 1;
                        }
                    }
                }
                if (__ballot(phase != kIdle) == 0) break;
                ++wave_iters;
                lane_iters += phase != kIdle;
                int complete = 0, tmode = 0;
                if (phase == kStep) {
                    const float dx = normal_scaled_f32(rng, nscale);
                    const float dy = normal_scaled_f32(rng, nscale);
                    const float dz = normal_scaled_f32(rng, nscale);
                    const uint64_t ku = rng_next(rng) >> 11;
                    rj = (ku >= p.rej1 ? 1 : 0) | (ku >= p.rej2 ? 2 : 0);
                    P.qx = px + dx;
                    P.qy = py + dy;
                    P.qz = pz + dz;
                    P.ox = ox;
                    P.oy = oy;
                    P.oz = oz;
                    P.na = a;
                    const float nDx = ox + P.qx, nDy = oy + P.qy, nDz = oz + P.qz;
                    if (!g.periodic && (nDx < lox || nDx >= hix || nDy < loy || nDy >= hiy || nDz < loz || nDz >= hiz)) {
                        escaped = 1;
                        phase = kDone;
                        // samples after an escape keep the last label
                        if (in)
                            for (int j = j_sample + 1; j < p.n_grid; ++j) {
                                if (occ_in_lds) atomicAdd(&occ[j], 1);
                                else atomicAdd((unsigned long long*)&W->occupancy[j], 1ull);
                            }
                    } else if (tier_query(t, a, P.qx, P.qy, P.qz, pcell, P.pin)) {
                        ++n_tier;
                        complete = 1;
                    } else if (EXIT && leave > 0.f && in && cell == a && rj == 3 &&
                               tier_exit(t, a, P.qx, P.qy, P.qz, leave)) {
                        // a certified exit completes as the rejected crossing "outside, cell a"
                        pcell = a;
                        P.pin = 0;
                        ++n_tier;
                        ++n_exit;
                        complete = 1;
                    } else {
                        ++n_class;
                        pending_start(P, S, phase);
                        if (t.a_off >= 0) scan_setup(S, t.a_off, t.a_end, t.a_hmax, t.alpha_a);  // the tier's anchor
                    }
                }
#pragma unroll 1
                for (int r = 0; r < R; ++r) {
                    if (phase == kScanA || phase == kScanK) {
                        const int res = classify_round(g, cap, phase, S, P, tmode, nfac, iters, facets, rechecks);
                        if (res > 0) {
                            complete = 1;
                            pcell = P.na;
                            phase = kStep;  // completed below
                        } else if (res < 0) {
                            err = 1;
                            escaped = 1;
                            phase = kDone;
                        }
                    }
                }
                if (complete && init) {
                    // the walker's first label: no step, no acceptance test
                    tier_of(g, P, tmode, S, cap, t);
                    a = P.na;
                    in = P.pin;
                    cell = P.na;
                    px = P.qx;
                    py = P.qy;
                    pz = P.qz;
                    ox = -P.qx;
                    oy = -P.qy;
                    oz = -P.qz;
                    if (in) {
                        if (occ_in_lds) atomicAdd(&occ[0], 1);
                        else atomicAdd((unsigned long long*)&W->occupancy[0], 1ull);
                    }
                    init = 0;
                    phase = p.n_steps > 0 ? kStep : kDone;
                } else if (complete) {
                    int m = 0;
                    if (in != P.pin) m = 1;
                    else if (in && P.pin && cell != pcell) m = 2;
                    bool accept = true;
                    if (m > 0) {
                        ++crossings;
                        if ((rj >> (m - 1)) & 1) {
                            accept = false;
                            ++rejects;
                        }
                    }
                    if (accept) {
                        if (tmode != 0) tier_of(g, P, tmode, S, cap, t);
                        a = P.na;
                        in = P.pin;
                        cell = pcell;
                        px = P.qx;
                        py = P.qy;
                        pz = P.qz;
                        ox = P.ox;
                        oy = P.oy;
                        oz = P.oz;
                    }
                    const float Dx = ox + px, Dy = oy + py, Dz = oz + pz;
                    hx += Dx;
                    hy += Dy;
                    hz += Dz;
                    ++step;
                    if (--until_sample == 0) {
                        until_sample = p.steps_per_h;
                        const int j = ++j_sample;
                        // fold the interval's plain sum into the compensated total
                        const float yx = hx - Cx, tx = Sx + yx;
                        Cx = (tx - Sx) - yx;
                        Sx = tx;
                        const float yy = hy - Cy, ty = Sy + yy;
                        Cy = (ty - Sy) - yy;
                        Sy = ty;
                        const float yz = hz - Cz, tz = Sz + yz;
                        Cz = (tz - Sz) - yz;
                        Sz = tz;
                        hx = hy = hz = 0.f;
                        if (in) {
                            if (occ_in_lds) atomicAdd(&occ[j], 1);
                            else atomicAdd((unsigned long long*)&W->occupancy[j], 1ull);
                        }
                        const float Yx = ts * ((Sx - Cx) - 0.5f * Dx);
                        const float Yy = ts * ((Sy - Cy) - 0.5f * Dy);
                        const float Yz = ts * ((Sz - Cz) - 0.5f * Dz);
                        if (W->y_out) {
                            // the walker's Y sample, widened to FP64
                            double* yo = W->y_out + ((walker_lo + w) * (int64_t)p.n_grid + j) * 3;
                            yo[0] = (double)Yx;
                            yo[1] = (double)Yy;
                            yo[2] = (double)Yz;
                        }
                        if (occ_in_lds) {
#pragma unroll
                            for (int u = 0; u < NP; ++u) {
                                const float sg = L.evc[j][u];
                                dM[u][0] = fmaf(sg, Yx, dM[u][0]);
                                dM[u][1] = fmaf(sg, Yy, dM[u][1]);
                                dM[u][2] = fmaf(sg, Yz, dM[u][2]);
                            }
                        } else {
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
                if (phase == kDone) {
                    // the walker's moments and active flag, by its index in the entry
                    const int64_t ew = walker_lo + w;
                    float* m = W->moments + ew * (int64_t)(n_pairs * 3);
#pragma unroll
                    for (int u = 0; u < NP; ++u)
                        if (u < n_pairs) {
                            m[3 * u] = dM[u][0];
                            m[3 * u + 1] = dM[u][1];
                            m[3 * u + 2] = dM[u][2];
                        }
                    W->active[ew] = escaped ? 0 : 1;
                    n_esc += escaped;
                    n_err += err;
                    phase = kIdle;
                }
            }
            // item done: its inside counts and flags
            if (occ_in_lds) {
                wave_lds_fence();
                for (int j = lane; j < p.n_grid; j += 64) {
                    const int v = occ[j];
                    if (v) atomicAdd((unsigned long long*)&W->occupancy[j], (unsigned long long)v);
                    occ[j] = 0;
                }
                wave_lds_fence();
            }
            {
                const double ne = wave_sum_f64((double)n_esc), nr = wave_sum_f64((double)n_err);
                if (lane == 0) {
                    if (ne > 0.0) atomicAdd((unsigned long long*)&W->flags[MADI_FLAG_ESCAPED], (unsigned long long)ne);
                    if (nr > 0.0)
                        atomicAdd((unsigned long long*)&W->flags[MADI_FLAG_CLASSIFIER_ERR], (unsigned long long)nr);
                }
            }
            e = e + 1 == n_work ? 0 : e + 1;
        }
    }
    facets += nfac;
    if (counters) {
        const double v0 = wave_sum_f64((double)n_class), v1 = wave_sum_f64((double)n_tier),
                     v2 = wave_sum_f64((double)iters), v3 = wave_sum_f64((double)rechecks),
                     v4 = wave_sum_f64((double)facets), v5 = wave_sum_f64((double)crossings),
                     v6 = wave_sum_f64((double)rejects), v7 = wave_sum_f64((double)lane_iters);
        if (lane == 0) {
            atomicAdd(&counters[MADI_CNT_CLASSIFY], (unsigned long long)v0);
            atomicAdd(&counters[MADI_CNT_SAFE_HIT], (unsigned long long)v1);
            atomicAdd(&counters[MADI_CNT_GRAPH_ITERS], (unsigned long long)v2);
            atomicAdd(&counters[MADI_CNT_FP64_RECHECK], (unsigned long long)v3);
            atomicAdd(&counters[MADI_CNT_FACETS], (unsigned long long)v4);
            atomicAdd(&counters[MADI_CNT_CROSSINGS], (unsigned long long)v5);
            atomicAdd(&counters[MADI_CNT_REJECTS], (unsigned long long)v6);
            atomicAdd(&counters[MADI_CNT_WAVE_ITERS], (unsigned long long)wave_iters);
            atomicAdd(&counters[MADI_CNT_LANE_ITERS], (unsigned long long)v7);
        }
        if (EXIT) {
            const double v8 = wave_sum_f64((double)n_exit);
            if (lane == 0) atomicAdd(&counters[MADI_CNT_EXITS], (unsigned long long)v8);
        }
    }
}

// ------------------------------------------------------------- split walk
//
// Walker waves and classifier waves (classify_mode 4).  In the per-lane state
// machine every wave runs the step code and the classification code in each
// iteration, each with only part of its lanes active.  Here the WW walker
// waves of a workgroup only step (tier hits complete at once) and send their
// tier misses, with the walker's first classification, through an LDS
// request ring to the workgroup's classifier wave, which keeps its 64 lanes
// busy with requests (the same per-lane classification state machine, one
// round of up to two edge groups per iteration) and writes each answer (label, moved
// point, new tier) into the asking lane's answer slot.  A walker lane waits
// for its answer; walker lanes are refilled with new walkers as in the
// streamed walk, and moments are stored for k_emit_tiles, so results are
// bitwise those of every other variant.
//
// LDS protocol: a producer reserves ring tickets with an atomic on `tail`,
// writes the request and then publishes it by storing ticket + 1 to the
// slot's `seq` with release order; the classifier reads slots in ticket
// order and takes one once `seq` (acquire) shows it.  Answers are published
// the same way through the answer slot's `status`.  The ring has a slot per
// walker lane and a lane has at most one request outstanding, so a slot is
// always consumed before its ticket comes round again.  No wait is
// unbounded: a walker wave that sees no lane progress for kSplitStallLimit
// iterations abandons its waiting walkers as classifier failures, and the
// classifier wave exits once every walker wave has finished and the ring is
// empty.

constexpr int kReqSlots = 256;  // >= walker lanes per workgroup (power of two)
constexpr uint32_t kSplitStallLimit = 1u << 26;

// 48 bytes: a ring of kReqSlots per classifier wave
struct ReqSlot {
    int na;
    int off;       // the anchor's edge list (< 0: not known, the classifier loads it)
    int len_own;   // bits 0-23: that list's padded length; bits 24-31: the requesting lane
    float hmax, alpha;  // the anchor's largest half length and annulus
    float qx, qy, qz, ox, oy, oz;
    int seq;       // last: published after the fields
};

// 144 bytes, the tier first (16-byte aligned in the array)
struct RespSlot {
    Tier t;
    int status, na, flags;  // flags: bit 0 inside, bits 1-3 tmode, bit 4 classifier failure
    float qx, qy, qz, ox, oy, oz;
};
static_assert(sizeof(ReqSlot) == 48 && sizeof(RespSlot) == 144, "split ring slots");
static_assert(sizeof(Tier) == 108, "tier layout of tier_lds_load / tier_lds_store");

typedef float v4f_t __attribute__((ext_vector_type(4)));
typedef float v3f_t __attribute__((ext_vector_type(3)));
typedef int v4i_t __attribute__((ext_vector_type(4)));

// A tier to or from LDS as six 16-byte accesses and one 12-byte access from
// one address (immediate offsets; the slots are 16-byte aligned).
__device__ __forceinline__ Tier tier_lds_load(const Tier& src)
{
    const v4f_t* s4 = (const v4f_t*)&src;
    const v4f_t a = s4[0], b = s4[1], c = s4[2], d = s4[3], e = s4[4], f = s4[5];
    const v3f_t g = *(const v3f_t*)((const float*)&src + 24);
    Tier t;
    t.ux = a.x, t.uy = a.y, t.uz = a.z, t.h = a.w;
    t.ux2 = b.x, t.uy2 = b.y, t.uz2 = b.z, t.h2 = b.w;
    t.ux3 = c.x, t.uy3 = c.y, t.uz3 = c.z, t.h3 = c.w;
    t.cx = d.x, t.cy = d.y, t.cz = d.z, t.ra_in2 = d.w;
    t.ra_out2 = e.x, t.rk_in2 = e.y, t.rk_out2 = e.z, t.rout2 = e.w;
    t.alpha_a = f.x, t.alpha_k = f.y, t.eps = f.z, t.k = __float_as_int(f.w);
    t.a_off = __float_as_int(g.x), t.a_end = __float_as_int(g.y), t.a_hmax = g.z;
    return t;
}

__device__ __forceinline__ void tier_lds_store(Tier& dst, const Tier& t)
{
    v4f_t* d4 = (v4f_t*)&dst;
    d4[0] = v4f_t{t.ux, t.uy, t.uz, t.h};
    d4[1] = v4f_t{t.ux2, t.uy2, t.uz2, t.h2};
    d4[2] = v4f_t{t.ux3, t.uy3, t.uz3, t.h3};
    d4[3] = v4f_t{t.cx, t.cy, t.cz, t.ra_in2};
    d4[4] = v4f_t{t.ra_out2, t.rk_in2, t.rk_out2, t.rout2};
    d4[5] = v4f_t{t.alpha_a, t.alpha_k, t.eps, __int_as_float(t.k)};
    *(v3f_t*)((float*)&dst + 24) = v3f_t{__int_as_float(t.a_off), __int_as_float(t.a_end), t.a_hmax};
}

template <int NP, int WW, int CW>
struct SplitLds {
    float evc[kMaxLdsGrid][NP];
    int occ[WW][kMaxLdsGrid];
    Rng states[WW][kClaimWalkers];
    ReqSlot req[CW][kReqSlots];  // one ring per classifier wave; walker wave w uses ring w mod CW
    RespSlot resp[WW * 64];
    GeomDev geo[WW];        // each walker wave's current entry (its requests are classified on it)
    unsigned int tail[CW];  // request tickets issued per ring
    unsigned int done;      // walker waves finished
};

__device__ __forceinline__ int lds_load_acquire(const int* p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void lds_store_release(int* p, int v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Verification build (VERIFY): the label the walk takes for a proposal,
// from the plane tier (kind 0), the exit certificate (1) or the classifier
// wave (2), against the exact FP64 KD classifier of madi_classify on the same
// represented point: the anchor's seed plus the FP32 offset q, summed in
// FP64.  A tier or classifier label (cell, inside) agrees when the inside
// flags are equal and, for an inside label, the cells are; a certified exit
// claims only that q is not inside the walker's cell a.  Counts per kind:
// checked, disagreeing; plus exact-classifier failures.
constexpr int kVerifyKinds = 3;

__device__ __forceinline__ void verify_label(const GeomDev& g, const KdView& v, int anchor, float qx, float qy,
                                             float qz, int kind, int cell, int inside, int a, uint32_t (&checked)[3],
                                             uint32_t (&bad)[3], uint32_t& failed)
{
    const double* c = g.cell_xyz + 3 * anchor;
    const double wx = gload(c) + (double)qx, wy = gload(c + 1) + (double)qy, wz = gload(c + 2) + (double)qz;
    int ecell = 0, ein = 0;
    long long nodes = 0;
    const int ok = classify_point_f64(g, v, wx, wy, wz, ecell, ein, nodes);
    const bool agree =
        kind == 1 ? !(ein && ecell == a) : (ein == inside && (!inside || ecell == cell));
    failed += !ok;
#pragma unroll
    for (int k = 0; k < kVerifyKinds; ++k)
        if (k == kind) {
            ++checked[k];
            bad[k] += ok && !agree;
        }
}

// One walker wave (wv < WW): the streamed walk's walker life cycle with the
// classification replaced by a request to the classifier wave.  VERIFY
// checks every label against the exact classifier (verify_label) and adds
// the counts to `verify` (MADI_N_VERIFY).
template <int NP, int WW, int CW, bool EXIT, bool VERIFY>
__device__ __forceinline__ void split_walker(const WorkDev* __restrict__ work, unsigned long long* item_next, const BatchHdr* hdr,
                             const PlanDev& plan, const Gf2Col* __restrict__ jump_pow, const Events& ev,
                             unsigned long long* counters, unsigned long long* verify, SplitLds<NP, WW, CW>& L,
                             int wv)
{
    enum : int { kStep = 0, kWait = 1, kDone = 3, kIdle = 4 };
    uint32_t v_checked[kVerifyKinds] = {0, 0, 0}, v_bad[kVerifyKinds] = {0, 0, 0}, v_failed = 0;
    const int lane = threadIdx.x & 63;
    const int32_t n_work = hdr->n_work;
    const int64_t total_w = hdr->total_walkers;
    int* occ = L.occ[wv];
    Rng* st = L.states[wv];
    RespSlot& my_resp = L.resp[wv * 64 + lane];
    for (int j = lane; j < kMaxLdsGrid; j += 64) occ[j] = 0;
    wave_lds_fence();
    uint32_t n_class = 0, n_tier = 0, crossings = 0, rejects = 0, n_exit = 0;
    uint32_t wave_iters = 0, lane_iters = 0;
    const int n_pairs = plan.n_pairs;
    if (n_work > 0 && total_w > 0) {
        const int64_t n_waves = (int64_t)gridDim.x * WW;
        const int64_t pos = ((int64_t)blockIdx.x * WW + wv) * total_w / n_waves;
        int lo = 0, hi = n_work - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (work[mid].walker_begin <= pos) lo = mid;
            else hi = mid - 1;
        }
        int e = __builtin_amdgcn_readfirstlane(lo);
        int dry = 0;
        bool aborted = false;
        while (dry < n_work) {
            const WorkDev* __restrict__ W = work + e;
            unsigned long long* next = item_next + e;
            int64_t claim = stream_claim(W, next, jump_pow, st, lane);
            if (claim >= W->n_walk) {
                ++dry;
                e = e + 1 == n_work ? 0 : e + 1;
                continue;
            }
            dry = 0;
            const GeomDev& g = W->g;
            const WalkDev& p = W->w;
            KdView kdv;  // the exact classifier reads every node from memory
            kdv.gf = g.kd_f;
            kdv.gi = g.kd_i;
            kdv.sf = nullptr;
            kdv.si = nullptr;
            kdv.n_lds = 0;
            // the classifier reads this wave's entry from LDS; no request of
            // this wave is outstanding while it changes
            if (lane == 0) L.geo[wv] = g;
            wave_lds_fence();
            const int64_t n_walk = W->n_walk, walker_lo = W->walker_lo;
            const float nscale = (float)(-2.0 * p.sigma * p.sigma * 0.6931471805599453);
            const float leave = EXIT ? exit_bound(g, p) : 0.f;  // EXIT: the certificate is compiled in
            const float ts = (float)p.ts;
            int64_t hand = claim, hand_end = claim + kClaimWalkers < n_walk ? claim + kClaimWalkers : n_walk;
            bool item_dry = false;

            int phase = kIdle, init = 0, enq = 0;
            unsigned int ticket = 0;
            int64_t w = 0;
            Rng rng{0, 0};
            int a = 0, in = 0, cell = 0, err = 0, escaped = 0;
            uint32_t n_esc = 0, n_err = 0, stall = 0;
            float px = 0.f, py = 0.f, pz = 0.f, ox = 0.f, oy = 0.f, oz = 0.f;
            float lox = 0.f, loy = 0.f, loz = 0.f, hix = 0.f, hiy = 0.f, hiz = 0.f;
            Tier t;
            tier_clear(t);
            float Sx = 0.f, Sy = 0.f, Sz = 0.f, Cx = 0.f, Cy = 0.f, Cz = 0.f, hx = 0.f, hy = 0.f, hz = 0.f;
            float dM[NP][3];
#pragma unroll
            for (int q = 0; q < NP; ++q) dM[q][0] = dM[q][1] = dM[q][2] = 0.f;
            int step = 0, until_sample = 0, j_sample = 0, next_ev = 0, rj = 0;
            float qx = 0.f, qy = 0.f, qz = 0.f, qox = 0.f, qoy = 0.f, qoz = 0.f;
            int qna = 0;
            for (;;) {
                uint64_t need = __ballot(phase == kIdle);
                if (need != 0 && !item_dry) {
                    bool got = false;
                    while (need != 0) {
                        if (hand >= hand_end) {
                            const int64_t nb = stream_claim(W, next, jump_pow, st, lane);
                            if (nb >= n_walk) {
                                item_dry = true;
                                break;
                            }
                            claim = nb;
                            hand = nb;
                            hand_end = nb + kClaimWalkers < n_walk ? nb + kClaimWalkers : n_walk;
                        }
                        const int64_t rank = (int64_t)lanes_below(need);
                        const bool mine = ((need >> lane) & 1ull) && rank < hand_end - hand;
                        if (mine) {
                            w = hand + rank;
                            rng = st[w - claim];
                            got = true;
                        }
                        const uint64_t taken = __ballot(mine);
                        hand += __popcll(taken);
                        need &= ~taken;
                        wave_lds_fence();
                    }
                    if (got) {
                        const double span = g.source_hi - g.source_lo;
                        const double x0 = g.source_lo + uniform_f64(rng) * span;
                        const double y0 = g.source_lo + uniform_f64(rng) * span;
                        const double z0 = g.source_lo + uniform_f64(rng) * span;
                        lox = (float)(-x0);
                        loy = (float)(-y0);
                        loz = (float)(-z0);
                        hix = (float)(g.L - x0);
                        hiy = (float)(g.L - y0);
                        hiz = (float)(g.L - z0);
                        Sx = Sy = Sz = Cx = Cy = Cz = hx = hy = hz = 0.f;
#pragma unroll
                        for (int q = 0; q < NP; ++q) dM[q][0] = dM[q][1] = dM[q][2] = 0.f;
                        step = 0;
                        until_sample = p.steps_per_h;
                        j_sample = 0;
                        next_ev = 0;
                        err = escaped = 0;
                        if (g.free_water) {
                            tier_free_water(t);
                            a = cell = in = 0;
                            px = py = pz = ox = oy = oz = 0.f;
                            init = 0;
                            phase = p.n_steps > 0 ? kStep : kDone;
                        } else {
                            anchor_world(g, x0, y0, z0, qna, qx, qy, qz);
                            qox = -qx;
                            qoy = -qy;
                            qoz = -qz;
                            ++n_class;
                            init = 1;
                            enq = 1;
                            phase = kWait;
                        }
                    }
                }
                if (__ballot(phase != kIdle) == 0) break;
                ++wave_iters;
                lane_iters += phase != kIdle;
                int complete = 0, fromq = 0, rflags = 0, rna = 0, vkind = 2;
                if (phase == kStep) {
                    const float dx = normal_scaled_f32(rng, nscale);
                    const float dy = normal_scaled_f32(rng, nscale);
                    const float dz = normal_scaled_f32(rng, nscale);
                    const uint64_t ku = rng_next(rng) >> 11;
                    rj = (ku >= p.rej1 ? 1 : 0) | (ku >= p.rej2 ? 2 : 0);
                    qx = px + dx;
                    qy = py + dy;
                    qz = pz + dz;
                    qox = ox;
                    qoy = oy;
                    qoz = oz;
                    qna = a;
                    const float nDx = ox + qx, nDy = oy + qy, nDz = oz + qz;
                    int pcell = 0, pin = 0;
                    if (!g.periodic && (nDx < lox || nDx >= hix || nDy < loy || nDy >= hiy || nDz < loz || nDz >= hiz)) {
                        escaped = 1;
                        phase = kDone;
                        if (in)
                            for (int j = j_sample + 1; j < p.n_grid; ++j) atomicAdd(&occ[j], 1);
                    } else if (tier_query(t, a, qx, qy, qz, pcell, pin)) {
                        ++n_tier;
                        vkind = 0;
                        complete = 1;
                        rna = pcell;
                        rflags = pin;
                    } else if (EXIT && leave > 0.f && in && cell == a && rj == 3 && tier_exit(t, a, qx, qy, qz, leave)) {
                        ++n_tier;  // a certified rejected exit (see stream_wave)
                        ++n_exit;
                        vkind = 1;
                        complete = 1;
                        rna = a;
                        rflags = 0;
                    } else {
                        ++n_class;
                        enq = 1;
                        phase = kWait;
                    }
                }
                // publish this iteration's requests
                {
                    const uint64_t em = __ballot(enq != 0);
                    if (em != 0) {
                        const int first = __builtin_ctzll(em);
                        unsigned int base = 0;
                        if (lane == first) base = atomicAdd(&L.tail[wv % CW], (unsigned int)__popcll(em));
                        base = __builtin_amdgcn_readfirstlane(__shfl(base, first, 64));
                        if (enq) {
                            ticket = base + lanes_below(em);
                            ReqSlot& rq = L.req[wv % CW][ticket & (kReqSlots - 1)];
                            // a step's anchor is the tier's: its record rides along
                            *(v4i_t*)&rq.na = v4i_t{qna, init ? -1 : t.a_off,
                                                    ((t.a_end - t.a_off) & 0xffffff) | ((wv * 64 + lane) << 24),
                                                    __float_as_int(t.a_hmax)};
                            *(v4f_t*)&rq.alpha = v4f_t{t.alpha_a, qx, qy, qz};
                            *(v3f_t*)&rq.ox = v3f_t{qox, qoy, qoz};
                            lds_store_release(&rq.seq, (int)(ticket + 1u));
                            enq = 0;
                        }
                    }
                }
                if (phase == kWait && lds_load_acquire(&my_resp.status) == (int)(ticket + 1u)) {
                    rflags = my_resp.flags;
                    rna = my_resp.na;
                    qx = my_resp.qx;
                    qy = my_resp.qy;
                    qz = my_resp.qz;
                    qox = my_resp.ox;
                    qoy = my_resp.oy;
                    qoz = my_resp.oz;
                    complete = 1;
                    fromq = 1;
                    if (rflags & 16) {
                        complete = 0;
                        err = 1;
                        escaped = 1;
                        phase = kDone;
                    } else {
                        phase = kStep;
                    }
                }
                // a tier label or a certified exit is for q relative to a; a
                // classifier answer moved q to its anchor rna
                if (VERIFY && complete)
                    verify_label(g, kdv, vkind == 2 ? rna : a, qx, qy, qz, vkind, rna, rflags & 1, a, v_checked,
                                 v_bad, v_failed);
                if (complete && init) {
                    // the walker's first label: no step, no acceptance test
                    t = tier_lds_load(my_resp.t);
                    a = rna;
                    in = rflags & 1;
                    cell = rna;
                    px = qx;
                    py = qy;
                    pz = qz;
                    ox = -qx;
                    oy = -qy;
                    oz = -qz;
                    if (in) atomicAdd(&occ[0], 1);
                    init = 0;
                    phase = p.n_steps > 0 ? kStep : kDone;
                } else if (complete) {
                    const int pin = rflags & 1, pcell = rna;
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
                        if (fromq) t = tier_lds_load(my_resp.t);
                        a = fromq ? rna : a;
                        in = pin;
                        cell = pcell;
                        px = qx;
                        py = qy;
                        pz = qz;
                        ox = qox;
                        oy = qoy;
                        oz = qoz;
                    }
                    const float Dx = ox + px, Dy = oy + py, Dz = oz + pz;
                    hx += Dx;
                    hy += Dy;
                    hz += Dz;
                    ++step;
                    if (--until_sample == 0) {
                        until_sample = p.steps_per_h;
                        const int j = ++j_sample;
                        // fold the interval's plain sum into the compensated total
                        const float yx = hx - Cx, tx = Sx + yx;
                        Cx = (tx - Sx) - yx;
                        Sx = tx;
                        const float yy = hy - Cy, ty = Sy + yy;
                        Cy = (ty - Sy) - yy;
                        Sy = ty;
                        const float yz = hz - Cz, tz = Sz + yz;
                        Cz = (tz - Sz) - yz;
                        Sz = tz;
                        hx = hy = hz = 0.f;
                        if (in) atomicAdd(&occ[j], 1);
                        const float Yx = ts * ((Sx - Cx) - 0.5f * Dx);
                        const float Yy = ts * ((Sy - Cy) - 0.5f * Dy);
                        const float Yz = ts * ((Sz - Cz) - 0.5f * Dz);
                        if (W->y_out) {
                            // the walker's Y sample, widened to FP64
                            double* yo = W->y_out + ((walker_lo + w) * (int64_t)p.n_grid + j) * 3;
                            yo[0] = (double)Yx;
                            yo[1] = (double)Yy;
                            yo[2] = (double)Yz;
                        }
#pragma unroll
                        for (int u = 0; u < NP; ++u) {
                            const float sg = L.evc[j][u];
                            dM[u][0] = fmaf(sg, Yx, dM[u][0]);
                            dM[u][1] = fmaf(sg, Yy, dM[u][1]);
                            dM[u][2] = fmaf(sg, Yz, dM[u][2]);
                        }
                    }
                    phase = step >= p.n_steps ? kDone : kStep;
                }
                if (phase == kDone) {
                    const int64_t ew = walker_lo + w;
                    float* m = W->moments + ew * (int64_t)(n_pairs * 3);
#pragma unroll
                    for (int u = 0; u < NP; ++u)
                        if (u < n_pairs) {
                            m[3 * u] = dM[u][0];
                            m[3 * u + 1] = dM[u][1];
                            m[3 * u + 2] = dM[u][2];
                        }
                    W->active[ew] = escaped ? 0 : 1;
                    n_esc += escaped;
                    n_err += err;
                    phase = kIdle;
                }
                // bounded wait: a wave whose lanes all wait makes no progress;
                // past the limit its waiting walkers count as classifier
                // failures (fatal for the entry) and the wave stops walking,
                // keeping its geometry slot for the requests still queued
                if (__ballot(phase != kWait && phase != kIdle) == 0 && __ballot(phase == kWait) != 0) {
                    __builtin_amdgcn_s_sleep(1);  // leave the issue slots to the classifier wave
                    if (++stall > kSplitStallLimit) {
                        n_err += phase == kWait;
                        n_esc += phase == kWait;
                        aborted = true;
                        break;
                    }
                } else {
                    stall = 0;
                }
            }
            {
                wave_lds_fence();
                for (int j = lane; j < p.n_grid; j += 64) {
                    const int v = occ[j];
                    if (v) atomicAdd((unsigned long long*)&W->occupancy[j], (unsigned long long)v);
                    occ[j] = 0;
                }
                wave_lds_fence();
            }
            {
                const double ne = wave_sum_f64((double)n_esc), nr = wave_sum_f64((double)n_err);
                if (lane == 0) {
                    if (ne > 0.0) atomicAdd((unsigned long long*)&W->flags[MADI_FLAG_ESCAPED], (unsigned long long)ne);
                    if (nr > 0.0)
                        atomicAdd((unsigned long long*)&W->flags[MADI_FLAG_CLASSIFIER_ERR], (unsigned long long)nr);
                }
            }
            if (aborted) break;
            e = e + 1 == n_work ? 0 : e + 1;
        }
    }
    if (lane == 0) __hip_atomic_fetch_add(&L.done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (VERIFY && verify) {
#pragma unroll
        for (int k = 0; k < kVerifyKinds; ++k) {
            const double c = wave_sum_f64((double)v_checked[k]), b = wave_sum_f64((double)v_bad[k]);
            if (lane == 0) {
                atomicAdd(&verify[2 * k], (unsigned long long)c);
                atomicAdd(&verify[2 * k + 1], (unsigned long long)b);
            }
        }
        const double f = wave_sum_f64((double)v_failed);
        if (lane == 0) atomicAdd(&verify[MADI_VERIFY_EXACT_FAILED], (unsigned long long)f);
    }
    if (counters) {
        const double v0 = wave_sum_f64((double)n_class), v1 = wave_sum_f64((double)n_tier),
                     v5 = wave_sum_f64((double)crossings), v6 = wave_sum_f64((double)rejects),
                     v7 = wave_sum_f64((double)lane_iters);
        if (lane == 0) {
            atomicAdd(&counters[MADI_CNT_CLASSIFY], (unsigned long long)v0);
            atomicAdd(&counters[MADI_CNT_SAFE_HIT], (unsigned long long)v1);
            atomicAdd(&counters[MADI_CNT_CROSSINGS], (unsigned long long)v5);
            atomicAdd(&counters[MADI_CNT_REJECTS], (unsigned long long)v6);
            atomicAdd(&counters[MADI_CNT_WAVE_ITERS], (unsigned long long)wave_iters);
            atomicAdd(&counters[MADI_CNT_LANE_ITERS], (unsigned long long)v7);
        }
        if (EXIT) {
            const double v8 = wave_sum_f64((double)n_exit);
            if (lane == 0) atomicAdd(&counters[MADI_CNT_EXITS], (unsigned long long)v8);
        }
    }
}

// The classifier wave: requests taken in ticket order into idle lanes,
// classified with the per-lane state machine on the asking wave's entry (its
// geometry slot), answered with the new tier.  Lanes may work on different
// entries.
template <int NP, int WW, int CW>
__device__ __forceinline__ void split_classifier(float cap_width, unsigned long long* counters, SplitLds<NP, WW, CW>& L, int cw)
{
    const int lane = threadIdx.x & 63;
    // walker lanes wait for this wave's answers: let it issue ahead of the
    // walker waves sharing its SIMD
    __builtin_amdgcn_s_setprio(3);
    uint32_t nfac = 0, iters = 0, facets = 0, rechecks = 0, c_iters = 0, c_busy = 0;
    int busy = 0, owner = 0, cphase = 0;
    unsigned int ticket = 0, head = 0;
    Pending P;
    P.qx = P.qy = P.qz = P.ox = P.oy = P.oz = 0.f;
    P.na = P.pin = P.moves = 0;
    P.sa_c = P.sa_c2 = P.sa_c3 = -1;
    P.sa_m1 = P.sa_m2 = P.sa_r = P.sa_eps = 0.f;
    ScanState S;
    S.cur = S.end = 0;
    S.c = S.c2 = S.c3 = -1;
    S.cell = 0;
    S.m1 = S.m2 = S.m3 = S.m4 = S.xb = S.pe = S.eps = S.nx = S.ny = S.nz = 0.f;
    for (;;) {
        const uint64_t need = __ballot(!busy);
        if (need != 0) {
            const unsigned int tk = head + (unsigned int)lane;
            const int sq = lds_load_acquire(&L.req[cw][tk & (kReqSlots - 1)].seq);
            const uint64_t rm = __ballot(sq == (int)(tk + 1u));
            const int n_ready = ~rm == 0ull ? 64 : (int)__builtin_ctzll(~rm);
            const int n_need = (int)__popcll(need);
            const int take = n_ready < n_need ? n_ready : n_need;
            const int rank = (int)lanes_below(need);
            if (!busy && rank < take) {
                ticket = head + (unsigned int)rank;
                const ReqSlot& rq = L.req[cw][ticket & (kReqSlots - 1)];
                const v4i_t r0 = *(const v4i_t*)&rq.na;
                const v4f_t r1 = *(const v4f_t*)&rq.alpha;
                const v3f_t r2 = *(const v3f_t*)&rq.ox;
                P.na = r0.x;
                const int len_own = r0.z;
                owner = (int)((unsigned)len_own >> 24);
                P.qx = r1.y;
                P.qy = r1.z;
                P.qz = r1.w;
                P.ox = r2.x;
                P.oy = r2.y;
                P.oz = r2.z;
                const int off = r0.y, end = off + (len_own & 0xffffff);
                const float hmax = __int_as_float(r0.w), alpha = r1.x;
                busy = 1;
                pending_start(P, S, cphase);
                if (off >= 0) scan_setup(S, off, end, hmax, alpha);
            }
            head += (unsigned int)take;
        }
        if (__ballot(busy) == 0) {
            if (lds_load_acquire((const int*)&L.done) == WW &&
                head == (unsigned int)__hip_atomic_load(&L.tail[cw], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP))
                break;
            __builtin_amdgcn_s_sleep(2);
            // no lane holds a classification: its state is dead here
            P.qx = P.qy = P.qz = P.ox = P.oy = P.oz = 0.f;
            P.na = P.pin = P.moves = 0;
            P.sa_c = P.sa_c2 = P.sa_c3 = -1;
            P.sa_m1 = P.sa_m2 = P.sa_r = P.sa_eps = 0.f;
            S.cur = S.end = 0;
            S.c = S.c2 = S.c3 = -1;
            S.cell = 0;
            S.m1 = S.m2 = S.m3 = S.m4 = S.xb = S.pe = S.eps = S.nx = S.ny = S.nz = S.alpha = 0.f;
            owner = cphase = 0;
            continue;
        }
        ++c_iters;
        c_busy += busy;
        const GeomDev& g = L.geo[owner >> 6];
        const float cap = g.alpha_max + cap_width;
        // an anchor outside the entry's cells cannot come from a live request;
        // keep every access in bounds regardless
        if (busy && (unsigned)P.na >= (unsigned)g.n_cells) P.na = 0;
        int res = 0, tmode = 0;
        if (busy) res = classify_round<true>(g, cap, cphase, S, P, tmode, nfac, iters, facets, rechecks);
        if (res != 0) {
            RespSlot& rs = L.resp[owner];
            if (res > 0) {
                Tier t;
                tier_of(g, P, tmode, S, cap, t);
                tier_lds_store(rs.t, t);
            }
            rs.na = P.na;
            rs.flags = (P.pin & 1) | (tmode << 1) | (res < 0 ? 16 : 0);
            rs.qx = P.qx;
            rs.qy = P.qy;
            rs.qz = P.qz;
            rs.ox = P.ox;
            rs.oy = P.oy;
            rs.oz = P.oz;
            lds_store_release(&rs.status, (int)(ticket + 1u));
            busy = 0;
        }
    }
    facets += nfac;
    if (counters) {
        const double v2 = wave_sum_f64((double)iters), v3 = wave_sum_f64((double)rechecks),
                     v4 = wave_sum_f64((double)facets), v8 = wave_sum_f64((double)c_busy);
        if (lane == 0) {
            atomicAdd(&counters[MADI_CNT_GRAPH_ITERS], (unsigned long long)v2);
            atomicAdd(&counters[MADI_CNT_FP64_RECHECK], (unsigned long long)v3);
            atomicAdd(&counters[MADI_CNT_FACETS], (unsigned long long)v4);
            atomicAdd(&counters[MADI_CNT_CLASS_ITERS], (unsigned long long)c_iters);
            atomicAdd(&counters[MADI_CNT_CLASS_BUSY], (unsigned long long)v8);
        }
    }
}

// One entry: block b runs tile b.
template <int NP, int R>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_walk_f32(
    GeomDev g, WalkDev p, const Gf2Col* __restrict__ jump_pow, FusedOut fo, float cap_width,
    unsigned long long* counters)
{
    __shared__ TileLds<NP> lds;
    TileOut out;
    out.slot = fo.partial + (size_t)(fo.tile_base + blockIdx.x) * fo.plan->n_cols * 2;
    out.occupancy = fo.occupancy;
    out.flags = fo.flags;
    walk_tile_f32<NP, R>(g, p, jump_pow, (int64_t)blockIdx.x * kBlock, *fo.plan, fo.ev, out, cap_width, counters, lds);
}

// Persistent work-stealing launch over many entries: each workgroup takes
// the next global tile index from the header's counter, finds its work item
// by binary search on work_begin, and runs that tile.  Every workgroup exits
// once the counter passes the total, so the grid drains (workgroups beyond
// the total exit at once).
template <int NP, int R>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_batch_f32(
    const WorkDev* __restrict__ work, BatchHdr* hdr, const PlanDev* __restrict__ plan, Events ev,
    const Gf2Col* __restrict__ jump_pow, float cap_width, unsigned long long* counters)
{
    __shared__ TileLds<NP> lds;
    __shared__ long long s_tile;
    __shared__ int s_item;
    const int64_t total = hdr->total;
    const int32_t n_work = hdr->n_work;
    for (;;) {
        if (threadIdx.x == 0) {
            const long long t = (long long)atomicAdd(&hdr->queue, 1ull);
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
        walk_tile_f32<NP, R>(W.g, W.w, jump_pow, tile * kBlock, *plan, ev, out, cap_width, counters, lds);
    }
}

// Streamed walk over a batch (see stream_wave).  The event table is the
// batch's, so every workgroup builds it once.  EXIT compiles in the exit
// certificate (tier_exit; it costs registers, so launches whose entries never
// use it go without).
template <int NP, int R, bool EXIT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_stream_f32(
    const WorkDev* __restrict__ work, unsigned long long* __restrict__ item_next, const BatchHdr* __restrict__ hdr,
    const PlanDev* __restrict__ plan, Events ev, const Gf2Col* __restrict__ jump_pow, float cap_width,
    unsigned long long* counters)
{
    __shared__ StreamLds<NP> lds;
    for (int j = threadIdx.x; j < kMaxLdsGrid; j += kBlock)
#pragma unroll
        for (int u = 0; u < NP; ++u) lds.evc[j][u] = 0.f;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int e = 0; e < ev.n; ++e)
            if (ev.j[e] < kMaxLdsGrid) lds.evc[ev.j[e]][ev.pair[e]] += (float)ev.sign[e];
    __syncthreads();
    stream_wave<NP, R, EXIT>(work, item_next, hdr, *plan, jump_pow, ev, cap_width, counters, lds);
}

// Split walk over a batch (see split_walker / split_classifier): WW walker
// waves and CW classifier waves per workgroup; EXIT as for k_stream_f32.
template <int NP, int WW, int CW, bool EXIT, bool VERIFY>
__global__ __launch_bounds__((WW + CW) * 64) __attribute__((amdgpu_waves_per_eu(3))) void k_split_f32(
    const WorkDev* __restrict__ work, unsigned long long* __restrict__ item_next, const BatchHdr* __restrict__ hdr,
    const PlanDev* __restrict__ plan, Events ev, const Gf2Col* __restrict__ jump_pow, float cap_width,
    unsigned long long* counters, unsigned long long* verify)
{
    static_assert(kReqSlots >= 64 * ((WW + CW - 1) / CW), "a ring needs a slot per walker lane it serves");
    static_assert(WW <= 4, "a request names its lane in 8 bits");
    __shared__ SplitLds<NP, WW, CW> lds;
    constexpr int kThreads = (WW + CW) * 64;
    for (int j = threadIdx.x; j < kMaxLdsGrid; j += kThreads)
#pragma unroll
        for (int u = 0; u < NP; ++u) lds.evc[j][u] = 0.f;
    for (int j = threadIdx.x; j < CW * kReqSlots; j += kThreads) lds.req[j / kReqSlots][j % kReqSlots].seq = 0;
    for (int j = threadIdx.x; j < WW * 64; j += kThreads) lds.resp[j].status = 0;
    if (threadIdx.x < CW) lds.tail[threadIdx.x] = 0;
    if (threadIdx.x == 0) lds.done = 0;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int e = 0; e < ev.n; ++e)
            if (ev.j[e] < kMaxLdsGrid) lds.evc[ev.j[e]][ev.pair[e]] += (float)ev.sign[e];
    __syncthreads();
    // roles rotate with the workgroup index, so classifier waves spread over
    // the CU's SIMDs
    const int r = (int)((threadIdx.x / 64 + blockIdx.x) % (WW + CW));
    if (r >= WW) split_classifier<NP, WW, CW>(cap_width, counters, lds, r - WW);
    else split_walker<NP, WW, CW, EXIT, VERIFY>(work, item_next, hdr, *plan, jump_pow, ev, counters, verify, lds, r);
}

// Column sums of every tile of a streamed batch from its walkers' stored
// moments, bit for bit those of emit_columns: per column, each group of 64
// walkers is summed in FP64 as wave_sum_f64's butterfly does (lane l with
// l ^ 32 first, then with l ^ 16, ...: the pairwise tree over the walkers
// taken in bit-reversed order of their index in the group), and the four
// group sums are added to 0.0 in group order.  Here a thread owns whole
// columns and forms that tree with a carry stack, so the sums need no
// cross-lane traffic or barriers.
template <int NP>
__global__ __launch_bounds__(kBlock) void k_emit_tiles(const WorkDev* __restrict__ work, const BatchHdr* hdr,
                                                       const PlanDev* __restrict__ plan)
{
    __shared__ float mom[kBlock][3 * NP];
    __shared__ uint8_t act[kBlock];
    const int64_t total = hdr->total;
    const int32_t n_work = hdr->n_work;
    const PlanDev& P = *plan;
    const int n_pairs = P.n_pairs, n_b = P.n_b, n_dirs = P.n_dirs, n_cols = P.n_cols;
    for (int64_t t = blockIdx.x; t < total; t += gridDim.x) {
        int lo = 0, hi = n_work - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (work[mid].work_begin <= t) lo = mid;
            else hi = mid - 1;
        }
        const WorkDev& W = work[lo];
        const int64_t tile = W.tile_lo + (t - W.work_begin);
        {
            const int64_t w = tile * kBlock + threadIdx.x;
            const bool live = w < W.w.n_walkers;
            act[threadIdx.x] = live ? W.active[w] : 0;
            const float* m = W.moments + (live ? w : 0) * (int64_t)(n_pairs * 3);
#pragma unroll
            for (int u = 0; u < NP; ++u)
#pragma unroll
                for (int c = 0; c < 3; ++c) mom[threadIdx.x][3 * u + c] = (live && u < n_pairs) ? m[3 * u + c] : 0.f;
        }
        __syncthreads();
        double* slot = W.partial + (size_t)tile * n_cols * 2;
        for (int col = threadIdx.x; col < n_cols; col += kBlock) {
            const int d = col % n_dirs, pb = col / n_dirs, b = pb % n_b, p = pb / n_b;
            double a_c = 0.0, a_s = 0.0;
            for (int grp = 0; grp < kWaves; ++grp) {
                double st_c[7], st_s[7];
#pragma unroll
                for (int k = 0; k < 64; ++k) {
                    // bit-reversed 6-bit index: the butterfly's pairing order
                    const int l = ((k & 1) << 5) | ((k & 2) << 3) | ((k & 4) << 1) | ((k & 8) >> 1) |
                                  ((k & 16) >> 3) | ((k & 32) >> 5);
                    const int wi = grp * 64 + l;
                    float sv, cv;
                    sincos_f32(column_phase_f32(P, p, b, d, mom[wi][3 * p], mom[wi][3 * p + 1], mom[wi][3 * p + 2]),
                               sv, cv);
                    double vc = act[wi] ? (double)cv : 0.0, vs = act[wi] ? (double)sv : 0.0;
                    int lvl = 0;
#pragma unroll
                    for (int q = k; q & 1; q >>= 1, ++lvl) {
                        vc = st_c[lvl] + vc;
                        vs = st_s[lvl] + vs;
                    }
                    st_c[lvl] = vc;
                    st_s[lvl] = vs;
                }
                a_c += st_c[6];
                a_s += st_s[6];
            }
            slot[2 * col] = a_c;
            slot[2 * col + 1] = a_s;
        }
        __syncthreads();
    }
}

// Per column, an entry's tile slots are summed in a fixed order: chunks of
// kFinalChunk consecutive tiles from 0.0 in tile order, then the chunk sums
// from 0.0 in chunk order.  Pass 1: one thread per (chunk, column).
constexpr int kFinalChunk = 32;

__global__ void k_finalize_chunks(const double* __restrict__ partial, int64_t tile_lo, int64_t n_tiles, int32_t n_cols,
                                  int64_t n_chunks, double* __restrict__ chunk_out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_chunks * n_cols) return;
    const int64_t q = i / n_cols;
    const int col = (int)(i - q * n_cols);
    const int64_t t0 = q * kFinalChunk, t1 = t0 + kFinalChunk < n_tiles ? t0 + kFinalChunk : n_tiles;
    double a = 0.0, b = 0.0;
    for (int64_t t = t0; t < t1; ++t) {
        const double* s = partial + ((size_t)(tile_lo + t) * n_cols + col) * 2;
        a += s[0];
        b += s[1];
    }
    chunk_out[(q * n_cols + col) * 2] = a;
    chunk_out[(q * n_cols + col) * 2 + 1] = b;
}

// Pass 2: one thread per column adds the chunk sums in chunk order.
__global__ void k_finalize_sum(const double* __restrict__ chunk, int64_t n_chunks, int32_t n_cols,
                               double* __restrict__ cos_out, double* __restrict__ sin_out)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= n_cols) return;
    double a = 0.0, b = 0.0;
    for (int64_t q = 0; q < n_chunks; ++q) {
        a += chunk[(q * n_cols + col) * 2];
        b += chunk[(q * n_cols + col) * 2 + 1];
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

#define MADI_DISPATCH_R(NPV, R_VAL, LAUNCH) \
    do {                                    \
        if ((R_VAL) <= 1) LAUNCH(NPV, 1);   \
        else LAUNCH(NPV, 2);                \
    } while (0)
#define MADI_DISPATCH_NP(NP_VAL, R_VAL, LAUNCH)                      \
    do {                                                             \
        const int np_ = (NP_VAL);                                    \
        if (np_ <= 2) MADI_DISPATCH_R(2, R_VAL, LAUNCH);             \
        else if (np_ <= 4) MADI_DISPATCH_R(4, R_VAL, LAUNCH);        \
        else MADI_DISPATCH_R(MADI_MAX_PAIRS, R_VAL, LAUNCH);         \
    } while (0)

int launch_walk_f32_fused(const GeomDev& g, const WalkDev& p, const Gf2Col* jump_pow, const FusedOut& fo,
                          float cap_width, int rounds, unsigned long long* counters, hipStream_t s)
{
    if (p.n_walkers == 0) return MADI_OK;
    const int64_t blocks = (p.n_walkers + kBlock - 1) / kBlock;
#define MADI_LAUNCH_WALK(NPV, RV)                                                                                   \
    hipLaunchKernelGGL((k_walk_f32<NPV, RV>), dim3((unsigned)blocks), dim3(kBlock), 0, s, g, p, jump_pow, fo,       \
                       cap_width, counters)
    MADI_DISPATCH_NP(fo.ev.n_pairs, rounds, MADI_LAUNCH_WALK);
#undef MADI_LAUNCH_WALK
    return check_launch("k_walk_f32");
}

int launch_batch_f32(const WorkDev* work, BatchHdr* hdr, const PlanDev* plan, const Events& ev, const Gf2Col* jump_pow,
                     float cap_width, int rounds, int32_t grid, unsigned long long* counters, hipStream_t s)
{
#define MADI_LAUNCH_BATCH(NPV, RV)                                                                                  \
    hipLaunchKernelGGL((k_batch_f32<NPV, RV>), dim3(grid), dim3(kBlock), 0, s, work, hdr, plan, ev, jump_pow,       \
                       cap_width, counters)
    MADI_DISPATCH_NP(ev.n_pairs, rounds, MADI_LAUNCH_BATCH);
#undef MADI_LAUNCH_BATCH
    return check_launch("k_batch_f32");
}

int launch_stream_f32(const WorkDev* work, unsigned long long* item_next, const BatchHdr* hdr, const PlanDev* plan,
                      const Events& ev, const Gf2Col* jump_pow, float cap_width, int rounds, bool exits, int32_t grid,
                      unsigned long long* counters, hipStream_t s)
{
#define MADI_LAUNCH_STREAM_X(NPV, RV, XV)                                                                           \
    hipLaunchKernelGGL((k_stream_f32<NPV, RV, XV>), dim3(grid), dim3(kBlock), 0, s, work, item_next, hdr, plan, ev, \
                       jump_pow, cap_width, counters)
#define MADI_LAUNCH_STREAM(NPV, RV)                                \
    do {                                                           \
        if (exits) MADI_LAUNCH_STREAM_X(NPV, RV, true);            \
        else MADI_LAUNCH_STREAM_X(NPV, RV, false);                 \
    } while (0)
    MADI_DISPATCH_NP(ev.n_pairs, rounds, MADI_LAUNCH_STREAM);
#undef MADI_LAUNCH_STREAM
#undef MADI_LAUNCH_STREAM_X
    return check_launch("k_stream_f32");
}

int launch_split_f32(const WorkDev* work, unsigned long long* item_next, const BatchHdr* hdr, const PlanDev* plan,
                     const Events& ev, const Gf2Col* jump_pow, float cap_width, int shape, bool exits, int32_t grid,
                     unsigned long long* counters, unsigned long long* verify, hipStream_t s)
{
#define MADI_LAUNCH_SPLIT_X(NPV, WWV, CWV, XV, VV)                                                              \
    hipLaunchKernelGGL((k_split_f32<NPV, WWV, CWV, XV, VV>), dim3(grid), dim3((WWV + CWV) * 64), 0, s, work,    \
                       item_next, hdr, plan, ev, jump_pow, cap_width, counters, verify)
#define MADI_LAUNCH_SPLIT(NPV, WWV, CWV)                                 \
    do {                                                                 \
        if (exits) MADI_LAUNCH_SPLIT_X(NPV, WWV, CWV, true, false);      \
        else MADI_LAUNCH_SPLIT_X(NPV, WWV, CWV, false, false);           \
    } while (0)
    // the verification build exists in the three-walker shape only (results
    // are the same bits in every shape)
#define MADI_SPLIT_SHAPE(NPV)                                                 \
    do {                                                                      \
        if (verify) {                                                         \
            if (exits) MADI_LAUNCH_SPLIT_X(NPV, 3, 1, true, true);            \
            else MADI_LAUNCH_SPLIT_X(NPV, 3, 1, false, true);                 \
        } else if (shape <= 1) MADI_LAUNCH_SPLIT(NPV, 3, 1);                  \
        else MADI_LAUNCH_SPLIT(NPV, 2, 2);                                    \
    } while (0)
    const int np = ev.n_pairs;
    if (np <= 2) MADI_SPLIT_SHAPE(2);
    else if (np <= 4) MADI_SPLIT_SHAPE(4);
    else MADI_SPLIT_SHAPE(MADI_MAX_PAIRS);
#undef MADI_SPLIT_SHAPE
#undef MADI_LAUNCH_SPLIT
#undef MADI_LAUNCH_SPLIT_X
    return check_launch("k_split_f32");
}

int launch_emit_tiles(const WorkDev* work, const BatchHdr* hdr, const PlanDev* plan, int n_pairs, int32_t grid,
                      hipStream_t s)
{
    if (n_pairs <= 2) hipLaunchKernelGGL((k_emit_tiles<2>), dim3(grid), dim3(kBlock), 0, s, work, hdr, plan);
    else if (n_pairs <= 4) hipLaunchKernelGGL((k_emit_tiles<4>), dim3(grid), dim3(kBlock), 0, s, work, hdr, plan);
    else hipLaunchKernelGGL((k_emit_tiles<MADI_MAX_PAIRS>), dim3(grid), dim3(kBlock), 0, s, work, hdr, plan);
    return check_launch("k_emit_tiles");
}

int64_t finalize_scratch_doubles(int64_t n_tiles, int32_t n_cols)
{
    return ((n_tiles + kFinalChunk - 1) / kFinalChunk) * (int64_t)n_cols * 2;
}

int launch_finalize_tiles(const double* partial, int64_t tile_lo, int64_t n_tiles, int32_t n_cols, double* cos_out,
                          double* sin_out, double* scratch, hipStream_t s)
{
    if (n_cols == 0) return MADI_OK;
    const int64_t n_chunks = (n_tiles + kFinalChunk - 1) / kFinalChunk;
    if (n_chunks > 0) {
        const int64_t n = n_chunks * n_cols;
        hipLaunchKernelGGL(k_finalize_chunks, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, partial, tile_lo,
                           n_tiles, n_cols, n_chunks, scratch);
    }
    hipLaunchKernelGGL(k_finalize_sum, dim3((n_cols + 127) / 128), dim3(128), 0, s, scratch, n_chunks, n_cols, cos_out,
                       sin_out);
    return check_launch("k_finalize_tiles");
}

}  // namespace madi
