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

// ------------------------------------------------------ FP64 graph labels

// classify_mode 5 (the reference's "exact_cached"): labels from the Delaunay
// graph in FP64, equal to the KD classifier's wherever they are given.
//
// From anchor a (its seed image at sa, world coordinates), the margin of edge
// a -> k at q is m = h - (q - sa) . ê, the distance from q to the a-k bisector
// (ê, h in FP64; the edges of a complete list are a's Voronoi facets).  The
// most negative margin below -eps moves the anchor across that facet (a
// nearer seed, the greedy walk to the nearest one).  Otherwise q lies in
// a's Voronoi cell by more than eps, a is its nearest seed, and the smallest
// margin is the distance to the cell's boundary: q is inside the contracted
// cell iff it is at least alpha_a.  eps (1e-9 um, scaled with the box)
// exceeds the FP64 rounding of these margins and of the KD classifier's
// (both below 1e-12 um: coordinates below 1e4 um, seeds at least 2e-3 um
// apart), so any label given here is the one the KD classifier computes;
// a margin within eps of 0 or alpha, a cell whose list may be incomplete
// (cell_flags) or a walk that does not settle returns 0, and the caller runs
// the KD classifier.  The smallest margin also certifies a ball around q
// with q's label: an inside label holds within m1 - alpha_a, an outside one
// within min(m1, alpha_a - m1) (margins move by at most the distance moved).

// The image of cell c's seed nearest to the world point q.
__device__ __forceinline__ void seed_image(const GeomDev& g, int c, double qx, double qy, double qz, double& sx,
                                           double& sy, double& sz)
{
    sx = g.cell_xyz[3 * c];
    sy = g.cell_xyz[3 * c + 1];
    sz = g.cell_xyz[3 * c + 2];
    if (g.periodic) {
        sx += g.L * rint((qx - sx) / g.L);
        sy += g.L * rint((qy - sy) / g.L);
        sz += g.L * rint((qz - sz) / g.L);
    }
}

__device__ __forceinline__ int graph_label_f64(const GeomDev& g, int& a, double& sx, double& sy, double& sz, double qx,
                                               double qy, double qz, double eps, int& inside, double& m1,
                                               long long& moves)
{
    for (int it = 0; it <= kMaxGraphIters; ++it) {
        if ((unsigned)a >= (unsigned)g.n_cells || g.cell_flags[a]) return 0;
        const int lo = g.nb_off[a], hi = g.nb_off[a + 1];
        const double rx = qx - sx, ry = qy - sy, rz = qz - sz;
        double mm = __builtin_inf();
        int km = -1;
        for (int k = lo; k < hi; ++k) {
            const double4 u = g.nb_u64[k];
            const double m = u.w - (rx * u.x + ry * u.y + rz * u.z);
            if (m < mm) {
                mm = m;
                km = k;
            }
        }
        if (km >= 0 && mm < -eps) {
            // a nearer seed across the facet of edge km
            a = g.nb_id[km];
            if ((unsigned)a >= (unsigned)g.n_cells) return 0;
            seed_image(g, a, sx + g.nb_e64[3 * km], sy + g.nb_e64[3 * km + 1], sz + g.nb_e64[3 * km + 2], sx, sy,
                       sz);
            ++moves;
            continue;
        }
        const double al = g.cell_alpha64[a];
        if (!(mm > eps) || fabs(mm - al) <= eps) return 0;
        inside = mm > al;
        m1 = mm;
        return 1;
    }
    return 0;
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
                                                    unsigned long long* counters, MultiPP mp)
{
    int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // several walks that differ only in p_p (mp.n = K > 0, kY), walk-major:
    // thread t runs walker t / K in walk t % K with that walk's p_p and its
    // slices of the outputs and counters.  The walks share the walk seed, so
    // the K lanes of one walker follow one path until an acceptance draw
    // falls between their p_p thresholds, and stay near each other after:
    // while they agree their geometry reads are one coalesced access, and
    // after that they still share cache lines.
    const bool multi = MODE == kY && mp.n > 0;
    if (multi) {
        const int64_t t = w;
        w = t / mp.n;
        const int kw = (int)(t - w * mp.n);
        const size_t n = (size_t)p.n_walkers;
        p.pp = mp.pp[kw];
        Y += kw * n * (size_t)p.n_grid * 3;
        itr += kw * n * (size_t)p.n_grid;
        esc_out += kw * n;
        err_out += kw * n;
        if (counters) counters += kw * MADI_N_COUNTERS;
    }
    __shared__ KdNodeF sf[kLdsNodes];
    __shared__ KdNodeI si[kLdsNodes];
    __shared__ double lds_cols[kWaves * kColChunk * 2];
    const KdView v = kd_stage(g, sf, si, kLdsNodes);

    const bool live = w < p.n_walkers;
    const int G = p.n_grid;
    long long nodes = 0, crossings = 0, rejects = 0, classified = 0, cache_hits = 0, moves = 0, kd_calls = 0;
    // graph labels (classify_mode 5): the walker's anchor (the cell of its
    // position and that seed's image, asx..), and the certified ball of the
    // last graph label: every point within crho of (ccx, ccy, ccz) has label
    // (ccell, cin) and anchor image (csx, ..)
    const bool graph = p.graph_labels != 0 && !g.free_water;
    const double eps = 1e-9 * fmax(1.0, g.L / 100.0);
    double ccx = 0.0, ccy = 0.0, ccz = 0.0, crho = -1.0, csx = 0.0, csy = 0.0, csz = 0.0;
    double asx = 0.0, asy = 0.0, asz = 0.0;
    int ccell = 0, cin = 0;

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
    if (live) {
        ++classified;
        ++kd_calls;
        if (!classify_point_f64(g, v, px, py, pz, cell, in, nodes)) {
            err = 1;
            escaped = 1;
        } else if (graph) {
            seed_image(g, cell, px, py, pz, asx, asy, asz);
        }
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
                int ncell = 0, nin = 0, ok = 1;
                double nsx = asx, nsy = asy, nsz = asz;  // the proposal's anchor image
                const double ex = nx - ccx, ey = ny - ccy, ez = nz - ccz;
                if (graph && crho > 0.0 && ex * ex + ey * ey + ez * ez < crho * crho) {
                    ncell = ccell;
                    nin = cin;
                    nsx = csx;
                    nsy = csy;
                    nsz = csz;
                    ++cache_hits;
                } else if (graph) {
                    ++classified;
                    int na = cell;  // from the walker's own cell
                    double m1 = 0.0;
                    if (graph_label_f64(g, na, nsx, nsy, nsz, nx, ny, nz, eps, nin, m1, moves)) {
                        ncell = na;
                        const double al = g.cell_alpha64[na];
                        crho = (nin ? m1 - al : fmin(m1, al - m1)) - 2.0 * eps;
                    } else {
                        ++kd_calls;
                        ok = classify_point_f64(g, v, nx, ny, nz, ncell, nin, nodes);
                        if (ok) seed_image(g, ncell, nx, ny, nz, nsx, nsy, nsz);
                        crho = -1.0;
                    }
                    ccx = nx;
                    ccy = ny;
                    ccz = nz;
                    ccell = ncell;
                    cin = nin;
                    csx = nsx;
                    csy = nsy;
                    csz = nsz;
                } else {
                    ++classified;
                    ++kd_calls;
                    ok = classify_point_f64(g, v, nx, ny, nz, ncell, nin, nodes);
                }
                if (!ok) {
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
                            nsx = asx;
                            nsy = asy;
                            nsz = asz;
                            ++rejects;
                        }
                    }
                    asx = nsx;
                    asy = nsy;
                    asz = nsz;
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
    if (counters && multi) {
        // the lanes of a wave belong to different walks: one add per lane
        if (live) {
            atomicAdd(&counters[MADI_CNT_KD_NODES], (unsigned long long)nodes);
            atomicAdd(&counters[MADI_CNT_CROSSINGS], (unsigned long long)crossings);
            atomicAdd(&counters[MADI_CNT_REJECTS], (unsigned long long)rejects);
            atomicAdd(&counters[MADI_CNT_CLASSIFY], (unsigned long long)classified);
            atomicAdd(&counters[MADI_CNT_SAFE_HIT], (unsigned long long)cache_hits);
            atomicAdd(&counters[MADI_CNT_GRAPH_ITERS], (unsigned long long)moves);
            atomicAdd(&counters[MADI_CNT_FP64_RECHECK], (unsigned long long)kd_calls);
        }
    } else if (counters) {
        const double a = wave_sum_f64((double)nodes);
        const double b = wave_sum_f64((double)crossings);
        const double c = wave_sum_f64((double)rejects);
        const double d = wave_sum_f64((double)classified);
        const double e = wave_sum_f64((double)cache_hits);
        const double f = wave_sum_f64((double)moves);
        const double h = wave_sum_f64((double)kd_calls);
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&counters[MADI_CNT_KD_NODES], (unsigned long long)a);
            atomicAdd(&counters[MADI_CNT_CROSSINGS], (unsigned long long)b);
            atomicAdd(&counters[MADI_CNT_REJECTS], (unsigned long long)c);
            // every live walker's start and unescaped step: a classification
            // (graph or KD) or, with graph labels, a label from the certified
            // ball; KD classifications (all of them without graph labels:
            // the starts and the graph's deferrals with them) count as FP64
            // re-decisions
            atomicAdd(&counters[MADI_CNT_CLASSIFY], (unsigned long long)d);
            atomicAdd(&counters[MADI_CNT_SAFE_HIT], (unsigned long long)e);
            atomicAdd(&counters[MADI_CNT_GRAPH_ITERS], (unsigned long long)f);
            atomicAdd(&counters[MADI_CNT_FP64_RECHECK], (unsigned long long)h);
        }
    }
}

// -------------------------------------------------------------- reduce

// sin and cos of x in FP64 for |x| <= kSinCosFast.  Cody and Waite's
// additive reduction: n = rint(x * 2/pi) and r = x - n*pi/2 with pi/2 split
// into three doubles (pi/2 to 160 bits, each part applied with one FMA), so
// |r| <= pi/4 to within a few units of its last place; then the Taylor series
// of sin and cos on r through r^17 and r^16 (the first omitted terms are below
// 1e-17 there), Horner in r^2, and the quadrant n mod 4.  No branch: callers
// route larger or non-finite arguments to the device library's sincos.
constexpr double kSinCosFast = 0x1p20;

// The constants of sincos_cw.  Kernels take them by value, so they sit in
// scalar registers and each Horner step is one three-input FMA (an FP64
// literal cannot be an operand of one: a constant written in the code costs
// a copy into the destination register at every step).
struct SinCosCoef {
    double two_over_pi;
    double p1, p2, p3;  // pi/2 rounded to a double, then the rounded remainders
    double s[8];        // sin r = r + r^3 * (s[0] + z s[1] + ... + z^7 s[7]), s[k] = (-1)^(k+1) / (2k+3)!
    double c[8];        // cos r = 1 + z * (c[0] + z c[1] + ... + z^7 c[7]), c[k] = (-1)^(k+1) / (2k+2)!
};
constexpr SinCosCoef kSinCos = {
    0x1.45f306dc9c883p-1,
    0x1.921fb54442d18p+0, 0x1.1a62633145c07p-54, -0x1.f1976b7ed8fbcp-110,
    {-1.0 / 6.0, 1.0 / 120.0, -1.0 / 5040.0, 1.0 / 362880.0, -1.0 / 39916800.0, 1.0 / 6227020800.0,
     -1.0 / 1307674368000.0, 1.0 / 355687428096000.0},
    {-1.0 / 2.0, 1.0 / 24.0, -1.0 / 720.0, 1.0 / 40320.0, -1.0 / 3628800.0, 1.0 / 479001600.0,
     -1.0 / 87178291200.0, 1.0 / 20922789888000.0},
};

__device__ __forceinline__ void sincos_cw(double x, const SinCosCoef& k, double& s, double& c)
{
    const double n = __builtin_rint(x * k.two_over_pi);
    double r = __builtin_fma(-n, k.p1, x);
    r = __builtin_fma(-n, k.p2, r);
    r = __builtin_fma(-n, k.p3, r);
    const double z = r * r;
    double ps = k.s[7], pq = k.c[7];
#pragma unroll
    for (int j = 6; j >= 0; --j) {
        ps = __builtin_fma(ps, z, k.s[j]);
        pq = __builtin_fma(pq, z, k.c[j]);
    }
    const double sr = __builtin_fma(r * z, ps, r);
    const double cr = __builtin_fma(z, pq, 1.0);
    // x = n*pi/2 + r: quadrant q gives (sin, cos) = (sr, cr), (cr, -sr), (-sr, -cr), (-cr, sr)
    const int q = (int)n & 3;
    const double a = (q & 1) ? cr : sr, b = (q & 1) ? sr : cr;
    s = (q & 2) ? -a : a;
    c = ((q + 1) & 2) ? -b : b;
}

// Y (walker, grid, axis) to grid-major Yt (grid, sample), sample i = walker *
// 3 + axis: block (walker tile, grid tile) stages 16 walkers x 128 grid
// points through LDS, reading each walker's segment and writing each grid
// row's segment contiguously.
constexpr int kTrW = 16;
constexpr int kTrG = 128;

__global__ __launch_bounds__(kBlock) void k_reduce_transpose(const double* __restrict__ Y, int64_t n, int32_t G,
                                                            double* __restrict__ Yt)
{
    __shared__ double tile[kTrW][kTrG * 3 + 1];
    const int64_t w0 = (int64_t)blockIdx.x * kTrW;
    const int g0 = blockIdx.y * kTrG;
    const int wn = n - w0 < kTrW ? (int)(n - w0) : kTrW;
    const int gn = G - g0 < kTrG ? G - g0 : kTrG;
    for (int idx = threadIdx.x; idx < wn * gn * 3; idx += kBlock) {
        const int w = idx / (gn * 3), r = idx - w * (gn * 3);
        tile[w][r] = Y[(size_t)(w0 + w) * G * 3 + (size_t)g0 * 3 + r];
    }
    __syncthreads();
    const int64_t total = n * 3;
    for (int idx = threadIdx.x; idx < gn * wn * 3; idx += kBlock) {
        const int g = idx / (wn * 3), r = idx - g * (wn * 3);
        Yt[(size_t)(g0 + g) * total + w0 * 3 + r] = tile[r / 3][g * 3 + r % 3];
    }
}

// Column reduce over Yt: block b sums kRedCols consecutive columns (column
// block b / n_chunks) over the tiles of kRedTile samples of chunk b % n_chunks
// (sample i = walker i / 3, axis i % 3 as in the reference).  Workgroups go
// to the eight XCDs in turn and n_chunks is a multiple of 8, so a chunk stays
// on one XCD for every column block and neighbouring pairs' shared grid rows
// are reused from that XCD's L2.  A thread holds kRedPer samples of the tile
// (i = tile start + thread + k * kBlock) and their dM = Yt[jd] + Yt[jD] -
// Yt[js], reloaded only when a column's grid indices differ from the previous
// column's (the b values of one (delta, Delta) pair share them).  A column
// whose phases all lie within kSinCosFast on the wave (bounded by |pc| times
// the wave's largest |dM|) takes sincos_cw without a branch; any other column
// takes the device library's sincos for every sample.  Per column: the thread's sum in k order, a wave butterfly,
// added into the wave's LDS row; at the end the waves' rows in wave order
// give the (column, chunk) partial.  k_reduce_finish adds partials in chunk
// order, so the result does not depend on scheduling.
constexpr int kRedCols = 32;
constexpr int kRedPer = 16;
constexpr int kRedTile = kBlock * kRedPer;

__global__ __launch_bounds__(kBlock) void k_reduce_f64(const double* __restrict__ Yt, int64_t n,
                                                      const int32_t* __restrict__ jd, const int32_t* __restrict__ jD,
                                                      const int32_t* __restrict__ js, const double* __restrict__ pc,
                                                      int32_t n_cols, int32_t n_chunks, double* __restrict__ part,
                                                      SinCosCoef coef)
{
    __shared__ double red[kWaves][kRedCols][2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int chunk = (int)(blockIdx.x % (unsigned)n_chunks);
    const int c0 = (int)(blockIdx.x / (unsigned)n_chunks) * kRedCols;
    const int nc = n_cols - c0 < kRedCols ? n_cols - c0 : kRedCols;
    const int64_t total = n * 3;
    const int64_t n_tiles = (total + kRedTile - 1) / kRedTile;
    const int64_t per_chunk = (n_tiles + n_chunks - 1) / n_chunks;
    const int64_t t_lo = (int64_t)chunk * per_chunk;
    const int64_t t_hi = t_lo + per_chunk < n_tiles ? t_lo + per_chunk : n_tiles;
    for (int j = threadIdx.x; j < kWaves * kRedCols; j += kBlock) {
        red[j / kRedCols][j % kRedCols][0] = 0.0;
        red[j / kRedCols][j % kRedCols][1] = 0.0;
    }
    __syncthreads();
    double dm[kRedPer];
    for (int64_t tile = t_lo; tile < t_hi; ++tile) {
        const int64_t i0 = tile * kRedTile + threadIdx.x;
        const int64_t left = total - i0;
        const int nv = left <= 0 ? 0 : (left >= (int64_t)kRedTile ? kRedPer : (int)((left + kBlock - 1) / kBlock));
        const bool full = (tile + 1) * kRedTile <= total;  // every thread holds kRedPer samples
        int pa = -1, pb = -1, ps = -1;
        double dmax = 0.0;  // the wave's largest |dM| (NaN if any is NaN)
        for (int j = 0; j < nc; ++j) {
            const int col = c0 + j;
            const int a = jd[col], b = jD[col], c = js[col];
            if (a != pa || b != pb || c != ps) {
                pa = a;
                pb = b;
                ps = c;
                const double* Ya = Yt + (size_t)a * total + i0;
                const double* Yb = Yt + (size_t)b * total + i0;
                const double* Yc = Yt + (size_t)c * total + i0;
#pragma unroll
                for (int k = 0; k < kRedPer; ++k) {
                    dm[k] = 0.0;
                    if (k < nv) dm[k] = Ya[k * kBlock] + Yb[k * kBlock] - Yc[k * kBlock];
                }
                double m = 0.0;
#pragma unroll
                for (int k = 0; k < kRedPer; ++k) {
                    const double v = __builtin_fabs(dm[k]);
                    m = (v > m || v != v) ? v : m;
                }
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) {
                    const double o = __shfl_xor(m, off, 64);
                    m = (o > m || o != o) ? o : m;
                }
                dmax = m;
            }
            const double kc = pc[col];
            double cs = 0.0, sn = 0.0;
            const bool fast = __builtin_fabs(kc) * dmax <= kSinCosFast;
            if (fast && full) {
#pragma unroll
                for (int k = 0; k < kRedPer; ++k) {
                    double sv, cv;
                    sincos_cw(kc * dm[k], coef, sv, cv);
                    cs += cv;
                    sn += sv;
                }
            } else if (fast) {
#pragma unroll
                for (int k = 0; k < kRedPer; ++k) {
                    double sv, cv;
                    sincos_cw(kc * dm[k], coef, sv, cv);
                    if (k < nv) {
                        cs += cv;
                        sn += sv;
                    }
                }
            } else {
                for (int k = 0; k < nv; ++k) {
                    double sv, cv;
                    sincos(kc * dm[k], &sv, &cv);
                    cs += cv;
                    sn += sv;
                }
            }
            cs = wave_sum_f64(cs);
            sn = wave_sum_f64(sn);
            if (lane == 0) {
                red[wave][j][0] += cs;
                red[wave][j][1] += sn;
            }
        }
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t < nc) {
        double a = 0.0, b = 0.0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            a += red[w][t][0];
            b += red[w][t][1];
        }
        part[((size_t)(c0 + t) * n_chunks + chunk) * 2 + 0] = a;
        part[((size_t)(c0 + t) * n_chunks + chunk) * 2 + 1] = b;
    }
}

int32_t reduce_chunks(int64_t n, int32_t n_cols)
{
    const int64_t n_tiles = (n * 3 + kRedTile - 1) / kRedTile;
    const int64_t col_blocks = ((int64_t)n_cols + kRedCols - 1) / kRedCols;
    // enough blocks to fill the chip several times over, and partials of at
    // most a few tens of MB
    int64_t want = (16384 + col_blocks - 1) / col_blocks;
    if (want < 64) want = 64;
    int64_t chunks = n_tiles < want ? n_tiles : want;
    // a multiple of the XCD count (see k_reduce_f64); chunks past the last
    // tile add zeros
    chunks = (chunks + 7) / 8 * 8;
    if (chunks < 8) chunks = 8;
    return (int32_t)chunks;
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
    MultiPP mp{};
    hipLaunchKernelGGL(k_walk_f64<kY>, dim3((unsigned)blocks), dim3(kBlock), 0, s, g, p, jump_pow, gi, gd, gu, Y,
                       itr, esc, err, fo, counters, mp);
    return check_launch("k_walk_f64");
}

int launch_walk_f64_multi(const GeomDev& g, const WalkDev& p, const Gf2Col* jump_pow, const MultiPP& mp, double* Y,
                          int8_t* itr, int8_t* esc, int8_t* err, unsigned long long* counters, hipStream_t s)
{
    if (p.n_walkers == 0 || mp.n == 0) return MADI_OK;
    // one thread per (walker, walk), walk-major
    const int64_t blocks = (p.n_walkers * mp.n + kBlock - 1) / kBlock;
    if (blocks > 0x7fffffff) {
        set_error("k_walk_f64 (several walks): %lld walkers x %d walks exceed one grid", (long long)p.n_walkers, mp.n);
        return MADI_ERR_ARG;
    }
    FusedOut fo{};
    hipLaunchKernelGGL(k_walk_f64<kY>, dim3((unsigned)blocks), dim3(kBlock), 0, s, g, p, jump_pow, nullptr, nullptr,
                       nullptr, Y, itr, esc, err, fo, counters, mp);
    return check_launch("k_walk_f64 (several walks)");
}

int launch_walk_f64_fused(const GeomDev& g, const WalkDev& p, const Gf2Col* jump_pow, const FusedOut& fo,
                          unsigned long long* counters, hipStream_t s)
{
    if (p.n_walkers == 0) return MADI_OK;
    const int64_t blocks = (p.n_walkers + kBlock - 1) / kBlock;
    MultiPP mp{};
    hipLaunchKernelGGL(k_walk_f64<kFused>, dim3((unsigned)blocks), dim3(kBlock), 0, s, g, p, jump_pow, nullptr,
                       nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, fo, counters, mp);
    return check_launch("k_walk_f64_fused");
}

int launch_reduce_f64(const double* Y, int64_t n, int32_t G, const int32_t* jd, const int32_t* jD, const int32_t* js,
                      const double* pc, int32_t n_cols, double* Yt, double* part, int32_t n_chunks, double* cos_out,
                      double* sin_out, hipStream_t s)
{
    if (n_cols == 0) return MADI_OK;
    if (n > 0) {
        hipLaunchKernelGGL(k_reduce_transpose, dim3((unsigned)((n + kTrW - 1) / kTrW), (G + kTrG - 1) / kTrG),
                           dim3(kBlock), 0, s, Y, n, G, Yt);
        const int rc = check_launch("k_reduce_transpose");
        if (rc) return rc;
    }
    const int64_t blocks = (int64_t)n_chunks * ((n_cols + kRedCols - 1) / kRedCols);
    hipLaunchKernelGGL(k_reduce_f64, dim3((unsigned)blocks), dim3(kBlock), 0, s, Yt, n, jd, jD, js, pc, n_cols,
                       n_chunks, part, kSinCos);
    int rc = check_launch("k_reduce_f64");
    if (rc) return rc;
    hipLaunchKernelGGL(k_reduce_finish, dim3((n_cols + 127) / 128), dim3(128), 0, s, part, n_cols, n_chunks, cos_out,
                       sin_out);
    return check_launch("k_reduce_finish");
}

}  // namespace madi
