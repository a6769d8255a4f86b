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

template <int GRP>
__device__ __forceinline__ int graph_label_f64(const GeomDev& g, int& a, double& sx, double& sy, double& sz, double qx,
                                               double qy, double qz, double eps, int& inside, double& m1,
                                               long long& moves)
{
    for (int it = 0; it <= kMaxGraphIters; ++it) {
        if ((unsigned)a >= (unsigned)g.n_cells) return 0;
        // the cell's flag, list and annulus in one round trip
        const int8_t flag = g.cell_flags[a];
        const int lo = g.nb_off[a], hi = g.nb_off[a + 1];
        const double al = g.cell_alpha64[a];
        if (flag) return 0;
        const double rx = qx - sx, ry = qy - sy, rz = qz - sz;
        double mm = __builtin_inf();
        int km = -1;
        // edges in groups of GRP, a group's records fetched before any is
        // used (one round trip per group; the tail re-reads edge lo), then
        // compared in list order: the first smallest margin, as one by one
        for (int k0 = lo; k0 < hi; k0 += GRP) {
            double4 u[GRP];
#pragma unroll
            for (int j = 0; j < GRP; ++j) u[j] = g.nb_u64[k0 + j < hi ? k0 + j : lo];
#pragma unroll
            for (int j = 0; j < GRP; ++j) {
                const double m = u[j].w - (rx * u[j].x + ry * u[j].y + rz * u[j].z);
                if (k0 + j < hi && m < mm) {
                    mm = m;
                    km = k0 + j;
                }
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
        if (!(mm > eps) || fabs(mm - al) <= eps) return 0;
        inside = mm > al;
        m1 = mm;
        return 1;
    }
    return 0;
}

// Edges per memory round trip of a graph scan: the one-walk kernel's waves
// run about one per SIMD and gain from the deeper fetch; the k_io sweep fills
// the GPU, where four edges keep it at three waves per SIMD instead of two.
constexpr int kGraphGroupWalk = 8;
constexpr int kGraphGroupSweep = 4;

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

// One walker of the reference step loop (madi/walker_gpu.py:531-656): its
// position, its anchor (the cell of its position and that seed's image,
// asx..; graph labels), the certified ball of the last graph label (every
// point within crho of (ccx, ccy, ccz) has label (ccell, cin) and anchor
// image (csx, ..)), Y, and its counters.
struct WalkerF64 {
    double px, py, pz, asx, asy, asz;
    double ccx, ccy, ccz, crho, csx, csy, csz;
    double Yx, Yy, Yz;
    int cell, in, ccell, cin, err, escaped;
    long long nodes, crossings, rejects, classified, cache_hits, moves, kd_calls;
};

__device__ __forceinline__ void walker_clear(WalkerF64& s)
{
    s.px = s.py = s.pz = s.asx = s.asy = s.asz = 0.0;
    s.ccx = s.ccy = s.ccz = s.csx = s.csy = s.csz = 0.0;
    s.crho = -1.0;
    s.Yx = s.Yy = s.Yz = 0.0;
    s.cell = s.in = s.ccell = s.cin = s.err = s.escaped = 0;
    s.nodes = s.crossings = s.rejects = s.classified = s.cache_hits = s.moves = s.kd_calls = 0;
}

// The start at (px, py, pz): the exact classification and the anchor image.
__device__ __forceinline__ void walker_start(const GeomDev& g, const KdView& v, bool graph, WalkerF64& s)
{
    ++s.classified;
    ++s.kd_calls;
    if (!classify_point_f64(g, v, s.px, s.py, s.pz, s.cell, s.in, s.nodes)) {
        s.err = 1;
        s.escaped = 1;
    } else if (graph) {
        seed_image(g, s.cell, s.px, s.py, s.pz, s.asx, s.asy, s.asz);
    }
}

// One step with displacement (dx, dy, dz) and acceptance uniform u: the
// proposal's label (the certified ball, the graph or the KD classifier),
// the membrane count m and the acceptance test u < p_p^m, the move, and the
// trapezoid of Y (its previous point: the position before the step less
// L/2).
template <int GRP>
__device__ __forceinline__ void walker_step(const GeomDev& g, const KdView& v, const WalkDev& p, bool graph,
                                            double eps, double half, double dx, double dy, double dz, double u,
                                            WalkerF64& s)
{
    double nx = s.px + dx, ny = s.py + dy, nz = s.pz + dz;
    if (!g.periodic && (nx < 0.0 || nx >= g.L || ny < 0.0 || ny >= g.L || nz < 0.0 || nz >= g.L)) {
        s.escaped = 1;
        return;
    }
    int ncell = 0, nin = 0, ok = 1;
    double nsx = s.asx, nsy = s.asy, nsz = s.asz;  // the proposal's anchor image
    const double ex = nx - s.ccx, ey = ny - s.ccy, ez = nz - s.ccz;
    if (graph && s.crho > 0.0 && ex * ex + ey * ey + ez * ez < s.crho * s.crho) {
        ncell = s.ccell;
        nin = s.cin;
        nsx = s.csx;
        nsy = s.csy;
        nsz = s.csz;
        ++s.cache_hits;
    } else if (graph) {
        ++s.classified;
        int na = s.cell;  // from the walker's own cell
        double m1 = 0.0;
        if (graph_label_f64<GRP>(g, na, nsx, nsy, nsz, nx, ny, nz, eps, nin, m1, s.moves)) {
            ncell = na;
            const double al = g.cell_alpha64[na];
            s.crho = (nin ? m1 - al : fmin(m1, al - m1)) - 2.0 * eps;
        } else {
            ++s.kd_calls;
            ok = classify_point_f64(g, v, nx, ny, nz, ncell, nin, s.nodes);
            if (ok) seed_image(g, ncell, nx, ny, nz, nsx, nsy, nsz);
            s.crho = -1.0;
        }
        s.ccx = nx;
        s.ccy = ny;
        s.ccz = nz;
        s.ccell = ncell;
        s.cin = nin;
        s.csx = nsx;
        s.csy = nsy;
        s.csz = nsz;
    } else {
        ++s.classified;
        ++s.kd_calls;
        ok = classify_point_f64(g, v, nx, ny, nz, ncell, nin, s.nodes);
    }
    if (!ok) {
        s.err = 1;
        s.escaped = 1;
        return;
    }
    int m = 0;
    if (s.in != nin) m = 1;
    else if (s.in && nin && s.cell != ncell) m = 2;
    if (m > 0) ++s.crossings;
    if (m > 0 && p.pp < 1.0) {
        const double thr = m == 1 ? p.pp : p.pp * p.pp;
        if (u >= thr) {
            nx = s.px;
            ny = s.py;
            nz = s.pz;
            ncell = s.cell;
            nin = s.in;
            nsx = s.asx;
            nsy = s.asy;
            nsz = s.asz;
            ++s.rejects;
        }
    }
    const double xp = s.px - half, yp = s.py - half, zp = s.pz - half;
    s.asx = nsx;
    s.asy = nsy;
    s.asz = nsz;
    s.px = nx;
    s.py = ny;
    s.pz = nz;
    s.cell = ncell;
    s.in = nin;
    const double xn = s.px - half, yn = s.py - half, zn = s.pz - half;
    s.Yx += 0.5 * (xp + xn) * p.ts;
    s.Yy += 0.5 * (yp + yn) * p.ts;
    s.Yz += 0.5 * (zp + zn) * p.ts;
}

// The walker's Y and inside flag at grid point j (mode kY).
__device__ __forceinline__ void walker_sample(const WalkerF64& s, int64_t w, int G, int j, double* __restrict__ Y,
                                              int8_t* __restrict__ itr)
{
    double* Yw = Y + ((size_t)w * G + j) * 3;
    Yw[0] = s.Yx;
    Yw[1] = s.Yy;
    Yw[2] = s.Yz;
    itr[(size_t)w * G + j] = (int8_t)s.in;
}

// One walker per thread.  Mode kY writes Y / inside samples; mode kFused
// accumulates the plan's moments instead and reduces phases at the end.
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
    const bool graph = p.graph_labels != 0 && !g.free_water;
    const double eps = 1e-9 * fmax(1.0, g.L / 100.0);
    const double half = 0.5 * g.L;

    WalkerF64 s;
    walker_clear(s);
    Rng r{0, 0};
    if (live) {
        if (p.use_golden) {
            s.px = gi[3 * w];
            s.py = gi[3 * w + 1];
            s.pz = gi[3 * w + 2];
        } else {
            r = rng_for_walker(p.seed, p.first_walker + (uint64_t)w, jump_pow);
            const double span = g.source_hi - g.source_lo;
            s.px = g.source_lo + uniform_f64(r) * span;
            s.py = g.source_lo + uniform_f64(r) * span;
            s.pz = g.source_lo + uniform_f64(r) * span;
        }
        walker_start(g, v, graph, s);
    }
    double dM[MADI_MAX_PAIRS][3];
#pragma unroll
    for (int q = 0; q < MADI_MAX_PAIRS; ++q) dM[q][0] = dM[q][1] = dM[q][2] = 0.0;
    if (MODE == kY && live) {
        double* Yw = Y + (size_t)w * G * 3;
        Yw[0] = Yw[1] = Yw[2] = 0.0;
        itr[(size_t)w * G] = (int8_t)(s.err ? 0 : s.in);
    }
    if (MODE == kFused) {
        const long long c0 = wave_popc(live && !s.err && s.in);
        if ((threadIdx.x & 63) == 0 && c0) atomicAdd((unsigned long long*)&fo.occupancy[0], (unsigned long long)c0);
    }
    int next_ev = 0;
    for (int32_t step = 0; step < p.n_steps; ++step) {
        if (live && !s.escaped) {
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
            walker_step<kGraphGroupWalk>(g, v, p, graph, eps, half, dx, dy, dz, u, s);
        }
        const int32_t s1 = step + 1;
        if (s1 % p.steps_per_h == 0) {
            const int j = s1 / p.steps_per_h;
            if (MODE == kY) {
                if (live) walker_sample(s, w, G, j, Y, itr);
            } else {
                const long long c = wave_popc(live && !s.err && s.in);
                if ((threadIdx.x & 63) == 0 && c) atomicAdd((unsigned long long*)&fo.occupancy[j], (unsigned long long)c);
                while (next_ev < fo.ev.n && fo.ev.j[next_ev] == j) {
                    const int q = fo.ev.pair[next_ev];
                    const double sg = (double)fo.ev.sign[next_ev];
#pragma unroll
                    for (int t = 0; t < MADI_MAX_PAIRS; ++t)
                        if (t == q) {
                            dM[t][0] += sg * s.Yx;
                            dM[t][1] += sg * s.Yy;
                            dM[t][2] += sg * s.Yz;
                        }
                    ++next_ev;
                }
            }
        }
    }
    if (MODE == kY) {
        if (live) {
            esc_out[w] = (int8_t)s.escaped;
            err_out[w] = (int8_t)s.err;
        }
    } else {
        const long long ne = wave_popc(live && s.escaped);
        const long long nr = wave_popc(live && s.err);
        if ((threadIdx.x & 63) == 0) {
            if (ne) atomicAdd((unsigned long long*)&fo.flags[MADI_FLAG_ESCAPED], (unsigned long long)ne);
            if (nr) atomicAdd((unsigned long long*)&fo.flags[MADI_FLAG_CLASSIFIER_ERR], (unsigned long long)nr);
        }
        double* slot = fo.partial + (size_t)(fo.tile_base + blockIdx.x) * fo.plan->n_cols * 2;
        emit_columns<double>(*fo.plan, dM, live && !s.escaped, lds_cols, slot);
    }
    if (counters) {
        const double a = wave_sum_f64((double)s.nodes);
        const double b = wave_sum_f64((double)s.crossings);
        const double c = wave_sum_f64((double)s.rejects);
        const double d = wave_sum_f64((double)s.classified);
        const double e = wave_sum_f64((double)s.cache_hits);
        const double f = wave_sum_f64((double)s.moves);
        const double h = wave_sum_f64((double)s.kd_calls);
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

// The k_io sweep (madi_walk_f64_multi): K = mp.n walks of one geometry that
// differ only in p_p, walk-major.  A block holds M = kBlock / K walkers as K
// consecutive threads each (thread m K + k runs walker m in walk k; the
// threads past M K idle).  The walks share the walk seed, so the K threads
// of a walker make the same draws at every step.  In each batch of K steps
// every thread advances its copy of the walker's stream and keeps the draws
// of step base + k, turns them into that step's displacement and acceptance
// uniform (the FP64 logarithm, square root and cosine: once per step and
// walker instead of once per walk) in the step table in LDS, and after a
// barrier the walker's threads take each step's values from the table.  A
// value is computed by the same code as in k_walk_f64 on the same draws, so
// every walk's results are bitwise its k_walk_f64 walk's.  The threads of a
// walker also follow one path until an acceptance draw falls between their
// p_p thresholds, and stay near each other after: their geometry reads are
// one coalesced access, or share cache lines.
constexpr int kSweepLdsNodes = 255;  // top 8 KD levels: the step table shares the LDS

__global__ __launch_bounds__(kBlock) void k_walk_f64_sweep(GeomDev g, WalkDev p, const Gf2Col* __restrict__ jump_pow,
                                                          MultiPP mp, double* __restrict__ Y,
                                                          int8_t* __restrict__ itr, int8_t* __restrict__ esc_out,
                                                          int8_t* __restrict__ err_out, unsigned long long* counters)
{
    __shared__ KdNodeF sf[kSweepLdsNodes];
    __shared__ KdNodeI si[kSweepLdsNodes];
    __shared__ double4 table[kBlock];  // walker m's step base + t at m K + t: (dx, dy, dz, u)
    const KdView v = kd_stage(g, sf, si, kSweepLdsNodes);

    const int K = mp.n;
    const int M = kBlock / K;
    const int m = (int)threadIdx.x / K;
    const int k = (int)threadIdx.x - m * K;
    const int64_t w = (int64_t)blockIdx.x * M + m;
    const bool slot = m < M;  // a (walker, walk) thread; its walker may lie past the last
    const bool live = slot && w < p.n_walkers;
    const int G = p.n_grid;
    // the thread's row of the outputs: walk k's slice holds walker w at row k n + w
    const int64_t row = (int64_t)k * p.n_walkers + w;
    if (slot) p.pp = mp.pp[k];
    const bool graph = p.graph_labels != 0 && !g.free_water;
    const double eps = 1e-9 * fmax(1.0, g.L / 100.0);
    const double half = 0.5 * g.L;

    WalkerF64 s;
    walker_clear(s);
    Rng r{0, 0};
    if (live) {
        r = rng_for_walker(p.seed, p.first_walker + (uint64_t)w, jump_pow);
        const double span = g.source_hi - g.source_lo;
        s.px = g.source_lo + uniform_f64(r) * span;
        s.py = g.source_lo + uniform_f64(r) * span;
        s.pz = g.source_lo + uniform_f64(r) * span;
        walker_start(g, v, graph, s);
        double* Yw = Y + (size_t)row * G * 3;
        Yw[0] = Yw[1] = Yw[2] = 0.0;
        itr[(size_t)row * G] = (int8_t)(s.err ? 0 : s.in);
    }
    for (int32_t base = 0; base < p.n_steps; base += K) {
        const int B = p.n_steps - base < K ? p.n_steps - base : K;
        // the draws of step base + k (7 per step: three normals from two
        // each, then the acceptance uniform)
        uint64_t d0 = 0, d1 = 0, d2 = 0, d3 = 0, d4 = 0, d5 = 0, d6 = 0;
        for (int t = 0; t < B; ++t) {
            const uint64_t a0 = rng_next(r), a1 = rng_next(r), a2 = rng_next(r), a3 = rng_next(r);
            const uint64_t a4 = rng_next(r), a5 = rng_next(r), a6 = rng_next(r);
            if (t == k) {
                d0 = a0;
                d1 = a1;
                d2 = a2;
                d3 = a3;
                d4 = a4;
                d5 = a5;
                d6 = a6;
            }
        }
        __syncthreads();  // every read of the previous batch's table is done
        if (live && k < B)
            table[threadIdx.x] = make_double4(normal_f64_of(d0, d1) * p.sigma, normal_f64_of(d2, d3) * p.sigma,
                                              normal_f64_of(d4, d5) * p.sigma, u53_to_f64(d6));
        __syncthreads();
        for (int t = 0; t < B; ++t) {
            if (live && !s.escaped) {
                const double4 d = table[m * K + t];
                walker_step<kGraphGroupSweep>(g, v, p, graph, eps, half, d.x, d.y, d.z, d.w, s);
            }
            const int32_t s1 = base + t + 1;
            if (s1 % p.steps_per_h == 0 && live) walker_sample(s, row, G, s1 / p.steps_per_h, Y, itr);
        }
    }
    if (live) {
        esc_out[row] = (int8_t)s.escaped;
        err_out[row] = (int8_t)s.err;
        if (counters) {
            // the threads of a wave belong to different walks: one add each
            unsigned long long* c = counters + k * MADI_N_COUNTERS;
            atomicAdd(&c[MADI_CNT_KD_NODES], (unsigned long long)s.nodes);
            atomicAdd(&c[MADI_CNT_CROSSINGS], (unsigned long long)s.crossings);
            atomicAdd(&c[MADI_CNT_REJECTS], (unsigned long long)s.rejects);
            atomicAdd(&c[MADI_CNT_CLASSIFY], (unsigned long long)s.classified);
            atomicAdd(&c[MADI_CNT_SAFE_HIT], (unsigned long long)s.cache_hits);
            atomicAdd(&c[MADI_CNT_GRAPH_ITERS], (unsigned long long)s.moves);
            atomicAdd(&c[MADI_CNT_FP64_RECHECK], (unsigned long long)s.kd_calls);
        }
    }
}

// Occupancy of stacked walks: counts[k][j] = sum over walkers w of
// inside[k][w][j] (the reference's np.sum(inside, axis=0),
// madi/walker_gpu.py:1054).  Block (walk k, 64 grid points) with 4 rows of
// threads: consecutive threads read consecutive bytes of a walker's row,
// each row of threads strides over the walkers, then the rows are summed in
// LDS.  Integer sums, so the order does not matter.
constexpr int kCountCols = 64;
constexpr int kCountRows = kBlock / kCountCols;

__global__ __launch_bounds__(kBlock) void k_count_inside(const int8_t* __restrict__ itr, int64_t n, int32_t G,
                                                        long long* __restrict__ counts)
{
    __shared__ long long part[kCountRows][kCountCols];
    const int tx = threadIdx.x % kCountCols, ty = threadIdx.x / kCountCols;
    const int64_t k = blockIdx.y;
    const int j = blockIdx.x * kCountCols + tx;
    long long c = 0;
    if (j < G) {
        const int8_t* col = itr + k * n * G + j;
        for (int64_t w = ty; w < n; w += kCountRows) c += col[w * G];
    }
    part[ty][tx] = c;
    __syncthreads();
    if (ty == 0 && j < G) {
        long long t = 0;
#pragma unroll
        for (int r = 0; r < kCountRows; ++r) t += part[r][tx];
        counts[k * G + j] = t;
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
    hipLaunchKernelGGL(k_walk_f64<kY>, dim3((unsigned)blocks), dim3(kBlock), 0, s, g, p, jump_pow, gi, gd, gu, Y,
                       itr, esc, err, fo, counters);
    return check_launch("k_walk_f64");
}

int launch_walk_f64_multi(const GeomDev& g, const WalkDev& p, const Gf2Col* jump_pow, const MultiPP& mp, double* Y,
                          int8_t* itr, int8_t* esc, int8_t* err, unsigned long long* counters, hipStream_t s)
{
    if (p.n_walkers == 0 || mp.n == 0) return MADI_OK;
    if (mp.n < 1 || mp.n > MADI_MAX_MULTI_WALKS) {
        set_error("k_walk_f64_sweep: 1 to %d walks per launch, got %d", MADI_MAX_MULTI_WALKS, mp.n);
        return MADI_ERR_ARG;
    }
    const int64_t per_block = kBlock / mp.n;  // walkers per block
    const int64_t blocks = (p.n_walkers + per_block - 1) / per_block;
    if (blocks > 0x7fffffff) {
        set_error("k_walk_f64_sweep: %lld walkers exceed one grid", (long long)p.n_walkers);
        return MADI_ERR_ARG;
    }
    hipLaunchKernelGGL(k_walk_f64_sweep, dim3((unsigned)blocks), dim3(kBlock), 0, s, g, p, jump_pow, mp, Y, itr, esc,
                       err, counters);
    return check_launch("k_walk_f64_sweep");
}

int launch_count_inside(const int8_t* itr, int64_t n_walks, int64_t n, int32_t G, long long* counts, hipStream_t s)
{
    if (n_walks == 0 || G == 0) return MADI_OK;
    hipLaunchKernelGGL(k_count_inside, dim3((unsigned)((G + kCountCols - 1) / kCountCols), (unsigned)n_walks),
                       dim3(kBlock), 0, s, itr, n, G, counts);
    return check_launch("k_count_inside");
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
