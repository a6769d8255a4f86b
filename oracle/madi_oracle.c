/*
 * CPU ORACLE for the MADI random walk -- TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain FP64 C, the reference algorithm of the walk
 * path so that tests can check the HIP kernels against it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  Nothing in
 * the shipped package links or calls it.
 *
 * Reference functions restated here (paths relative to the reference tree):
 *   og_next / og_jump / og_seed_state   numba-src:45-125 (xoroshiro128+,
 *                                       splitmix64 seeding, 2^64 jump)
 *   og_uniform_f64 / og_normal_f64      numba-src:128-221
 *   og_rng_states                       numba-src:224-240 (walker i = jump^i)
 *   og_nearest / og_classify            madi/walker_gpu.py:222-335
 *   og_classify_brute                   tests/physics_audit/test_geometry_diagnostics.py:119-140
 *   og_walk                             madi/walker_gpu.py:520-656 (+ the CPU
 *                                       semantics of madi/walker_gpu.py:698-840
 *                                       for samples after an escape)
 *   og_reduce                           madi/walker_gpu.py:954-968
 *   og_map_match                        madi/fitters_gpu.py:96-154 (MAP
 *                                       matchers, fixed and free S0)
 *
 * Arithmetic follows the reference GPU formula order (dx*dx+dy*dy+dz*dz, no
 * fused multiply-add except the explicit fma() of the matchers).  Build with
 * -ffp-contract=off (see oracle/Makefile).
 *
 * Parity pins (tests/test_oracle_cpu.py, tests/test_match_cpu.py):
 * brute-force all-seed S2 labels, the trapezoid Y closed form (bitwise),
 * free-water Stejskal-Tanner, an independent GF(2) derivation of the jump,
 * the numpy restatement of the reference CPU walk on injected streams, and
 * the reference's own CPU walk and reduce on a seeded Eq. S2 population
 * (tests/golden/walk_vectors.npz: og_walk's Y equals it bit for bit,
 * tests/test_walk_golden_cpu.py); for the matchers the reference's
 * known-answer test, exact-copy and scaled-copy libraries with zero
 * residual, and the numpy form of the reference's CPU matcher path.
 *
 * Parity unpinned: og_normal_f64's float32 uniforms follow numba 0.54.1's
 * text; whether numba >= 0.57 (the reference's pin) still draws them that way
 * cannot be checked offline, and no reference test pins the numba stream.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Thread count of the OpenMP loops below (benchmark control). */
void og_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }

/* ---------------------------------------------------------------- RNG */

typedef struct { uint64_t s0, s1; } og_state;

static inline uint64_t og_rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

/* splitmix64 single output; both words take it (numba-src:45-68). */
og_state og_seed_state(uint64_t seed)
{
    uint64_t z = seed + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    og_state s = {z, z};
    return s;
}

/* xoroshiro128+ step with rotations 55/14/36 (numba-src:79-98). */
uint64_t og_next(og_state* st)
{
    uint64_t a = st->s0, b = st->s1;
    uint64_t out = a + b;
    b ^= a;
    st->s0 = og_rotl(a, 55) ^ b ^ (b << 14);
    st->s1 = og_rotl(b, 36);
    return out;
}

/* Advance by 2^64 steps with the published jump polynomial (numba-src:101-125). */
void og_jump(og_state* st)
{
    static const uint64_t J[2] = {0xbeac0467eba5facbull, 0xd86b048b86aa9922ull};
    uint64_t a = 0, b = 0;
    for (int w = 0; w < 2; ++w)
        for (int bit = 0; bit < 64; ++bit) {
            if (J[w] & (1ull << bit)) { a ^= st->s0; b ^= st->s1; }
            og_next(st);
        }
    st->s0 = a;
    st->s1 = b;
}

/* states[2*i], states[2*i+1] for walker i = jump^(first+i)(seed) (numba-src:224-240). */
void og_rng_states(uint64_t seed, uint64_t first, int64_t n, uint64_t* states)
{
    og_state s = og_seed_state(seed);
    for (uint64_t k = 0; k < first; ++k) og_jump(&s);
    for (int64_t i = 0; i < n; ++i) {
        if (i > 0) og_jump(&s);
        states[2 * i] = s.s0;
        states[2 * i + 1] = s.s1;
    }
}

static inline double og_u53(uint64_t x) { return (double)(x >> 11) * 0x1.0p-53; }

double og_uniform_f64(og_state* st) { return og_u53(og_next(st)); }

float og_uniform_f32(og_state* st) { return (float)og_u53(og_next(st)); }

/* Box-Muller from two float32 uniforms, first output only (numba-src:199-221).
 * The float32 logarithm is taken as the correctly rounded value
 * (float)log((double)u1), so CPU and GPU builds agree bit for bit. */
double og_normal_f64(og_state* st)
{
    float u1 = og_uniform_f32(st);
    float u2 = og_uniform_f32(st);
    float lg = (float)log((double)u1);
    return sqrt(-2.0 * (double)lg) * cos(6.283185307179586 * (double)u2);
}

/* Fill n draws of each kind (test helper). */
void og_draw(uint64_t s0, uint64_t s1, int kind, int64_t n, double* out)
{
    og_state st = {s0, s1};
    for (int64_t i = 0; i < n; ++i) {
        if (kind == 0) out[i] = og_uniform_f64(&st);
        else if (kind == 1) out[i] = (double)og_uniform_f32(&st);
        else if (kind == 2) out[i] = og_normal_f64(&st);
        else { uint64_t r = og_next(&st); memcpy(&out[i], &r, 8); }
    }
}

/* ---------------------------------------------------------- geometry */

typedef struct {
    const double* seeds;      /* n x 3 */
    const double* annulus;    /* n */
    const int32_t* node_seed; /* n, pre-order balanced KD tree */
    const int8_t* node_axis;
    const int32_t* node_left;
    const int32_t* node_right;
    const int32_t* node_parent;
    const int32_t* orig_id;   /* n or NULL: label map of a periodic padded set */
    int64_t n;
    double L;
    int32_t periodic;
    int32_t free_water;
} og_geom;

/* Stackless exact nearest seed (madi/walker_gpu.py:222-270). */
static void og_nearest(const og_geom* g, double px, double py, double pz, int32_t* sid_out, double* dsq_out)
{
    const double* S = g->seeds;
    int32_t node = 0, prev = -1, best = -1;
    double best_dsq = INFINITY;
    while (node >= 0) {
        int32_t sid = g->node_seed[node];
        int ax = g->node_axis[node];
        double c = ax == 0 ? px : (ax == 1 ? py : pz);
        double delta = c - S[3 * sid + ax];
        int32_t near = delta <= 0.0 ? g->node_left[node] : g->node_right[node];
        int32_t far = delta <= 0.0 ? g->node_right[node] : g->node_left[node];
        int32_t up = g->node_parent[node];
        if (prev == up) {
            double dx = px - S[3 * sid + 0];
            double dy = py - S[3 * sid + 1];
            double dz = pz - S[3 * sid + 2];
            double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 < best_dsq) { best_dsq = d2; best = sid; }
            prev = node;
            if (near >= 0) node = near;
            else if (far >= 0 && delta * delta < best_dsq) node = far;
            else node = up;
        } else if (prev == near && far >= 0 && delta * delta < best_dsq) {
            prev = node;
            node = far;
        } else {
            prev = node;
            node = up;
        }
    }
    *sid_out = best;
    *dsq_out = best_dsq;
}

/* Exact SI Eq. S2 label with the d1 + 2 alpha1 ball (madi/walker_gpu.py:272-335).
 * Returns ok (1) or 0 on coincident seeds / empty tree. */
static int og_classify_raw(const og_geom* g, double px, double py, double pz, int32_t* cell, int8_t* inside)
{
    const double* S = g->seeds;
    int32_t s1;
    double d1sq;
    og_nearest(g, px, py, pz, &s1, &d1sq);
    if (s1 < 0) { *cell = 0; *inside = 0; return 0; }
    double a1 = g->annulus[s1];
    double radius = sqrt(d1sq) + 2.0 * a1;
    double r2 = radius * radius;
    int32_t node = 0, prev = -1;
    while (node >= 0) {
        int32_t sid = g->node_seed[node];
        int ax = g->node_axis[node];
        double c = ax == 0 ? px : (ax == 1 ? py : pz);
        double delta = c - S[3 * sid + ax];
        int32_t near = delta <= 0.0 ? g->node_left[node] : g->node_right[node];
        int32_t far = delta <= 0.0 ? g->node_right[node] : g->node_left[node];
        int32_t up = g->node_parent[node];
        if (prev == up) {
            double dx = px - S[3 * sid + 0];
            double dy = py - S[3 * sid + 1];
            double dz = pz - S[3 * sid + 2];
            double dj2 = dx * dx + dy * dy + dz * dz;
            if (sid != s1 && dj2 <= r2) {
                double sx = S[3 * sid + 0] - S[3 * s1 + 0];
                double sy = S[3 * sid + 1] - S[3 * s1 + 1];
                double sz = S[3 * sid + 2] - S[3 * s1 + 2];
                double sep = sqrt(sx * sx + sy * sy + sz * sz);
                if (sep <= 0.0) { *cell = 0; *inside = 0; return 0; }
                double margin = (dj2 - d1sq) / (2.0 * sep);
                if (margin < a1) { *cell = s1; *inside = 0; return 1; }
            }
            prev = node;
            if (near >= 0) node = near;
            else if (far >= 0 && delta * delta <= r2) node = far;
            else node = up;
        } else if (prev == near && far >= 0 && delta * delta <= r2) {
            prev = node;
            node = far;
        } else {
            prev = node;
            node = up;
        }
    }
    *cell = s1;
    *inside = 1;
    return 1;
}

/* Periodic wrap into [0, L). */
static inline double og_wrap(double x, double L)
{
    double w = x - L * floor(x / L);
    return w >= L ? w - L : w;
}

static int og_classify_point(const og_geom* g, double px, double py, double pz, int32_t* cell, int8_t* inside)
{
    if (g->free_water) { *cell = 0; *inside = 0; return 1; }
    if (g->periodic) { px = og_wrap(px, g->L); py = og_wrap(py, g->L); pz = og_wrap(pz, g->L); }
    int ok = og_classify_raw(g, px, py, pz, cell, inside);
    if (ok && g->orig_id) *cell = g->orig_id[*cell];
    return ok;
}

void og_classify(const og_geom* g, const double* pts, int64_t m, int32_t* cell, int8_t* inside, int8_t* ok)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i)
        ok[i] = (int8_t)og_classify_point(g, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], &cell[i], &inside[i]);
}

/* All-seed S2 evaluation, an independent check of og_classify
 * (tests/physics_audit/test_geometry_diagnostics.py:119-140). */
void og_classify_brute(const double* seeds, const double* annulus, int64_t n, const double* pts, int64_t m,
                       int32_t* cell, int8_t* inside)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i) {
        const double* p = pts + 3 * i;
        int64_t best = 0;
        double bd = INFINITY;
        for (int64_t j = 0; j < n; ++j) {
            double dx = p[0] - seeds[3 * j], dy = p[1] - seeds[3 * j + 1], dz = p[2] - seeds[3 * j + 2];
            double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 < bd) { bd = d2; best = j; }
        }
        int8_t in = 1;
        for (int64_t j = 0; j < n && in; ++j) {
            if (j == best) continue;
            double dx = p[0] - seeds[3 * j], dy = p[1] - seeds[3 * j + 1], dz = p[2] - seeds[3 * j + 2];
            double d2 = dx * dx + dy * dy + dz * dz;
            double sx = seeds[3 * j] - seeds[3 * best], sy = seeds[3 * j + 1] - seeds[3 * best + 1],
                   sz = seeds[3 * j + 2] - seeds[3 * best + 2];
            double sep = sqrt(sx * sx + sy * sy + sz * sz);
            if ((d2 - bd) / (2.0 * sep) < annulus[best]) in = 0;
        }
        cell[i] = (int32_t)best;
        inside[i] = in;
    }
}

/* -------------------------------------------------------------- walk */

typedef struct {
    int64_t n_walkers;
    int32_t n_steps;
    int32_t steps_per_h;
    int32_t n_grid;
    int32_t use_golden;
    double sigma;
    double ts;
    double pp;
    double source_lo;
    double source_hi;
    uint64_t seed;         /* walk seed when use_golden == 0 */
    uint64_t first_walker; /* RNG subsequence of walker 0 */
} og_walk_params;

/* One walker of the reference step loop (madi/walker_gpu.py:531-656). */
static void og_walk_one(const og_geom* g, const og_walk_params* P, int64_t w, og_state st,
                        const double* gi, const double* gd, const double* gu,
                        double* Y, int8_t* itr, int8_t* esc, int8_t* err)
{
    const int64_t N = P->n_walkers;
    const int G = P->n_grid;
    double* Yw = Y + (size_t)w * G * 3;
    int8_t* Iw = itr + (size_t)w * G;
    double px, py, pz;
    double span = P->source_hi - P->source_lo;
    if (P->use_golden) {
        px = gi[3 * w]; py = gi[3 * w + 1]; pz = gi[3 * w + 2];
    } else {
        px = P->source_lo + og_uniform_f64(&st) * span;
        py = P->source_lo + og_uniform_f64(&st) * span;
        pz = P->source_lo + og_uniform_f64(&st) * span;
    }
    int32_t cell;
    int8_t in;
    if (!og_classify_point(g, px, py, pz, &cell, &in)) {
        err[w] = 1; esc[w] = 1;
        for (int j = 0; j < G; ++j) { Yw[3 * j] = Yw[3 * j + 1] = Yw[3 * j + 2] = 0.0; Iw[j] = 0; }
        return;
    }
    const double half = 0.5 * g->L;
    double xp = px - half, yp = py - half, zp = pz - half;
    double Yx = 0.0, Yy = 0.0, Yz = 0.0;
    Yw[0] = Yw[1] = Yw[2] = 0.0;
    Iw[0] = in;
    int escaped = 0;
    int j_done = 0;
    for (int32_t step = 0; step < P->n_steps; ++step) {
        double dx, dy, dz, u;
        if (P->use_golden) {
            size_t k = (size_t)step * N + w;
            dx = gd[3 * k]; dy = gd[3 * k + 1]; dz = gd[3 * k + 2];
            u = gu[k];
        } else {
            dx = og_normal_f64(&st) * P->sigma;
            dy = og_normal_f64(&st) * P->sigma;
            dz = og_normal_f64(&st) * P->sigma;
            u = og_uniform_f64(&st);
        }
        double nx = px + dx, ny = py + dy, nz = pz + dz;
        if (!g->periodic && (nx < 0.0 || nx >= g->L || ny < 0.0 || ny >= g->L || nz < 0.0 || nz >= g->L)) {
            escaped = 1;
            break;
        }
        int32_t ncell;
        int8_t nin;
        if (!og_classify_point(g, nx, ny, nz, &ncell, &nin)) { err[w] = 1; escaped = 1; break; }
        int m = 0;
        if (in != nin) m = 1;
        else if (in && nin && cell != ncell) m = 2;
        if (m > 0 && P->pp < 1.0) {
            double thr = m == 1 ? P->pp : P->pp * P->pp;
            if (u >= thr) { nx = px; ny = py; nz = pz; ncell = cell; nin = in; }
        }
        px = nx; py = ny; pz = nz;
        cell = ncell; in = nin;
        double xn = px - half, yn = py - half, zn = pz - half;
        Yx += 0.5 * (xp + xn) * P->ts;
        Yy += 0.5 * (yp + yn) * P->ts;
        Yz += 0.5 * (zp + zn) * P->ts;
        xp = xn; yp = yn; zp = zn;
        int32_t s1 = step + 1;
        if (s1 % P->steps_per_h == 0) {
            int j = s1 / P->steps_per_h;
            Yw[3 * j] = Yx; Yw[3 * j + 1] = Yy; Yw[3 * j + 2] = Yz;
            Iw[j] = in;
            j_done = j;
        }
    }
    /* Samples after an escape keep the frozen running integral and label,
     * as the reference CPU walk does (madi/walker_gpu.py:834-837). */
    for (int j = j_done + 1; j < G; ++j) { Yw[3 * j] = Yx; Yw[3 * j + 1] = Yy; Yw[3 * j + 2] = Yz; Iw[j] = in; }
    esc[w] = (int8_t)escaped;
}

void og_walk(const og_geom* g, const og_walk_params* P, const uint64_t* states,
             const double* golden_init, const double* golden_incr, const double* golden_u,
             double* Y, int8_t* inside_trace, int8_t* escaped, int8_t* err)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t w = 0; w < P->n_walkers; ++w) {
        og_state st = {0, 0};
        if (!P->use_golden) { st.s0 = states[2 * w]; st.s1 = states[2 * w + 1]; }
        err[w] = 0;
        og_walk_one(g, P, w, st, golden_init, golden_incr, golden_u, Y, inside_trace, escaped, err);
    }
}

/* ------------------------------------------------------------ reduce */

/* cos/sin sums of pc*(Y[jd]+Y[jD]-Y[js]) over walkers and axes
 * (madi/walker_gpu.py:954-968).  Summation is sequential in walker-axis
 * order, so tests compare sums with a relative tolerance. */
void og_reduce(const double* Y, int64_t n, int32_t n_grid, const int32_t* jd, const int32_t* jD, const int32_t* js,
               const double* pc, int32_t n_cols, double* cos_out, double* sin_out)
{
#pragma omp parallel for schedule(static)
    for (int32_t c = 0; c < n_cols; ++c) {
        double cs = 0.0, sn = 0.0;
        for (int64_t w = 0; w < n; ++w)
            for (int a = 0; a < 3; ++a) {
                const double* Yw = Y + (size_t)w * n_grid * 3;
                double dM = Yw[3 * jd[c] + a] + Yw[3 * jD[c] + a] - Yw[3 * js[c] + a];
                double ph = pc[c] * dM;
                cs += cos(ph);
                sn += sin(ph);
            }
        cos_out[c] = cs;
        sin_out[c] = sn;
    }
}

/* ------------------------------------------------------- MAP matching */

/* The reference's per-voxel GPU matchers (madi/fitters_gpu.py:96-118 fixed
 * S0, :120-154 free S0, called from madi/library.py:1402-1522): squared norms
 * and dot products accumulate over the features in order with fused
 * multiply-adds (the contraction GPU compilers apply to the kernels'
 * `dot += a * b`), and the scan over library rows keeps the first strict
 * minimum.  fit_s0 = 0: score = |r|^2 - 2 m.r, resid = |m|^2 + best score.
 * fit_s0 = 1: rr = max(|r|^2, 1e-30) (NaN kept, as np.maximum),
 * s0 = m.r / rr, resid = |m|^2 - (m.r)^2 / rr, +inf where s0 <= 0.
 * The reference's CPU path (BLAS) sums in another order; tests compare the
 * two within a tolerance. */
static double og_sumsq(const double* x, int32_t nf)
{
    double s = 0.0;
    for (int32_t k = 0; k < nf; ++k) s = fma(x[k], x[k], s);
    return s;
}

void og_map_match(const double* M, int64_t n_vox, const double* R, int64_t n_lib, int32_t nf, int32_t fit_s0,
                  int64_t* idx, double* resid, double* s0_out)
{
    double* rn = (double*)malloc((size_t)(n_lib > 0 ? n_lib : 1) * sizeof(double));
    if (!rn) return;
    for (int64_t j = 0; j < n_lib; ++j) {
        double s = og_sumsq(R + (size_t)j * nf, nf);
        if (fit_s0 && s < 1e-30) s = 1e-30;
        rn[j] = s;
    }
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < n_vox; ++v) {
        const double* m = M + (size_t)v * nf;
        const double mm = og_sumsq(m, nf);
        double best = INFINITY, bs0 = 0.0;
        int64_t bj = 0;
        for (int64_t j = 0; j < n_lib; ++j) {
            const double* r = R + (size_t)j * nf;
            double dot = 0.0;
            for (int32_t k = 0; k < nf; ++k) dot = fma(m[k], r[k], dot);
            if (fit_s0) {
                const double s0 = dot / rn[j];
                double res = mm - (dot * dot) / rn[j];
                if (s0 <= 0.0) res = INFINITY;
                if (res < best) {
                    best = res;
                    bj = j;
                    bs0 = s0;
                }
            } else {
                const double score = rn[j] - 2.0 * dot;
                if (score < best) {
                    best = score;
                    bj = j;
                }
            }
        }
        idx[v] = bj;
        resid[v] = fit_s0 ? best : mm + best;
        if (fit_s0 && s0_out) s0_out[v] = bs0;
    }
    free(rn);
}
