// Host-side geometry construction of libmadi_hip.so (CPU only, no device call):
//
// * madi_nearest_points: the nearest point of a set to each query, plain or
//   periodic (minimum image in a cube), on a uniform bin grid searched in
//   rings of growing Chebyshev radius;
// * madi_voronoi_graph: the Voronoi neighbours (the Delaunay graph's edges)
//   of the first n_owned points of a set, each cell cut out of a box by the
//   bisector planes of its candidates in order of distance until no further
//   point can cut it.
//
// A cell is held in the dual form: every vertex is the triangle of the three
// planes that meet there, oriented so that (n_a x n_b) . n_c > 0.  A cut by
// plane q removes the vertices beyond it; the directed plane pairs (a, b) of
// removed triangles whose reverse pair belongs to a kept triangle bound the
// hole, and each becomes the new vertex (a, b, q).  A candidate point at
// distance |v| can cut the cell only if |v| / 2 <= R, the largest vertex
// distance of the cell so far, so the search stops once every unvisited
// point is farther than 2 R (the cell's security radius).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/madi_hip.h"

namespace madi {
void set_error(const char* fmt, ...);
}

namespace {

using madi::set_error;

// Points sorted by bin.  Bins cover [lo, lo + nb * h) per axis; a periodic
// grid covers [0, period)^3 and wraps.
struct BinGrid {
    double lo[3] = {0, 0, 0};
    double h = 1.0;
    int nb[3] = {1, 1, 1};
    bool periodic = false;
    double period = 0.0;
    std::vector<int64_t> start;  // first sorted slot of each bin, plus the end
    std::vector<int32_t> ids;    // point ids, bin by bin (ascending within a bin)
    std::vector<double> xyz;     // the points in the same order (wrapped when periodic)

    int64_t bin_of(int i, int j, int k) const { return ((int64_t)k * nb[1] + j) * nb[0] + i; }

    int coord(double x, int d) const
    {
        int i = (int)std::floor((x - lo[d]) / h);
        if (periodic) {
            i %= nb[d];
            if (i < 0) i += nb[d];
            return i;
        }
        return i < 0 ? 0 : (i >= nb[d] ? nb[d] - 1 : i);
    }
};

double wrap_coord(double x, double L)
{
    double y = std::fmod(x, L);
    if (y < 0.0) y += L;
    if (y >= L) y = 0.0;
    return y;
}

void build_grid(BinGrid& G, const double* pts, int64_t n, double period, double points_per_bin)
{
    G.periodic = period > 0.0;
    G.period = period;
    double mn[3], mx[3];
    if (G.periodic) {
        for (int d = 0; d < 3; ++d) {
            mn[d] = 0.0;
            mx[d] = period;
        }
    } else {
        for (int d = 0; d < 3; ++d) {
            mn[d] = INFINITY;
            mx[d] = -INFINITY;
        }
        for (int64_t i = 0; i < n; ++i)
            for (int d = 0; d < 3; ++d) {
                mn[d] = std::min(mn[d], pts[3 * i + d]);
                mx[d] = std::max(mx[d], pts[3 * i + d]);
            }
    }
    double ext[3], vol = 1.0;
    for (int d = 0; d < 3; ++d) {
        ext[d] = mx[d] - mn[d];
        if (!(ext[d] > 0.0)) ext[d] = 0.0;
    }
    const double emax = std::max(ext[0], std::max(ext[1], ext[2]));
    const double floor_ext = emax > 0.0 ? emax * 1e-3 : 1.0;
    for (int d = 0; d < 3; ++d) vol *= std::max(ext[d], floor_ext);
    double h = std::cbrt(vol * points_per_bin / (double)std::max<int64_t>(n, 1));
    if (!(h > 0.0) || !std::isfinite(h)) h = 1.0;
    // at most ~2^22 bins
    for (;;) {
        double nbins = 1.0;
        for (int d = 0; d < 3; ++d) nbins *= std::max(1.0, std::ceil(ext[d] / h));
        if (nbins <= 4194304.0) break;
        h *= 1.25;
    }
    if (G.periodic) {
        const int m = std::max(1, (int)std::floor(period / h));
        G.h = period / m;
        for (int d = 0; d < 3; ++d) {
            G.nb[d] = m;
            G.lo[d] = 0.0;
        }
    } else {
        G.h = h;
        for (int d = 0; d < 3; ++d) {
            G.nb[d] = std::max(1, (int)std::ceil(ext[d] / h));
            G.lo[d] = mn[d];
        }
    }
    const int64_t nbins = (int64_t)G.nb[0] * G.nb[1] * G.nb[2];
    std::vector<int64_t> bin((size_t)n);
    G.start.assign((size_t)nbins + 1, 0);
    for (int64_t i = 0; i < n; ++i) {
        double p[3];
        for (int d = 0; d < 3; ++d) p[d] = G.periodic ? wrap_coord(pts[3 * i + d], period) : pts[3 * i + d];
        bin[(size_t)i] = G.bin_of(G.coord(p[0], 0), G.coord(p[1], 1), G.coord(p[2], 2));
        ++G.start[(size_t)bin[(size_t)i] + 1];
    }
    for (int64_t b = 0; b < nbins; ++b) G.start[(size_t)b + 1] += G.start[(size_t)b];
    std::vector<int64_t> fill(G.start.begin(), G.start.end() - 1);
    G.ids.resize((size_t)n);
    G.xyz.resize((size_t)(3 * n));
    for (int64_t i = 0; i < n; ++i) {
        const int64_t s = fill[(size_t)bin[(size_t)i]]++;
        G.ids[(size_t)s] = (int32_t)i;
        for (int d = 0; d < 3; ++d)
            G.xyz[(size_t)(3 * s + d)] = G.periodic ? wrap_coord(pts[3 * i + d], period) : pts[3 * i + d];
    }
}

// Calls f(bin) for every bin at Chebyshev distance exactly r from (ci, cj,
// ck); returns false when no such bin exists (a plain grid already covered).
template <typename F>
bool for_ring(const BinGrid& G, int ci, int cj, int ck, int r, F&& f)
{
    const int c[3] = {ci, cj, ck};
    if (!G.periodic) {
        bool any = false;
        for (int d = 0; d < 3; ++d)
            if (c[d] - r >= 0 || c[d] + r < G.nb[d]) any = true;
        if (!any) return false;
    } else {
        int nmax = std::max(G.nb[0], std::max(G.nb[1], G.nb[2]));
        if (2 * r - 1 >= nmax) return false;  // rings 0 .. r-1 already wrapped over every bin
    }
    auto visit = [&](int di, int dj, int dk) {
        int i = ci + di, j = cj + dj, k = ck + dk;
        if (G.periodic) {
            i = ((i % G.nb[0]) + G.nb[0]) % G.nb[0];
            j = ((j % G.nb[1]) + G.nb[1]) % G.nb[1];
            k = ((k % G.nb[2]) + G.nb[2]) % G.nb[2];
        } else if (i < 0 || j < 0 || k < 0 || i >= G.nb[0] || j >= G.nb[1] || k >= G.nb[2]) {
            return;
        }
        f(G.bin_of(i, j, k));
    };
    if (r == 0) {
        visit(0, 0, 0);
        return true;
    }
    for (int dk = -r; dk <= r; ++dk)
        for (int dj = -r; dj <= r; ++dj) {
            if (dk == -r || dk == r || dj == -r || dj == r) {
                for (int di = -r; di <= r; ++di) visit(di, dj, dk);
            } else {
                visit(-r, dj, dk);
                visit(r, dj, dk);
            }
        }
    return true;
}

// Squared distance in the order (dx^2 + dy^2) + dz^2, minimum image when periodic.
inline double dist2(const BinGrid& G, const double* a, const double* b)
{
    double s = 0.0;
    for (int d = 0; d < 3; ++d) {
        double x = a[d] - b[d];
        if (G.periodic) {
            const double hb = 0.5 * G.period;
            if (x < -hb) x = G.period + x;
            else if (x > hb) x = x - G.period;
        }
        s += x * x;
    }
    return s;
}

template <typename F>
void parallel_for(int64_t n, int n_threads, F&& f)
{
    int T = n_threads > 0 ? std::min(n_threads, 64) : (int)std::min(8u, std::max(1u, std::thread::hardware_concurrency()));
    if (T <= 1 || n < 256) {
        f(0, n, 0);
        return;
    }
    if ((int64_t)T > n / 64) T = (int)std::max<int64_t>(1, n / 64);
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t) {
        const int64_t lo = n * t / T, hi = n * (t + 1) / T;
        th.emplace_back([&f, lo, hi, t] { f(lo, hi, t); });
    }
    for (auto& x : th) x.join();
}

// ------------------------------------------------------------ Voronoi cell

struct Plane {
    double nx, ny, nz, d;  // half-space n . x <= d (unit n), seed-relative
    int32_t id;            // point id, or -1 for a face of the box
};

struct Tri {
    int32_t a, b, c;  // plane indices
    double x, y, z;   // the vertex
};

struct Cell {
    std::vector<Plane> planes;
    std::vector<Tri> tris, next;
    std::vector<uint8_t> out;
    std::vector<std::pair<int32_t, int32_t>> removed_edges, bound;
    std::vector<int32_t> touched;
    double R2 = 0.0;

    bool vertex(int a, int b, int c, double& x, double& y, double& z) const
    {
        const Plane &A = planes[a], &B = planes[b], &C = planes[c];
        // n_b x n_c, n_c x n_a, n_a x n_b
        const double bcx = B.ny * C.nz - B.nz * C.ny, bcy = B.nz * C.nx - B.nx * C.nz, bcz = B.nx * C.ny - B.ny * C.nx;
        const double cax = C.ny * A.nz - C.nz * A.ny, cay = C.nz * A.nx - C.nx * A.nz, caz = C.nx * A.ny - C.ny * A.nx;
        const double abx = A.ny * B.nz - A.nz * B.ny, aby = A.nz * B.nx - A.nx * B.nz, abz = A.nx * B.ny - A.ny * B.nx;
        const double det = A.nx * bcx + A.ny * bcy + A.nz * bcz;
        if (!(std::fabs(det) > 1e-300)) return false;
        x = (A.d * bcx + B.d * cax + C.d * abx) / det;
        y = (A.d * bcy + B.d * cay + C.d * aby) / det;
        z = (A.d * bcz + B.d * caz + C.d * abz) / det;
        return std::isfinite(x) && std::isfinite(y) && std::isfinite(z);
    }

    void update_radius()
    {
        R2 = 0.0;
        for (const Tri& t : tris) R2 = std::max(R2, t.x * t.x + t.y * t.y + t.z * t.z);
    }

    // the box [blo, bhi] in seed-relative coordinates
    bool reset(const double* blo, const double* bhi)
    {
        planes.clear();
        tris.clear();
        touched.clear();
        planes.push_back({1, 0, 0, bhi[0], -1});
        planes.push_back({-1, 0, 0, -blo[0], -1});
        planes.push_back({0, 1, 0, bhi[1], -1});
        planes.push_back({0, -1, 0, -blo[1], -1});
        planes.push_back({0, 0, 1, bhi[2], -1});
        planes.push_back({0, 0, -1, -blo[2], -1});
        for (int sx = 0; sx < 2; ++sx)
            for (int sy = 0; sy < 2; ++sy)
                for (int sz = 0; sz < 2; ++sz) {
                    int a = sx, b = 2 + sy, c = 4 + sz;  // planes +x / -x, +y / -y, +z / -z
                    const Plane &A = planes[a], &B = planes[b], &C = planes[c];
                    const double tp = (A.ny * B.nz - A.nz * B.ny) * C.nx + (A.nz * B.nx - A.nx * B.nz) * C.ny +
                                      (A.nx * B.ny - A.ny * B.nx) * C.nz;
                    if (tp < 0.0) std::swap(a, b);
                    Tri t{a, b, c, 0, 0, 0};
                    if (!vertex(a, b, c, t.x, t.y, t.z)) return false;
                    tris.push_back(t);
                }
        update_radius();
        return true;
    }

    // 0: no effect; 1: cut; 2: touched within eps (recorded); -1: failure
    int cut(const Plane& q, double eps)
    {
        const size_t T = tris.size();
        out.assign(T, 0);
        size_t n_out = 0;
        double smax = -INFINITY;
        for (size_t t = 0; t < T; ++t) {
            const double s = q.nx * tris[t].x + q.ny * tris[t].y + q.nz * tris[t].z - q.d;
            smax = std::max(smax, s);
            if (s > eps) {
                out[t] = 1;
                ++n_out;
            }
        }
        if (smax <= -eps) return 0;
        if (n_out == 0) {
            touched.push_back(q.id);
            return 2;
        }
        if (n_out == T) return -1;
        removed_edges.clear();
        for (size_t t = 0; t < T; ++t)
            if (out[t]) {
                const Tri& r = tris[t];
                removed_edges.emplace_back(r.a, r.b);
                removed_edges.emplace_back(r.b, r.c);
                removed_edges.emplace_back(r.c, r.a);
            }
        std::sort(removed_edges.begin(), removed_edges.end());
        bound.clear();
        for (const auto& e : removed_edges)
            if (!std::binary_search(removed_edges.begin(), removed_edges.end(), std::make_pair(e.second, e.first)))
                bound.push_back(e);
        if (bound.size() < 3) return -1;
        const int32_t qi = (int32_t)planes.size();
        planes.push_back(q);
        next.clear();
        for (size_t t = 0; t < T; ++t)
            if (!out[t]) next.push_back(tris[t]);
        for (const auto& e : bound) {
            Tri t{e.first, e.second, qi, 0, 0, 0};
            if (!vertex(t.a, t.b, t.c, t.x, t.y, t.z)) return -1;
            next.push_back(t);
        }
        tris.swap(next);
        update_radius();
        return 1;
    }
};

}  // namespace

extern "C" {

int madi_nearest_points(const double* pts, int64_t n, const double* q, int64_t nq, double period,
                        int32_t exclude_self, int32_t n_threads, int32_t* idx, double* dist)
{
    if (n < 0 || nq < 0 || n > INT32_MAX || (n > 0 && !pts) || (nq > 0 && !q) || !(period >= 0.0) ||
        !std::isfinite(period)) {
        set_error("madi_nearest_points: invalid arguments");
        return MADI_ERR_ARG;
    }
    if (nq == 0) return MADI_OK;
    if (n == 0 || (exclude_self && n < 2)) {
        set_error("madi_nearest_points: the point set has no candidate");
        return MADI_ERR_ARG;
    }
    for (int64_t i = 0; i < 3 * n; ++i)
        if (!std::isfinite(pts[i])) {
            set_error("madi_nearest_points: non-finite point coordinate");
            return MADI_ERR_ARG;
        }
    BinGrid G;
    build_grid(G, pts, n, period, 2.0);
    parallel_for(nq, n_threads, [&](int64_t lo, int64_t hi, int) {
        for (int64_t i = lo; i < hi; ++i) {
            double p[3];
            for (int d = 0; d < 3; ++d) p[d] = G.periodic ? wrap_coord(q[3 * i + d], period) : q[3 * i + d];
            const int ci = G.coord(p[0], 0), cj = G.coord(p[1], 1), ck = G.coord(p[2], 2);
            double best = INFINITY;
            int32_t arg = -1;
            const int32_t self = exclude_self ? (int32_t)i : -1;
            for (int r = 0;; ++r) {
                if (r >= 1 && arg >= 0) {
                    const double lb = (r - 1) * G.h;
                    if (lb * lb > best) break;
                }
                const bool more = for_ring(G, ci, cj, ck, r, [&](int64_t b) {
                    for (int64_t s = G.start[(size_t)b]; s < G.start[(size_t)b + 1]; ++s) {
                        const int32_t id = G.ids[(size_t)s];
                        if (id == self) continue;
                        const double d2 = dist2(G, p, &G.xyz[(size_t)(3 * s)]);
                        if (d2 < best || (d2 == best && id < arg)) {
                            best = d2;
                            arg = id;
                        }
                    }
                });
                if (!more) break;
            }
            if (idx) idx[i] = arg;
            if (dist) dist[i] = std::sqrt(best);
        }
    });
    return MADI_OK;
}

int madi_voronoi_graph(const double* pts, int64_t n, int64_t n_owned, double pad, int32_t n_threads,
                       int64_t* offsets, int32_t* ids, int64_t capacity, int64_t* n_edges, double* radius,
                       int8_t* boxed)
{
    if (n < 0 || n > INT32_MAX || n_owned < 0 || n_owned > n || (n > 0 && !pts) || !offsets || !n_edges ||
        capacity < 0 || (capacity > 0 && !ids) || !(pad > 0.0) || !std::isfinite(pad)) {
        set_error("madi_voronoi_graph: invalid arguments");
        return MADI_ERR_ARG;
    }
    *n_edges = 0;
    offsets[0] = 0;
    if (n_owned == 0) return MADI_OK;
    for (int64_t i = 0; i < 3 * n; ++i)
        if (!std::isfinite(pts[i])) {
            set_error("madi_voronoi_graph: non-finite point coordinate");
            return MADI_ERR_ARG;
        }
    BinGrid G;
    build_grid(G, pts, n, 0.0, 1.0);
    double blo[3], bhi[3];
    for (int d = 0; d < 3; ++d) {
        blo[d] = INFINITY;
        bhi[d] = -INFINITY;
    }
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            blo[d] = std::min(blo[d], pts[3 * i + d]);
            bhi[d] = std::max(bhi[d], pts[3 * i + d]);
        }
    for (int d = 0; d < 3; ++d) {
        blo[d] -= pad;
        bhi[d] += pad;
    }
    // a decision within eps of a plane counts as touching it (the plane is
    // then kept as a neighbour: a superset of the Delaunay edges is safe)
    const double eps = 1e-10 * G.h;
    std::vector<std::vector<int32_t>> nb((size_t)n_owned);
    std::vector<int64_t> bad_coincident(64, -1), bad_numeric(64, -1);

    parallel_for(n_owned, n_threads, [&](int64_t lo, int64_t hi, int t) {
        Cell C;
        std::vector<std::pair<double, int32_t>> cand;
        std::vector<int32_t> list;
        for (int64_t i = lo; i < hi; ++i) {
            const double* s = pts + 3 * i;
            double rl[3], rh[3];
            for (int d = 0; d < 3; ++d) {
                rl[d] = blo[d] - s[d];
                rh[d] = bhi[d] - s[d];
            }
            bool ok = C.reset(rl, rh);
            bool coincident = false;
            const int ci = G.coord(s[0], 0), cj = G.coord(s[1], 1), ck = G.coord(s[2], 2);
            // rings 0-2 in one batch sorted by distance (the nearest points,
            // which bound the cell, cut first), then one ring at a time
            for (int r = 0; ok;) {
                const int r_hi = r == 0 ? 2 : r;
                if (r >= 1) {
                    const double lb = (r - 1) * G.h;
                    if (lb > 2.0 * std::sqrt(C.R2) + 4.0 * eps) break;
                }
                cand.clear();
                // only points within twice the current vertex distance can cut
                const double reach = 2.0 * std::sqrt(C.R2) + 4.0 * eps, reach2 = reach * reach;
                bool more = false;
                for (int rr = r; rr <= r_hi; ++rr)
                    more |= for_ring(G, ci, cj, ck, rr, [&](int64_t b) {
                        for (int64_t k = G.start[(size_t)b]; k < G.start[(size_t)b + 1]; ++k) {
                            const int32_t id = G.ids[(size_t)k];
                            if (id == (int32_t)i) continue;
                            const double* p = &G.xyz[(size_t)(3 * k)];
                            const double vx = p[0] - s[0], vy = p[1] - s[1], vz = p[2] - s[2];
                            const double d2 = (vx * vx + vy * vy) + vz * vz;
                            if (d2 <= reach2) cand.emplace_back(d2, id);
                        }
                    });
                r = r_hi + 1;
                if (!more) break;
                std::sort(cand.begin(), cand.end());
                for (const auto& cd : cand) {
                    const double len = std::sqrt(cd.first);
                    if (len > 2.0 * std::sqrt(C.R2) + 4.0 * eps) break;
                    if (!(len > 0.0)) {
                        coincident = true;
                        ok = false;
                        break;
                    }
                    const double* p = pts + 3 * (int64_t)cd.second;
                    const Plane q{(p[0] - s[0]) / len, (p[1] - s[1]) / len, (p[2] - s[2]) / len, 0.5 * len, cd.second};
                    if (C.cut(q, eps) < 0) {
                        ok = false;
                        break;
                    }
                }
            }
            list.clear();
            bool at_box = false;
            if (ok) {
                for (const Tri& tr : C.tris)
                    for (int32_t pi : {tr.a, tr.b, tr.c}) {
                        const int32_t id = C.planes[(size_t)pi].id;
                        if (id < 0) at_box = true;
                        else list.push_back(id);
                    }
                // a touched plane stays a neighbour while some vertex lies within eps of it
                for (int32_t id : C.touched) {
                    const double* p = pts + 3 * (int64_t)id;
                    const double vx = p[0] - s[0], vy = p[1] - s[1], vz = p[2] - s[2];
                    const double len = std::sqrt((vx * vx + vy * vy) + vz * vz);
                    double smax = -INFINITY;
                    for (const Tri& tr : C.tris)
                        smax = std::max(smax, (vx * tr.x + vy * tr.y + vz * tr.z) / len - 0.5 * len);
                    if (smax > -eps) list.push_back(id);
                }
                std::sort(list.begin(), list.end());
                list.erase(std::unique(list.begin(), list.end()), list.end());
                if (list.empty()) ok = false;
            }
            if (!ok) {
                std::vector<int64_t>& bad = coincident ? bad_coincident : bad_numeric;
                if (bad[(size_t)t] < 0) bad[(size_t)t] = i;
                list.clear();
            }
            nb[(size_t)i] = list;
            if (radius) radius[i] = std::sqrt(C.R2);
            if (boxed) boxed[i] = at_box ? 1 : 0;
        }
    });
    int64_t first_coincident = -1, first_numeric = -1;
    for (size_t t = 0; t < bad_coincident.size(); ++t) {
        if (bad_coincident[t] >= 0 && (first_coincident < 0 || bad_coincident[t] < first_coincident))
            first_coincident = bad_coincident[t];
        if (bad_numeric[t] >= 0 && (first_numeric < 0 || bad_numeric[t] < first_numeric))
            first_numeric = bad_numeric[t];
    }
    if (first_coincident >= 0) {
        int64_t count = 0;
        for (int64_t i = 0; i < n_owned; ++i) count += nb[(size_t)i].empty();
        set_error("GPU full-facet classification failed: %lld seed(s) (first: %lld) coincide with another seed, "
                  "so their cells have no facets",
                  (long long)count, (long long)first_coincident);
        return MADI_ERR_CLASSIFIER;
    }
    if (first_numeric >= 0) {
        set_error("madi_voronoi_graph: the cell of point %lld could not be cut (degenerate input)",
                  (long long)first_numeric);
        return MADI_ERR_CLASSIFIER;
    }
    int64_t total = 0;
    for (int64_t i = 0; i < n_owned; ++i) {
        total += (int64_t)nb[(size_t)i].size();
        offsets[i + 1] = total;
    }
    *n_edges = total;
    if (total > capacity) {
        set_error("madi_voronoi_graph: %lld edges exceed the capacity %lld", (long long)total, (long long)capacity);
        return MADI_ERR_ARG;
    }
    for (int64_t i = 0; i < n_owned; ++i)
        std::copy(nb[(size_t)i].begin(), nb[(size_t)i].end(), ids + offsets[i]);
    return MADI_OK;
}

}  // extern "C"
