// FP64 exact SI Eq. S2 classifier over a balanced KD tree (device code).
//
// Semantics follow _nearest_one_kdtree / _classify_full_facet
// (madi/walker_gpu.py:222-335): the exact nearest seed s1, then every seed
// within d1 + 2*alpha1 of the point is tested with the shifted-facet margin
// (dj^2 - d1^2) / (2 |c1 - cj|) >= alpha1.  Both passes are stackless
// parent-pointer traversals; the first levels of the tree (breadth-first
// prefix of the node arrays) are read from LDS.
//
// This header must be compiled with FP contraction off so its FP64
// arithmetic keeps the reference's operation order.
#pragma once

#include "common.hpp"

namespace madi {

struct KdView {
    const KdNodeF* gf;
    const KdNodeI* gi;
    const KdNodeF* sf;  // LDS copy of nodes [0, n_lds)
    const KdNodeI* si;
    int n_lds;
};

__device__ __forceinline__ void kd_load(const KdView& v, int node, KdNodeF& f, KdNodeI& i)
{
    if (node < v.n_lds) {
        f = v.sf[node];
        i = v.si[node];
    } else {
        f = v.gf[node];
        i = v.gi[node];
    }
}

__device__ __forceinline__ double axis_of(double x, double y, double z, int ax) { return ax == 0 ? x : (ax == 1 ? y : z); }

// Nearest KD point: returns its index (KD point id, not node) and d^2.
__device__ __forceinline__ int kd_nearest(const KdView& v, double px, double py, double pz, double& best_d2,
                                          double& bx, double& by, double& bz, long long& nodes)
{
    int node = 0, prev = -1, best = -1;
    double bd = __builtin_inf();
    bx = by = bz = 0.0;
    while (node >= 0) {
        KdNodeF f;
        KdNodeI m;
        kd_load(v, node, f, m);
        ++nodes;
        const int ax = m.sid_axis & 3;
        const double delta = axis_of(px, py, pz, ax) - axis_of(f.x, f.y, f.z, ax);
        const int near = delta <= 0.0 ? m.left : m.right;
        const int far = delta <= 0.0 ? m.right : m.left;
        if (prev == m.parent) {
            const double dx = px - f.x;
            const double dy = py - f.y;
            const double dz = pz - f.z;
            const double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 < bd) {
                bd = d2;
                best = m.sid_axis >> 2;
                bx = f.x;
                by = f.y;
                bz = f.z;
            }
            prev = node;
            if (near >= 0) node = near;
            else if (far >= 0 && delta * delta < bd) node = far;
            else node = m.parent;
        } else if (prev == near && far >= 0 && delta * delta < bd) {
            prev = node;
            node = far;
        } else {
            prev = node;
            node = m.parent;
        }
    }
    best_d2 = bd;
    return best;
}

// Exact label.  Returns ok (0 on coincident seeds or an empty tree); cell is
// the KD point id of the nearest seed (callers map it through kd_label).
__device__ __forceinline__ int kd_classify(const KdView& v, const double* annulus, double px, double py, double pz,
                                           int& cell, int& inside, long long& nodes)
{
    double d1sq, cx, cy, cz;
    const int s1 = kd_nearest(v, px, py, pz, d1sq, cx, cy, cz, nodes);
    if (s1 < 0) {
        cell = 0;
        inside = 0;
        return 0;
    }
    const double a1 = annulus[s1];
    const double radius = sqrt(d1sq) + 2.0 * a1;
    const double r2 = radius * radius;
    int node = 0, prev = -1;
    while (node >= 0) {
        KdNodeF f;
        KdNodeI m;
        kd_load(v, node, f, m);
        ++nodes;
        const int ax = m.sid_axis & 3;
        const double delta = axis_of(px, py, pz, ax) - axis_of(f.x, f.y, f.z, ax);
        const int near = delta <= 0.0 ? m.left : m.right;
        const int far = delta <= 0.0 ? m.right : m.left;
        if (prev == m.parent) {
            const int sid = m.sid_axis >> 2;
            const double dx = px - f.x;
            const double dy = py - f.y;
            const double dz = pz - f.z;
            const double dj2 = dx * dx + dy * dy + dz * dz;
            if (sid != s1 && dj2 <= r2) {
                const double sx = f.x - cx;
                const double sy = f.y - cy;
                const double sz = f.z - cz;
                const double sep = sqrt(sx * sx + sy * sy + sz * sz);
                if (sep <= 0.0) {
                    cell = 0;
                    inside = 0;
                    return 0;
                }
                const double margin = (dj2 - d1sq) / (2.0 * sep);
                if (margin < a1) {
                    cell = s1;
                    inside = 0;
                    return 1;
                }
            }
            prev = node;
            if (near >= 0) node = near;
            else if (far >= 0 && delta * delta <= r2) node = far;
            else node = m.parent;
        } else if (prev == near && far >= 0 && delta * delta <= r2) {
            prev = node;
            node = far;
        } else {
            prev = node;
            node = m.parent;
        }
    }
    cell = s1;
    inside = 1;
    return 1;
}

__device__ __forceinline__ double wrap_periodic(double x, double L)
{
    double w = x - L * floor(x / L);
    return w >= L ? w - L : w;
}

// Full point classification with periodic wrap and label map.
__device__ __forceinline__ int classify_point_f64(const GeomDev& g, const KdView& v, double px, double py, double pz,
                                                 int& cell, int& inside, long long& nodes)
{
    if (g.free_water) {
        cell = 0;
        inside = 0;
        return 1;
    }
    if (g.periodic) {
        px = wrap_periodic(px, g.L);
        py = wrap_periodic(py, g.L);
        pz = wrap_periodic(pz, g.L);
    }
    const int ok = kd_classify(v, g.kd_annulus, px, py, pz, cell, inside, nodes);
    if (ok) cell = g.kd_label[cell];
    return ok;
}

// Stage the first n_lds KD nodes into LDS (call with all threads of a block).
__device__ __forceinline__ KdView kd_stage(const GeomDev& g, KdNodeF* sf, KdNodeI* si, int cap)
{
    KdView v;
    v.gf = g.kd_f;
    v.gi = g.kd_i;
    v.sf = sf;
    v.si = si;
    int n = g.n_kd_lds < cap ? g.n_kd_lds : cap;
    if (g.free_water) n = 0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        sf[k] = g.kd_f[k];
        si[k] = g.kd_i[k];
    }
    __syncthreads();
    v.n_lds = n;
    return v;
}

}  // namespace madi
