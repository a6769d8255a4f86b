"""Native host geometry of libmadi_hip.so (CPU only, no device call).

* ``madi_nearest_points`` against scipy's cKDTree (plain and periodic, a
  point set's own nearest other point and free queries): bitwise equal
  distances and indices.
* ``madi_voronoi_graph`` against two independent statements of the
  Voronoi neighbour relation:
  - qhull's Delaunay triangulation (scipy.spatial.Delaunay) on random point
    sets: equal neighbour sets for every cell that does not reach the box,
    a subset of qhull's for those that do, and equal sets for the owned
    cells of periodic image-padded boxes;
  - the definition itself, by linear programming: j is a neighbour of i iff
    some point of the i-j bisector plane is strictly nearer to i and j than
    to every other point (the facet has an interior), i.e. the LP
    max t s.t. |x - p_k|^2 - |x - p_i|^2 >= t (all k != i, j) on that plane
    has t* > 0.
* The periodic padding certificate: the graph of a box does not change when
  the image margin is doubled.
"""

from __future__ import annotations

import numpy as np
import pytest
from scipy.optimize import linprog
from scipy.spatial import Delaunay, cKDTree

from custom_madi_amd import ensemble as E


def _qhull_sets(pts, owned):
    ip, ind = Delaunay(pts).vertex_neighbor_vertices
    return [set(ind[ip[i]:ip[i + 1]].tolist()) for i in range(owned)]


def _native_sets(offs, ids):
    return [set(ids[offs[i]:offs[i + 1]].tolist()) for i in range(len(offs) - 1)]


@pytest.mark.parametrize("period", [0.0, 37.5])
def test_nearest_points_equal_a_kd_tree(period):
    rng = np.random.default_rng(11)
    hi = period if period > 0 else 50.0
    pts = rng.uniform(0.0, hi, size=(2500, 3))
    qs = rng.uniform(0.0, hi, size=(4000, 3))
    tree = cKDTree(pts, boxsize=period if period > 0 else None)
    d, i = tree.query(qs, k=1)
    mi, md = E.nearest_points(pts, qs, period=period)
    assert np.array_equal(mi, i) and np.array_equal(md, d)
    d2, i2 = tree.query(pts, k=2)
    si, sd = E.nearest_points(pts, period=period)
    assert np.array_equal(si, i2[:, 1]) and np.array_equal(sd, d2[:, 1])
    # queries far outside a plain point set still find the nearest point
    if period == 0.0:
        far = rng.uniform(-500.0, 550.0, size=(300, 3))
        d, i = tree.query(far, k=1)
        fi, fd = E.nearest_points(pts, far)
        assert np.array_equal(fi, i) and np.array_equal(fd, d)


def test_nearest_points_single_thread_equals_threaded():
    rng = np.random.default_rng(3)
    pts = rng.uniform(0.0, 20.0, size=(3000, 3))
    a = E.nearest_points(pts, n_threads=1)
    b = E.nearest_points(pts, n_threads=8)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("n,seed", [(200, 1), (1500, 2), (4000, 3)])
def test_voronoi_graph_equals_qhull(n, seed):
    pts = np.random.default_rng(seed).uniform(0.0, 100.0, size=(n, 3))
    spacing = (1e6 / n) ** (1.0 / 3.0)
    offs, ids, radius, boxed = E.voronoi_graph(pts, n, pad=spacing)
    mine, ref = _native_sets(offs, ids), _qhull_sets(pts, n)
    inner = np.flatnonzero(boxed == 0)
    assert len(inner) > n // 3
    for i in range(n):
        assert i not in mine[i]
        if boxed[i]:
            assert mine[i] <= ref[i]  # only edges whose facets lie beyond the box are missing
        else:
            assert mine[i] == ref[i], i
    # every vertex of a cell is within its radius, so every neighbour is
    # within twice the radius
    for i in inner:
        d = np.linalg.norm(pts[list(mine[i])] - pts[i], axis=1)
        assert np.all(d <= 2.0 * radius[i] * (1 + 1e-12))
    # threads give the same lists
    o1, i1, _, _ = E.voronoi_graph(pts, n, pad=spacing, n_threads=1)
    assert np.array_equal(o1, offs) and np.array_equal(i1, ids)


def _facet_has_interior(pts, i, j):
    """LP on the i-j bisector: max t with |x-p_k|^2 - |x-p_i|^2 >= t for k != i, j."""
    pi, pj = pts[i], pts[j]
    others = np.array([k for k in range(len(pts)) if k not in (i, j)])
    pk = pts[others]
    # |x-p_k|^2 - |x-p_i|^2 = 2 x.(p_i - p_k) + |p_k|^2 - |p_i|^2 >= t
    A_ub = np.column_stack((-2.0 * (pi - pk), np.ones(len(pk))))
    b_ub = np.sum(pk * pk, axis=1) - pi @ pi
    A_eq = np.array([[*(2.0 * (pj - pi)), 0.0]])  # |x-p_i|^2 = |x-p_j|^2
    b_eq = np.array([pj @ pj - pi @ pi])
    bounds = [(-1e4, 1e4)] * 3 + [(None, 1e3)]
    r = linprog(c=[0, 0, 0, -1.0], A_ub=A_ub, b_ub=b_ub, A_eq=A_eq, b_eq=b_eq, bounds=bounds, method="highs")
    assert r.status == 0
    return -r.fun > 1e-9


def test_voronoi_graph_equals_the_definition_by_linear_programming():
    pts = np.random.default_rng(7).uniform(0.0, 10.0, size=(60, 3))
    offs, ids, _, boxed = E.voronoi_graph(pts, len(pts), pad=2.5)
    mine = _native_sets(offs, ids)
    inner = [i for i in range(len(pts)) if not boxed[i]]
    assert len(inner) >= 5
    for i in inner:
        for j in range(len(pts)):
            if j != i:
                assert (j in mine[i]) == _facet_has_interior(pts, i, j), (i, j)


def test_periodic_graph_is_certified_by_its_margin():
    """The owned cells of an image-padded box equal qhull's and do not change
    when the padding is doubled (every point that can cut a cell is in the
    padded set)."""
    ens = E.create_periodic_ensemble(300, 5.0e5, 0.7, seed=4, with_kd=False)
    arr = E.geometry_arrays(ens, with_kd=False)
    n, L = len(ens.seeds), ens.L
    pts, label, _ = E.periodic_images(ens.seeds, L, arr.pad_margin)
    ref = _qhull_sets(pts, n)
    spacing = (L ** 3 / n) ** (1.0 / 3.0)
    offs, ids, radius, boxed = E.voronoi_graph(pts, n, pad=spacing)
    assert not boxed.any() and 2.0 * radius.max() <= arr.pad_margin
    assert _native_sets(offs, ids) == ref
    big, blabel, _ = E.periodic_images(ens.seeds, L, 2.0 * arr.pad_margin)
    o2, i2, _, _ = E.voronoi_graph(big, n, pad=spacing)

    def canon(p, lab, o, ii):
        return [sorted((int(lab[k]), tuple(np.round(p[k] - p[a], 9))) for k in ii[o[a]:o[a + 1]]) for a in range(n)]

    assert canon(pts, label, offs, ids) == canon(big, blabel, o2, i2)
    # the device arrays carry the same edges (label and image vector)
    got = [sorted((int(arr.nb_ids[k]), tuple(np.round(arr.nb_vec[k], 9)))
                  for k in range(arr.nb_offsets[a], arr.nb_offsets[a + 1])) for a in range(n)]
    assert got == canon(pts, label, offs, ids)


def test_voronoi_graph_argument_errors():
    pts = np.random.default_rng(0).uniform(0.0, 5.0, size=(20, 3))
    with pytest.raises(ValueError):
        E.voronoi_graph(pts, 21, pad=1.0)
    with pytest.raises(ValueError):
        E.voronoi_graph(pts, 20, pad=0.0)
    bad = pts.copy()
    bad[3, 1] = np.nan
    with pytest.raises(ValueError):
        E.voronoi_graph(bad, 20, pad=1.0)
    dup = pts.copy()
    dup[7] = dup[2]
    with pytest.raises(RuntimeError, match="classification failed"):
        E.voronoi_graph(dup, 20, pad=1.0)
    with pytest.raises(ValueError):
        E.nearest_points(pts[:1])  # no other point
