"""The plane tier of the FP32 walk, checked against brute force on the CPU.

A tier built by a full classification at x answers later proposals q near x
without classifying them (csrc/prod.hip: tier_make / tier_query).  This test
restates both in FP64 numpy (exact arithmetic, so without the FP32 rounding
slack, and with the exact fourth margin where the kernel uses a lower bound
of it) and checks every answer the tier gives against the all-seed SI Eq. S2
labels of the 27-image tiled periodic box (oracle.periodic_brute).  It pins
the planes and radii: an inside answer trusted up to m4 instead of
m4 - alpha, a third plane left out of the answers, a radius from the fifth
margin, an outside ball not bounded by alpha_min (plus the distance to any
third Voronoi cell), or a k-side answer without the k scan would each give
wrong labels here.
"""

from __future__ import annotations

import numpy as np
import pytest

from custom_madi_amd.ensemble import geometry_arrays
from oracle import oracle as O
from tests import _geom

CAP_WIDTH = 3.0


def _margins(arr, a, p):
    lo, hi = arr.nb_offsets[a], arr.nb_offsets[a + 1]
    e = arr.nb_vec[lo:hi]
    ln = np.linalg.norm(e, axis=1)
    u = e / ln[:, None]
    return 0.5 * ln - u @ p, u, 0.5 * ln, lo


def _second(m):
    s = np.sort(m)
    return s[0], (s[1] if len(s) > 1 else np.inf), int(np.argmin(m))


def _locate(arr, a, p):
    """Greedy walk to the nearest seed (exact), returning the anchor and p relative to it."""
    for _ in range(512):
        m, _, _, lo = _margins(arr, a, p)
        k = int(np.argmin(m))
        if m[k] >= 0:
            return a, p
        p = p - arr.nb_vec[lo + k]
        a = int(arr.nb_ids[lo + k])
    raise AssertionError("greedy walk did not settle")


def tier_make(arr, a, x, cap):
    """The tier built at x (relative to anchor a, a the nearest seed): the
    three planes of smallest margin (m1, m2, m3) and the fourth margin m4."""
    m, u, h, lo = _margins(arr, a, x)
    order = np.argsort(m, kind="stable")
    m1, c = m[order[0]], int(order[0])
    m2 = min(m[order[1]], cap) if len(m) > 1 else cap
    c2 = int(order[1]) if len(m) > 1 else None
    c3 = int(order[2]) if len(m) > 2 else None
    m4 = min(m[order[3]], cap) if len(m) > 3 else cap
    aa = arr.cell_annulus[a]
    amin = float(arr.cell_annulus.min())
    inside = m1 >= aa
    t = dict(c=x.copy(), a=a, alpha_a=aa)
    if m1 >= cap:
        t.update(u=np.zeros(3), h=cap, u2=np.zeros(3), h2=cap, u3=np.zeros(3), h3=cap, k=a, alpha_k=0.0,
                 ra_in=cap - max(aa, 0), ra_out=cap - 0.0, rk_in=0.0, rk_out=0.0, rout=0.0)
        return t, inside
    k = int(arr.nb_ids[lo + c])
    ak = arr.cell_annulus[k]
    u2, h2 = (u[c2], h[c2]) if c2 is not None else (np.zeros(3), cap)
    u3, h3 = (u[c3], h[c3]) if c3 is not None else (np.zeros(3), cap)
    t.update(u=u[c], h=h[c], u2=u2, h2=h2, u3=u3, h3=h3, k=k, alpha_k=ak, ra_in=m4 - max(aa, 0.0), ra_out=m4)
    if inside:
        t.update(rk_in=0.0, rk_out=0.0, rout=0.0)
    else:
        xk = x - arr.nb_vec[lo + c]
        n = _margins(arr, k, xk)[0]
        n2 = min(_second(n)[1], cap)
        t.update(rk_in=n2 - max(ak, 0.0), rk_out=n2, rout=min(aa - m1, ak + m1, amin + max(min(m2, n2), 0.0)))
    return t, inside


def tier_query(t, q):
    """(certified, cell, inside) as csrc/prod.hip:tier_query, without rounding slack."""
    d = np.linalg.norm(q - t["c"])
    if d < max(t["rout"], 0.0):
        return True, None, False
    s = t["h"] - q @ t["u"]
    s23 = min(t["h2"] - q @ t["u2"], t["h3"] - q @ t["u3"])
    if s >= 0 and s23 >= 0:
        ins = min(s, s23) >= t["alpha_a"]
        ok = d < max(t["ra_in"] if ins else t["ra_out"], 0.0)
        return ok, t["a"], ins
    if s >= 0:
        return False, None, None  # across the second or third plane: no certificate
    r = -s
    ins = r >= t["alpha_k"]
    ok = d < max(t["rk_in"] if ins else t["rk_out"], 0.0)
    return ok, t["k"], ins


@pytest.mark.parametrize("spread", [False, True])
def test_every_tier_answer_equals_the_brute_force_label(spread):
    ens = _geom.periodic(n_cells=80, seed=3)
    if spread:
        # annuli differing from cell to cell, so alpha_min bounds the outside ball
        ens.annulus = ens.annulus * np.random.default_rng(4).uniform(0.15, 1.0, size=len(ens.annulus))
    arr = geometry_arrays(ens, with_kd=False)
    cap = float(arr.cell_annulus.max()) + CAP_WIDTH
    rng = np.random.default_rng(17)

    def anchored(x_world):
        a0 = int(np.argmin(np.linalg.norm(((arr.cell_xyz - x_world + ens.L / 2) % ens.L) - ens.L / 2, axis=1)))
        return _locate(arr, a0, ((x_world - arr.cell_xyz[a0] + ens.L / 2) % ens.L) - ens.L / 2)

    # half the centres anywhere, half in the gap near a Voronoi edge (second
    # margin below 1 um), where third cells and the alpha_min bound matter
    pool = rng.uniform(0, ens.L, size=(3000, 3))
    near_edge = []
    for x_world in pool:
        a, x = anchored(x_world)
        m = np.sort(_margins(arr, a, x)[0])
        if m[0] < arr.cell_annulus[a] and m[1] < 1.0:
            near_edge.append(x_world)
        if len(near_edge) == 300:
            break
    centres = np.concatenate([pool[-300:], np.asarray(near_edge)])
    c_lab, c_in = O.periodic_brute(ens.seeds, ens.annulus, ens.L, centres)
    queries, expect, kinds = [], [], []
    n_out = 0
    for x_world, cl, ci in zip(centres, c_lab, c_in):
        a, x = anchored(x_world)
        t, inside = tier_make(arr, a, x, cap)
        assert inside == bool(ci) and (not inside or a == int(cl))
        n_out += not inside
        rmax = max(t["ra_out"], t["rk_out"], t["rout"], 0.5) * 1.3
        dirs = rng.normal(size=(60, 3))
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
        qs = x + dirs * rng.uniform(0, rmax, size=(60, 1))
        if t["rout"] > 0:
            # and uniformly inside the outside ball, whose third-cell bound
            # binds near Voronoi edges
            b = rng.normal(size=(40, 3))
            b *= (t["rout"] * rng.uniform(0, 1, size=(40, 1)) ** (1 / 3)) / np.linalg.norm(b, axis=1, keepdims=True)
            qs = np.concatenate([qs, x + b])
        for q in qs:
            ok, cell, ins = tier_query(t, q)
            if ok:
                queries.append((arr.cell_xyz[a] + q) % ens.L)
                expect.append((cell, ins))
                kinds.append("ball" if cell is None else ("k" if cell == t["k"] and cell != a else "a"))
    assert n_out > 60  # centres in the gap exercise the k side and the outside ball
    lab, ins = O.periodic_brute(ens.seeds, ens.annulus, ens.L, np.asarray(queries))
    kinds = np.asarray(kinds)
    for kind in ("a", "k", "ball"):
        assert (kinds == kind).sum() > 200, kind
    for (cell, e_in), l, i in zip(expect, lab, ins):
        assert bool(i) == e_in
        if e_in:
            assert int(l) == cell


def test_outside_ball_third_cell_term_is_tight():
    """Near-cubic cells: a gap point x of cell a, 0.2 from the a->k face (m1) and 0.3 from the a->j face
    (m2), with j's annulus (0.1) the smallest.  x lies farther than alpha_j from j's face along the a->k
    plane, so the nearest contracted cell is j's, at exactly alpha_j + m2 = 0.4 (alpha_a - m1 = 0.7,
    alpha_k + m1 = 1.1): the tier's ball reaches it but not beyond, so a third-cell bound taken any larger
    would certify points inside cell j."""
    n, h = 4, 2.0
    ens = _geom.lattice(n=n, h=h, jitter=1e-4, annulus=0.9, seed=1)
    idx = lambda i, j, k: (i * n + j) * n + k  # noqa: E731
    a, k_cell, j_cell = idx(1, 1, 1), idx(2, 1, 1), idx(1, 2, 1)
    ens.annulus[j_cell] = 0.1
    arr = geometry_arrays(ens, with_kd=False)
    cap = float(arr.cell_annulus.max()) + CAP_WIDTH
    x = np.array([0.5 * h - 0.2, 0.5 * h - 0.3, 0.0])  # relative to seed a
    a2, x2 = _locate(arr, a, x)
    assert a2 == a
    t, inside = tier_make(arr, a, x, cap)
    assert not inside and t["k"] == k_cell
    assert abs(t["rout"] - 0.4) < 2e-3
    rng = np.random.default_rng(5)
    d = rng.normal(size=(4000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    for radius, expect_inside in ((t["rout"] * (1 - 1e-6), False), (t["rout"] + 0.02, True)):
        pts = (arr.cell_xyz[a] + x + d * radius) % ens.L
        lab, ins = O.periodic_brute(ens.seeds, ens.annulus, ens.L, pts)
        assert bool(np.any(ins)) == expect_inside
        if expect_inside:
            assert set(np.asarray(lab)[np.asarray(ins).astype(bool)].tolist()) == {j_cell}


def tier_leaves(t, q, L):
    """csrc/prod.hip:tier_exit without rounding slack: with k != a, the smallest margin of planes 1-3 (on k's side
    plane 1 alone) puts q outside a's contracted cell (a missing plane is (0, cap) and never does), and
    |q|_1 < 0.49 L keeps a's own seed nearer to q than any periodic image of it."""
    if t["k"] == t["a"]:
        return False
    s = t["h"] - q @ t["u"]
    sm = min(s, t["h2"] - q @ t["u2"], t["h3"] - q @ t["u3"]) if s >= 0 else s  # on k's side plane 1 alone
    return bool(sm < t["alpha_a"] and np.abs(q).sum() < 0.49 * L)


def _exit_cases(ens, n_centres, reach, seed):
    """Inside centres x of random cells and proposals q = x + d (|d| uniform up to reach): the certified exits
    and the brute-force labels of those q."""
    arr = geometry_arrays(ens, with_kd=False)
    cap = float(arr.cell_annulus.max()) + CAP_WIDTH
    rng = np.random.default_rng(seed)
    pts, cells, far, n_exit = [], [], 0, 0
    for x_world in rng.uniform(0, ens.L, size=(20 * n_centres, 3)):
        a0 = int(np.argmin(np.linalg.norm(((arr.cell_xyz - x_world + ens.L / 2) % ens.L) - ens.L / 2, axis=1)))
        a, x = _locate(arr, a0, ((x_world - arr.cell_xyz[a0] + ens.L / 2) % ens.L) - ens.L / 2)
        t, inside = tier_make(arr, a, x, cap)
        if not inside:
            continue
        d = rng.normal(size=(80, 3))
        d *= (reach * rng.uniform(0, 1, size=(80, 1))) / np.linalg.norm(d, axis=1, keepdims=True)
        for q in x + d:
            m = _margins(arr, a, q)[0]
            n_exit += bool(m.min() < t["alpha_a"])
            if tier_leaves(t, q, ens.L):
                pts.append((arr.cell_xyz[a] + q) % ens.L)
                cells.append(a)
                far += bool(np.linalg.norm(q) > 0.3 * ens.L)
        n_centres -= 1
        if n_centres == 0:
            break
    lab, ins = O.periodic_brute(ens.seeds, ens.annulus, ens.L, np.asarray(pts))
    return np.asarray(cells), np.asarray(lab), np.asarray(ins).astype(bool), far, n_exit


def test_certified_exits_leave_the_contracted_cell():
    """Every proposal the tier certifies as an exit of an inside walker's cell a lies in another cell or outside
    a's contracted cell (brute-force labels of the 27-image box), so a step rejected for both crossing kinds
    needs no classification.  Most geometric exits near the centre are certified."""
    ens = _geom.periodic(n_cells=80, seed=5)
    cells, lab, ins, _, n_exit = _exit_cases(ens, 400, 2.0, seed=23)
    assert len(cells) > 2000 and len(cells) > 0.5 * n_exit
    assert not np.any((lab == cells) & ins)


def test_certified_exits_exclude_periodic_images_of_the_cell():
    """A 2 x 2 x 2 lattice box (L = 4): proposals reaching a periodic image of the walker's own cell carry its
    index, so an exit certificate there would skip a move with no crossing.  The |q| < 0.49 L bound excludes
    them; certified exits still include points well away from the centre."""
    ens = _geom.lattice(n=2, h=2.0, jitter=1e-3, annulus=0.3, seed=2)
    cells, lab, ins, far, _ = _exit_cases(ens, 60, 1.6 * ens.L, seed=29)
    assert len(cells) > 200 and far > 40
    assert not np.any((lab == cells) & ins)
