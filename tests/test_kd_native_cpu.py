"""The native KD build (madi_kd_build, host code) against the numpy form of
the reference's build_kdtree (madi/ensemble.py:525-563): per node, the axis
of largest extent (first on ties) and the element of rank len//2 along it
(np.argpartition), ids in pre-order.  Bitwise equal on point sets whose split
coordinates are distinct (random and periodic-padded seeds, anisotropic
boxes that exercise every axis).  Where coordinates tie (periodic images
repeat their seed's coordinates on the unshifted axes) np.argpartition does
not fix which equal element it returns, and the native build orders equal
coordinates by point index: there the tree is checked as a valid,
deterministic split tree of the set."""

from __future__ import annotations

import numpy as np
import pytest

from custom_madi_amd.ensemble import build_kdtree, periodic_images


def _kd_numpy(points):
    pts = np.asarray(points, dtype=np.float64)
    n = len(pts)
    seed = np.empty(n, dtype=np.int32)
    axis = np.empty(n, dtype=np.int8)
    left = np.full(n, -1, dtype=np.int32)
    right = np.full(n, -1, dtype=np.int32)
    parent = np.full(n, -1, dtype=np.int32)
    next_id = 0
    stack = [(np.arange(n, dtype=np.int64), -1, -1)]
    while stack:
        idx, par, side = stack.pop()
        if len(idx) == 0:
            continue
        sub = pts[idx]
        ax = int(np.argmax(sub.max(axis=0) - sub.min(axis=0)))
        mid = len(idx) // 2
        arranged = idx[np.argpartition(sub[:, ax], mid)]
        node = next_id
        next_id += 1
        seed[node], axis[node], parent[node] = arranged[mid], ax, par
        if side == 0:
            left[par] = node
        elif side == 1:
            right[par] = node
        stack.append((arranged[mid + 1:], node, 1))
        stack.append((arranged[:mid], node, 0))
    return seed, axis, left, right, parent


def _same(points):
    got, want = build_kdtree(points), _kd_numpy(points)
    for g, w, name in zip(got, want, ("seed", "axis", "left", "right", "parent")):
        assert g.dtype == w.dtype and np.array_equal(g, w), name


@pytest.mark.parametrize("n", [0, 1, 2, 3, 7, 64, 1000, 20_000])
def test_native_kd_equals_numpy_form_on_random_points(n):
    _same(np.random.default_rng(n).uniform(0.0, 50.0, size=(n, 3)))


def test_native_kd_equals_numpy_form_on_anisotropic_sets():
    rng = np.random.default_rng(5)
    for scale in ((1.0, 5.0, 0.2), (0.3, 0.3, 9.0), (7.0, 1.0, 1.0)):
        _same(rng.uniform(0.0, 1.0, size=(5000, 3)) * np.asarray(scale))


def _periodic_padded():
    # seeds and their periodic images: an image repeats the seed's
    # coordinates on its unshifted axes, so this set is full of ties
    seeds = np.random.default_rng(6).uniform(0.0, 20.0, size=(800, 3))
    pts, _, _ = periodic_images(seeds, 20.0, 6.0)
    return pts


@pytest.mark.parametrize("points", ["random", "periodic"])
def test_native_kd_tree_is_a_valid_split_tree(points):
    pts = np.random.default_rng(9).uniform(0.0, 10.0, size=(3000, 3)) if points == "random" else _periodic_padded()
    seed, axis, left, right, parent = build_kdtree(pts)
    assert build_kdtree(pts)[0].tobytes() == seed.tobytes()      # deterministic
    assert sorted(seed.tolist()) == list(range(len(pts)))
    # every left descendant <= node <= every right descendant along the node's axis
    for node in range(0, len(pts), 97):
        v = pts[seed[node], axis[node]]
        for child, ok in ((left[node], lambda x: x <= v), (right[node], lambda x: x >= v)):
            stack = [child]
            while stack:
                c = stack.pop()
                if c < 0:
                    continue
                assert ok(pts[seed[c], axis[node]])
                assert parent[c] >= 0
                stack += [left[c], right[c]]
