"""Pins of the CPU oracle against what the reference and the mathematics fix.

Each oracle function is checked against an independent derivation chosen so
that a plausible slip (a dropped term, a wrong index or shift, a transposed
operand) fails:
  RNG       pure-Python xoroshiro128+ (numba-src:45-125) and a GF(2) matrix
            derivation of the 2^64 jump (T^(2^64) by 64 squarings);
  classify  all-seed brute force (reference test
            tests/physics_audit/test_geometry_diagnostics.py:143-152), the
            alpha = 0 two-nearest limit (:188-207), the 27-image tiling for
            periodic boxes;
  walk      the trapezoid Y closed form, bitwise
            (tests/physics_audit/test_finite_delta_reference.py:64-99),
            free-water Stejskal-Tanner (:35-61), the fatal escape
            (tests/physics_audit/test_p0_remediation.py:66-84), and the numpy
            restatement of the reference CPU walk with cKDTree labels;
  reduce    the numpy restatement of _reduce_cpu.
"""

from __future__ import annotations

import numpy as np
import pytest
from scipy.spatial import cKDTree

from oracle import oracle as O
from custom_madi_amd.ensemble import geometry_arrays
from tests import _geom

M64 = (1 << 64) - 1


# ---------------------------------------------------------------------- RNG

def test_stream_matches_pure_python_xoroshiro():
    for seed in (0, 1, 20_260_901, 2 ** 31 - 1):
        py = O.Xoroshiro128p(seed=seed)
        st = O.lib().og_seed_state(seed)
        assert st.s0 == st.s1 == py.s0  # splitmix64 single output in both words
        raw = O.draws(st.s0, st.s1, 3, 200)
        assert [int(v) for v in raw] == [py.next() for _ in range(200)]


def _matrix_of_next():
    """Transition matrix of the xoroshiro128+ state update, column j = image of bit j."""
    cols = []
    for j in range(128):
        s = O.Xoroshiro128p(state=((1 << j) if j < 64 else 0, (1 << (j - 64)) if j >= 64 else 0))
        s.step_linear()
        cols.append(s.s0 | (s.s1 << 64))
    return cols


def _mat_vec(cols, v):
    out = 0
    j = 0
    while v:
        if v & 1:
            out ^= cols[j]
        v >>= 1
        j += 1
    return out


def _mat_mul(a, b):
    return [_mat_vec(a, c) for c in b]


def test_jump_equals_2_to_the_64_steps_by_gf2_squaring():
    T = _matrix_of_next()
    P = T
    for _ in range(64):
        P = _mat_mul(P, P)
    rng = np.random.default_rng(7)
    for _ in range(4):
        s0, s1 = (int(x) for x in rng.integers(0, 2 ** 63, size=2))
        st = O.OState(s0, s1)
        O.lib().og_jump(st)
        v = _mat_vec(P, s0 | (s1 << 64))
        assert (st.s0, st.s1) == (v & M64, v >> 64)


def test_walker_states_are_successive_jumps():
    states = O.rng_states(12345, 3, 5)
    st = O.lib().og_seed_state(12345)
    for _ in range(3):
        O.lib().og_jump(st)
    for i in range(5):
        if i:
            O.lib().og_jump(st)
        assert (int(states[i, 0]), int(states[i, 1])) == (st.s0, st.s1)


def test_uniform_and_normal_follow_the_numba_definitions():
    st = O.lib().og_seed_state(99)
    raw = [int(v) for v in O.draws(st.s0, st.s1, 3, 64)]
    u64 = O.draws(st.s0, st.s1, 0, 64)
    u32 = O.draws(st.s0, st.s1, 1, 64)
    z = O.draws(st.s0, st.s1, 2, 32)
    ref_u = np.array([(x >> 11) * 2.0 ** -53 for x in raw])
    assert np.array_equal(u64, ref_u)
    assert np.array_equal(u32, ref_u.astype(np.float32).astype(np.float64))
    f = ref_u.astype(np.float32)
    u1, u2 = f[0::2], f[1::2]
    lg = np.log(u1.astype(np.float64)).astype(np.float32).astype(np.float64)
    ref_z = np.sqrt(-2.0 * lg) * np.cos(6.283185307179586 * u2.astype(np.float64))
    assert np.allclose(z, ref_z, rtol=1e-15, atol=1e-15)


def test_normal_draws_are_standard_normal():
    """The oracle's normals (numba's Box-Muller on float32 uniforms) pinned by
    what the mathematics fixes, not by their formula: 2e6 draws have mean 0
    and variance 1 within 5 standard errors, fourth moment 3 within 5 SE,
    pass a Kolmogorov-Smirnov test against N(0, 1), and successive draws are
    uncorrelated.  A dropped factor (sqrt(-log u) instead of sqrt(-2 log u)),
    log10 for log, or a missing square root each fails (checked on the CPU).
    The angle's period is not visible here (cos of a uniform angle on [0, pi)
    is distributed as on [0, 2 pi)); the formula test above pins it."""
    from scipy import stats
    st = O.lib().og_seed_state(2024)
    z = np.asarray(O.draws(st.s0, st.s1, 2, 2_000_000), dtype=np.float64)
    n = z.size
    assert abs(z.mean()) <= 5.0 / np.sqrt(n)
    assert abs(z.var() - 1.0) <= 5.0 * np.sqrt(2.0 / n)
    assert abs(np.mean(z ** 4) - 3.0) <= 5.0 * np.sqrt(96.0 / n)   # Var(z^4) = 105 - 9
    assert stats.kstest(z, "norm").pvalue > 1e-4
    assert abs(np.corrcoef(z[:-1], z[1:])[0, 1]) <= 5.0 / np.sqrt(n)


# ----------------------------------------------------------------- classify

def test_kd_classifier_matches_all_seed_brute_force():
    ens = _geom.cellular(L=50.0, seed=3)
    g = O.Geometry.from_arrays(geometry_arrays(ens))
    pts = np.random.default_rng(20_260_902).uniform(0.0, ens.L, size=(2000, 3))
    cell, inside, ok = g.classify(pts)
    bcell, binside = O.classify_brute(ens.seeds, ens.annulus, pts)
    assert ok.all()
    assert np.array_equal(cell, bcell)
    assert np.array_equal(inside, binside)
    assert 0.3 < inside.mean() < 0.9
    c2, i2 = O.classify_brute_c(ens.seeds, ens.annulus, pts)
    assert np.array_equal(c2, bcell) and np.array_equal(i2, binside)


def test_alpha_zero_collapses_to_voronoi_two_nearest():
    ens = _geom.cellular(L=40.0, seed=5, alpha=0.0)
    ens.annulus[:] = 0.0
    g = O.Geometry.from_arrays(geometry_arrays(ens))
    pts = np.random.default_rng(11).uniform(0.0, ens.L, size=(3000, 3))
    cell, inside, _ = g.classify(pts)
    d, idx = cKDTree(ens.seeds).query(pts, k=2)
    sep = np.linalg.norm(ens.seeds[idx[:, 1]] - ens.seeds[idx[:, 0]], axis=1)
    two_inside = (d[:, 1] ** 2 - d[:, 0] ** 2) / (2.0 * sep) >= 0.0
    assert np.array_equal(cell, idx[:, 0])
    assert np.array_equal(inside, two_inside)


def test_periodic_padding_equals_27_image_tiling():
    ens = _geom.periodic(n_cells=64, seed=0)
    arr = geometry_arrays(ens)
    g = O.Geometry.from_arrays(arr)
    pts = np.random.default_rng(4).uniform(-0.5 * ens.L, 1.5 * ens.L, size=(3000, 3))
    cell, inside, ok = g.classify(pts)
    bcell, binside = O.periodic_brute(ens.seeds, ens.annulus, ens.L, pts)
    assert ok.all()
    assert np.array_equal(cell, bcell)
    assert np.array_equal(inside, binside)


def test_ckdtree_restatement_agrees_with_brute_force():
    ens = _geom.cellular(L=40.0, seed=8)
    pts = np.random.default_rng(12).uniform(0.0, ens.L, size=(800, 3))
    c1, i1 = O.classify_ckdtree(ens.seeds, ens.annulus, pts)
    c2, i2 = O.classify_brute(ens.seeds, ens.annulus, pts)
    assert np.array_equal(c1, c2) and np.array_equal(i1, i2)


# --------------------------------------------------------------------- walk

def _free_water_geometry(L):
    from custom_madi_amd.config import SimConfig
    from custom_madi_amd.ensemble import create_dummy_ensemble
    cfg = SimConfig(L=L, T_max_ms=2.0, small_deltas=[1.0], big_deltas=[1.0], b_values=[0.0])
    return O.Geometry.from_arrays(geometry_arrays(create_dummy_ensemble(cfg)))


def test_Y_is_the_microsecond_trapezoid_bitwise():
    steps, ts, L = 2000, 1e-3, 200.0
    t = np.arange(steps) * ts
    inc = np.zeros((steps, 1, 3))
    inc[:, 0, 0] = 4.0e-3 + 2.0e-3 * np.sin(2.0 * np.pi * t / 0.37)
    inc[:, 0, 1] = -1.0e-3 + 1.0e-3 * np.cos(2.0 * np.pi * t / 0.23)
    x0 = np.array([[100.0, 100.0, 100.0]])
    Y, itr, esc, err = O.walk(_free_water_geometry(L), 1, steps, 1000, 3, 0.0, ts, 0.0, 0.0, L,
                              stream=(x0, inc, np.zeros((steps, 1))))
    pos = x0[0].copy()
    prev = pos - L / 2.0
    integral = np.zeros(3)
    expected = [np.zeros(3)]
    for k, d in enumerate(inc[:, 0, :], start=1):
        pos += d
        cur = pos - L / 2.0
        integral += 0.5 * (prev + cur) * ts
        prev = cur
        if k % 1000 == 0:
            expected.append(integral.copy())
    assert esc[0] == 0 and err[0] == 0
    assert np.array_equal(Y[0], np.asarray(expected))


def test_fatal_escape_is_flagged():
    steps = 2000
    inc = np.zeros((steps, 1, 3))
    inc[0, 0, 0] = 2.0
    Y, itr, esc, err = O.walk(_free_water_geometry(2.0), 1, steps, 1000, 3, 0.0, 1e-3, 0.0, 0.0, 2.0,
                              stream=(np.array([[1.0, 1.0, 1.0]]), inc, np.zeros((steps, 1))))
    assert esc[0] == 1 and err[0] == 0


def _finite_signal(Y, jd, jD, js, b, delta, Delta):
    from custom_madi_amd.engine import phase_coef
    ph = phase_coef(b, delta, Delta) * (Y[:, jd, :] + Y[:, jD, :] - Y[:, js, :])
    s = np.cos(ph).ravel()
    return s.mean(), s.std(ddof=1) / np.sqrt(s.size)


def test_free_water_oracle_walk_matches_stejskal_tanner():
    D0, ts, T = 3.0, 1e-3, 30.0
    steps, sph = int(T / ts), int(1.0 / ts)
    L = 20.0 * np.sqrt(2.0 * D0 * T)
    Y, itr, esc, err = O.walk(_free_water_geometry(L), 512, steps, sph, int(T) + 1, np.sqrt(2 * D0 * ts), ts, 0.0,
                              0.3 * L, 0.7 * L, seed=20_260_901)
    assert not esc.any()
    for delta, Delta in ((1.0, 20.0), (5.0, 20.0), (10.0, 20.0), (7.0, 15.0)):
        for b in (0.0, 500.0, 1000.0, 2000.0, 4000.0):
            obs, se = _finite_signal(Y, int(delta), int(Delta), int(delta + Delta), b, delta, Delta)
            assert abs(obs - np.exp(-b * D0 * 1e-3)) <= 5.0 * se + 0.012, (delta, Delta, b, obs)


def test_oracle_walk_matches_numpy_reference_cpu_walk_on_injected_stream():
    ens = _geom.cellular(L=40.0, seed=21)
    arr = geometry_arrays(ens)
    g = O.Geometry.from_arrays(arr)
    n, steps, sph, G = 96, 400, 50, 9
    sigma, ts, pp = np.sqrt(2 * 2.0 * 0.02), 0.02, 0.05
    rng = np.random.default_rng(5)
    x0 = rng.uniform(ens.source_lo, ens.source_hi, size=(n, 3))
    inc = rng.normal(0.0, sigma, size=(steps, n, 3))
    u = rng.uniform(0.0, 1.0, size=(steps, n))
    Y, itr, esc, err = O.walk(g, n, steps, sph, G, sigma, ts, pp, ens.source_lo, ens.source_hi, stream=(x0, inc, u))
    Yr, itr_r, esc_r = O.walk_cpu_reference(ens.seeds, ens.annulus, ens.L, ens.source_lo, ens.source_hi, n, steps,
                                            sph, G, sigma, ts, pp, stream=(x0, inc, u))
    assert not err.any()
    assert np.array_equal(esc, esc_r)
    assert np.array_equal(itr, itr_r)
    assert np.allclose(Y, Yr, rtol=5e-11, atol=5e-12)
    assert 0 < itr.mean() < 1


def test_oracle_periodic_walk_matches_numpy_reference_walk_on_the_image_tiling():
    """The numpy walk over the image-padded seeds (the reference's CPU
    algorithm seeing the periodic tiling) equals the C oracle's periodic walk."""
    ens = _geom.periodic(n_cells=64, seed=3)
    arr = geometry_arrays(ens)
    g = O.Geometry.from_arrays(arr)
    n, steps, sph, G = 64, 500, 50, 11
    sigma, ts, pp = np.sqrt(2 * 2.0 * 0.02), 0.02, 0.05
    rng = np.random.default_rng(8)
    x0 = rng.uniform(0.0, ens.L, size=(n, 3))
    inc = rng.normal(0.0, sigma, size=(steps, n, 3))
    u = rng.uniform(0.0, 1.0, size=(steps, n))
    Y, itr, esc, err = O.walk(g, n, steps, sph, G, sigma, ts, pp, 0.0, ens.L, stream=(x0, inc, u))
    Yr, itr_r, esc_r = O.walk_cpu_reference(arr.kd_xyz, arr.kd_annulus, ens.L, 0.0, ens.L, n, steps, sph, G, sigma,
                                            ts, pp, stream=(x0, inc, u), label=arr.kd_label)
    assert not err.any() and not esc.any() and not esc_r.any()
    assert np.array_equal(itr, itr_r)
    assert np.allclose(Y, Yr, rtol=5e-11, atol=5e-12)
    # proposals cross the box faces, so the wrap is exercised
    path = x0[None] + np.cumsum(inc, axis=0)
    assert np.any((path < 0.0) | (path >= ens.L)) and 0 < itr.mean() < 1


def test_oracle_reduce_matches_numpy_restatement():
    rng = np.random.default_rng(2)
    Y = rng.normal(0.0, 30.0, size=(300, 9, 3))
    jd, jD, js = np.array([1, 1, 2]), np.array([4, 6, 5]), np.array([5, 7, 7])
    pc = np.array([0.0, 0.013, 0.029])
    a = O.reduce(Y, jd, jD, js, pc)
    b = O.reduce_reference(Y, jd, jD, js, pc)
    assert np.allclose(a[0], b[0], rtol=1e-12) and np.allclose(a[1], b[1], rtol=1e-12, atol=1e-10)
