"""HIP kernels vs the CPU oracle on identical inputs (run on an MI355X).

Bars: labels and RNG states bit-exact; FP64 walks within the reference's
golden tolerance rtol 5e-11 / atol 5e-12 (scripts/run_gpu_golden_check.py:27-28)
with identical inside traces; FP64 sums within 1e-10 relative; the FP32
production walk within |dS| <= 2e-3 and 3 Monte-Carlo standard errors of the
FP64 walk on the same seeds.
"""

from __future__ import annotations

import sys

import numpy as np
import pytest

from oracle import oracle as O
from custom_madi_amd import _hip, engine, workloads
from custom_madi_amd.ensemble import DeviceGeometry, geometry_arrays
from tests import _geom

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def cell_geom():
    ens = _geom.cellular(L=50.0, seed=3)
    arr = geometry_arrays(ens)
    return ens, arr, DeviceGeometry(arr), O.Geometry.from_arrays(arr)


@pytest.fixture(scope="module")
def big_geom():
    ens = _geom.cellular(L=150.0, seed=9)
    arr = geometry_arrays(ens)
    return ens, arr, DeviceGeometry(arr), O.Geometry.from_arrays(arr)


@pytest.fixture(scope="module")
def periodic_geom():
    ens = _geom.periodic(n_cells=64, seed=0)
    arr = geometry_arrays(ens)
    return ens, arr, DeviceGeometry(arr), O.Geometry.from_arrays(arr)


def test_library_abi_version():
    lib = _hip.load_library()
    assert lib.madi_abi_version() == 9
    info = _hip.device_info()
    say(f"device: {info}")
    assert info["cu_count"] > 0


def test_exact_classifier_bit_exact_vs_oracle_and_brute(cell_geom):
    ens, arr, dg, og = cell_geom
    pts = np.random.default_rng(20_260_902).uniform(0.0, ens.L, size=(20_000, 3))
    cell, inside, ok = dg.classify(pts)
    ocell, oinside, ook = og.classify(pts)
    assert ok.all() and ook.all()
    assert np.array_equal(cell, ocell)
    assert np.array_equal(inside, oinside)
    bcell, binside = O.classify_brute_c(ens.seeds, ens.annulus, pts[:4000])
    assert np.array_equal(cell[:4000], bcell) and np.array_equal(inside[:4000], binside)


def test_graph_classifier_equals_exact(cell_geom):
    ens, arr, dg, og = cell_geom
    pts = np.random.default_rng(77).uniform(0.0, ens.L, size=(50_000, 3))
    cell, inside, ok, cnt = dg.classify(pts, graph=True, counters=True)
    ecell, einside, eok = dg.classify(pts)
    say(f"graph classifier counters: {cnt}")
    assert ok.all()
    assert np.array_equal(inside, einside)
    assert np.array_equal(cell, ecell)


def test_periodic_classifiers_equal_tiling(periodic_geom):
    ens, arr, dg, og = periodic_geom
    pts = np.random.default_rng(4).uniform(-0.5 * ens.L, 1.5 * ens.L, size=(20_000, 3))
    cell, inside, ok = dg.classify(pts)
    gcell, ginside, gok = dg.classify(pts, graph=True)
    ocell, oinside, _ = og.classify(pts)
    assert ok.all() and gok.all()
    assert np.array_equal(cell, ocell) and np.array_equal(inside, oinside)
    assert np.array_equal(inside, ginside) and np.array_equal(cell, gcell)


def test_rng_states_match_oracle():
    for seed, first, n in ((0, 0, 1000), (20_260_901, 17, 513), (2 ** 31 - 1, 100_000, 64)):
        dev = engine.rng_states(seed, first, n)
        ref = O.rng_states(seed, first, n)
        assert np.array_equal(dev, ref), (seed, first)


def _walk_pair(ens, arr, dg, og, n, steps, sph, G, sigma, ts, pp, seed=0, stream=None):
    p = engine.walk_params(n, _Cfg(steps, sph, G, sigma, ts), pp, seed)
    out = engine.walk_y(dg, p, None if stream is None else _Stream(*stream))
    Y, itr, esc, err = O.walk(og, n, steps, sph, G, sigma, ts, pp, arr.source_lo, arr.source_hi, seed=seed,
                              stream=stream)
    return out, (Y, itr, esc, err)


class _Cfg:
    def __init__(self, steps, sph, G, sigma, ts):
        self.n_steps, self.steps_per_h, self.n_grid, self.sigma, self.ts = steps, sph, G, sigma, ts


class _Stream:
    def __init__(self, a, b, c):
        self.initial_positions, self.increments, self.acceptance_uniforms = a, b, c


def test_fp64_walk_injected_stream_matches_oracle(cell_geom):
    ens, arr, dg, og = cell_geom
    n, steps, sph, G = 256, 600, 50, 13
    sigma, ts, pp = np.sqrt(2 * 2.0 * 0.02), 0.02, 0.05
    rng = np.random.default_rng(5)
    stream = (rng.uniform(ens.source_lo, ens.source_hi, size=(n, 3)), rng.normal(0.0, sigma, size=(steps, n, 3)),
              rng.uniform(0.0, 1.0, size=(steps, n)))
    out, (Y, itr, esc, err) = _walk_pair(ens, arr, dg, og, n, steps, sph, G, sigma, ts, pp, stream=stream)
    assert not err.any() and not out.err.any()
    assert np.array_equal(out.escaped, esc)
    assert np.array_equal(out.inside_trace, itr)
    assert np.allclose(out.Y.cpu().numpy(), Y, rtol=5e-11, atol=5e-12)


def test_fp64_walk_device_rng_matches_oracle(cell_geom):
    ens, arr, dg, og = cell_geom
    n, steps, sph, G = 512, 1000, 50, 21
    sigma, ts, pp = np.sqrt(2 * 2.0 * 0.02), 0.02, 0.02
    out, (Y, itr, esc, err) = _walk_pair(ens, arr, dg, og, n, steps, sph, G, sigma, ts, pp, seed=104_729)
    assert not err.any() and not out.err.any()
    assert np.array_equal(out.escaped, esc)
    assert np.array_equal(out.inside_trace, itr)
    assert np.allclose(out.Y.cpu().numpy(), Y, rtol=5e-11, atol=5e-12)


def test_fp64_walk_periodic_matches_oracle(periodic_geom):
    ens, arr, dg, og = periodic_geom
    n, steps, sph, G = 256, 1500, 50, 31
    sigma, ts, pp = np.sqrt(2 * 2.0 * 0.02), 0.02, 0.01
    out, (Y, itr, esc, err) = _walk_pair(ens, arr, dg, og, n, steps, sph, G, sigma, ts, pp, seed=11)
    assert not out.err.any() and not out.escaped.any()
    assert np.array_equal(out.inside_trace, itr)
    assert np.allclose(out.Y.cpu().numpy(), Y, rtol=5e-11, atol=5e-12)


def test_reduce_matches_oracle():
    torch = _hip.require_gpu()
    rng = np.random.default_rng(2)
    Y = rng.normal(0.0, 30.0, size=(5000, 9, 3))
    jd, jD, js = np.array([1, 1, 2, 6]), np.array([4, 6, 5, 7]), np.array([5, 7, 7, 8])
    pc = np.array([0.0, 0.013, 0.029, 0.05])
    cs, sn = engine.reduce_columns(torch.from_numpy(Y).cuda(), jd, jD, js, pc)
    ocs, osn = O.reduce(Y, jd, jD, js, pc)
    assert np.allclose(cs, ocs, rtol=1e-11, atol=1e-9)
    assert np.allclose(sn, osn, rtol=1e-11, atol=1e-9)


def test_reduce_universal_grid_production_size():
    """The universal library's reduce at production size (50,000 walkers x 129
    Y samples x 31,125 columns; reference kernel madi/walker_gpu.py:658-695):
    every 61st column against the oracle's sequential sums, b = 0 columns
    exactly 3n / 0 (phase 0 everywhere), and two launches bitwise equal."""
    torch = _hip.require_gpu()
    Y, cols = workloads.universal_reduce()
    n = Y.shape[0]
    dY = torch.from_numpy(Y).cuda()
    cs, sn = engine.reduce_columns(dY, cols.j_delta, cols.j_Delta, cols.j_sum, cols.phase_coef)
    cs2, sn2 = engine.reduce_columns(dY, cols.j_delta, cols.j_Delta, cols.j_sum, cols.phase_coef)
    assert np.array_equal(cs, cs2) and np.array_equal(sn, sn2)
    sub = np.arange(0, len(cols.phase_coef), 61)
    ocs, osn = O.reduce(Y, cols.j_delta[sub], cols.j_Delta[sub], cols.j_sum[sub], cols.phase_coef[sub])
    # sums of 150,000 unit-bounded terms in different orders
    assert np.all(np.abs(cs[sub] - ocs) <= 1e-11 * np.maximum(np.abs(ocs), 1.0) + 1e-9)
    assert np.all(np.abs(sn[sub] - osn) <= 1e-11 * np.maximum(np.abs(osn), 1.0) + 1e-9)
    zero = cols.phase_coef == 0.0
    assert zero.sum() == cols.n_pairs
    assert np.all(cs[zero] == 3.0 * n) and np.all(sn[zero] == 0.0)


def test_reduce_sincos_elementwise():
    """One walker with Y[0] = 0 and Y[1] = (1, 0, 0), columns (jd, jD, js) =
    (1, 0, 0): column c's phase is exactly pc[c] on axis 0 and 0 on the others,
    so its sums are sin(pc) and 2 + cos(pc).  Against numpy's sin and cos over
    the fast range (to 2^20), across quadrant boundaries (multiples of pi/2),
    tiny arguments and the library path beyond 2^20."""
    torch = _hip.require_gpu()
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-60, 60, 200_000), rng.uniform(-2e5, 2e5, 20_000),
                         rng.uniform(-1e-6, 1e-6, 5_000), np.arange(-200, 201) * (np.pi / 2),
                         np.array([0.0, -0.0, 1e-300, 7.0e5, -3.3e6, 1e12, 1048576.0, 1048577.0, -1048576.5])])
    Y1 = np.zeros((1, 2, 3))
    Y1[0, 1, 0] = 1.0
    one, zero = np.ones(len(xs), dtype=np.int32), np.zeros(len(xs), dtype=np.int32)
    cs, sn = engine.reduce_columns(torch.from_numpy(Y1).cuda(), one, zero, zero, xs)
    ref_s, ref_c = np.sin(xs), np.cos(xs)
    fast = np.abs(xs) <= 2.0 ** 20
    ulps = np.abs(sn - ref_s) / np.spacing(np.maximum(np.abs(ref_s), 1e-300))
    assert ulps[fast].max() <= 2.0
    assert np.abs(sn - ref_s).max() <= 2e-16
    # cos carries the rounding of 2 + cos (half an ulp of 2..3)
    assert np.abs((cs - 2.0) - ref_c).max() <= 5e-16


def _spec(G, sph_ms=1.0, dirs=engine.AXIS_DIRECTIONS, bs=(0.0, 1000.0, 2500.0, 6000.0)):
    delta, Deltas = 2.0, (5.0, 8.0, 10.0)
    pc = np.array([[engine.phase_coef(b, delta, D) for b in bs] for D in Deltas])
    return engine.PhaseSpec(
        j_delta=np.array([2, 2, 2]), j_Delta=np.array([5, 8, 10]), j_sum=np.array([7, 10, 12]), pc=pc, dirs=dirs)


# _spec's b values followed by the 4b values that give each signal's Monte
# Carlo variance (Var cos(phi) = (1 + S(4b)) / 2 - S(b)^2, since the phase
# grows as sqrt(b)); idx / idx4: positions of b and of 4b
_B_VAR = (0.0, 1000.0, 2500.0, 6000.0, 4000.0, 10000.0, 24000.0)
_IDX, _IDX4 = np.array([0, 1, 2, 3]), np.array([0, 4, 5, 6])


def test_fused_fp64_equals_walk_then_reduce(big_geom):
    ens, arr, dg, og = big_geom
    n, steps, sph, G = 4096, 650, 50, 14
    sigma, ts, pp = np.sqrt(2 * 2.0 * 0.02), 0.02, 0.03
    cfg = _Cfg(steps, sph, G, sigma, ts)
    spec = _spec(G)
    p = engine.walk_params(n, cfg, pp, 321)
    fused = engine.walk_fused(dg, p, spec, precision=64)
    out = engine.walk_y(dg, p)
    cols = [(k, b) for k in range(3) for b in range(4)]
    jd = [spec.j_delta[k] for k, b in cols]
    jD = [spec.j_Delta[k] for k, b in cols]
    js = [spec.j_sum[k] for k, b in cols]
    pcs = [spec.pc[k, b] for k, b in cols]
    cs, sn = engine.reduce_columns(out.Y, jd, jD, js, pcs)
    fc = fused.cos.sum(axis=2).ravel()
    fs = fused.sin.sum(axis=2).ravel()
    assert fused.n_escaped == 0 and fused.n_error == 0
    assert np.array_equal(fused.occupancy, out.inside_trace.sum(axis=0))
    assert np.allclose(fc, cs, rtol=1e-10, atol=1e-8)
    assert np.allclose(fs, sn, rtol=1e-10, atol=1e-8)


@pytest.mark.parametrize("which", ["big_geom", "periodic_geom"])
def test_fp32_walk_matches_fp64_statistically(which, request):
    ens, arr, dg, og = request.getfixturevalue(which)
    n, steps, sph, G = 65_536, 650, 50, 14
    sigma, ts, pp = np.sqrt(2 * 2.0 * 0.02), 0.02, 0.03
    cfg = _Cfg(steps, sph, G, sigma, ts)
    spec = _spec(G)
    p = engine.walk_params(n, cfg, pp, 987)
    f64 = engine.walk_fused(dg, p, _spec(G, bs=_B_VAR), precision=64)
    f32 = engine.walk_fused(dg, p, spec, precision=32)
    say(f"fp32 counters: {f32.counters}")
    assert f32.n_escaped == 0 and f32.n_error == 0
    S64, se = workloads.signal_se(f64.cos, n, _IDX, _IDX4)
    S32 = f32.cos / n
    dS = np.abs(S64 - S32)
    say(f"max |dS| fp32 vs fp64 = {dS.max():.3e}, max |dS| / SE = {(dS / np.maximum(se, 1e-300)).max():.3f}")
    assert dS.max() <= 2e-3
    assert np.all(dS <= 3.0 * se)
    occ_rel = np.abs(f32.occupancy - f64.occupancy) / n
    assert occ_rel.max() < 0.01


def test_config0_entry_fp64_matches_oracle():
    """The CPU-runnable configuration end to end: periodic cube of 64 cells
    (seeds from default_rng(0)), rho = 6e5 / uL, V = 1 pL (v_i = 0.6),
    k_io = 20 / s (p_p by Eq. 5), D = 2.0, dt = 0.05 ms x 920 steps, 20,000
    walkers, FP64; 4 Delta x 5 b x 24 directions.  The fused FP64 entry
    (walk + phase sums in one launch) against the oracle's walk with the
    directional sums formed in numpy from its Y."""
    from custom_madi_amd.config import SimConfig
    from custom_madi_amd.ensemble import create_periodic_ensemble
    from custom_madi_amd.walker_gpu import kio_to_pp
    cfg = SimConfig(D0=2.0, ts=0.05, T_max_ms=46.0, h_ms=1.0, n_walkers=20_000, small_deltas=[6.0],
                    big_deltas=[15.0, 25.0, 30.0, 40.0], b_values=[0.0, 1000.0, 2500.0, 4000.0, 6000.0],
                    periodic=True, precision="fp64")
    assert cfg.n_steps == 920 and cfg.steps_per_h == 20
    ens = create_periodic_ensemble(64, 6.0e5, 0.6, cfg, seed=0)
    pp = kio_to_pp(20.0, ens.mean_AV, cfg)
    dirs = np.random.default_rng(1).normal(size=(24, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    Deltas = [15.0, 25.0, 30.0, 40.0]
    pc = np.array([[engine.phase_coef(b, 6.0, D) for b in cfg.b_values] for D in Deltas])
    spec = engine.PhaseSpec(j_delta=np.array([6] * 4), j_Delta=np.array([int(D) for D in Deltas]),
                            j_sum=np.array([int(D + 6.0) for D in Deltas]), pc=pc, dirs=dirs)
    N = cfg.n_walkers
    say(f"config 0: L={ens.L:.3f} um, pp={pp:.6g}, {N} walkers x {cfg.n_steps} steps")
    p = engine.walk_params(N, cfg, pp, 0)
    fused = engine.walk_fused(ens.device_geometry(), p, spec, precision=64)
    arr = geometry_arrays(ens)
    Y, itr, esc, err = O.walk(O.Geometry.from_arrays(arr), N, cfg.n_steps, cfg.steps_per_h, cfg.n_grid, cfg.sigma,
                              cfg.ts, pp, arr.source_lo, arr.source_hi, seed=0)
    assert fused.n_escaped == 0 and fused.n_error == 0 and not esc.any() and not err.any()
    assert np.array_equal(fused.occupancy, itr.sum(axis=0))
    ocos = np.empty((4, 5, 24))
    osin = np.empty((4, 5, 24))
    for k in range(4):
        dM = Y[:, spec.j_delta[k]] + Y[:, spec.j_Delta[k]] - Y[:, spec.j_sum[k]]
        proj = dM @ dirs.T
        for b in range(5):
            ph = pc[k, b] * proj
            ocos[k, b] = np.cos(ph).sum(axis=0)
            osin[k, b] = np.sin(ph).sum(axis=0)
    assert np.allclose(fused.cos, ocos, rtol=1e-9, atol=1e-7)
    assert np.allclose(fused.sin, osin, rtol=1e-9, atol=1e-7)
    S = fused.cos / N
    say(f"config 0: S(b=6000, Delta=40) mean over directions {S[3, 4].mean():.5f}")
    assert 0.0 < S[3, 4].mean() < S[0, 1].mean() < 1.0
