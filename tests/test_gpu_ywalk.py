"""The FP32 walk with Y output (madi_walk_y_f32) followed by the FP64 column
reduce: the FP32 production path for column sets the fused walk cannot take
(more than 8 pairs or 16 b values, the universal grid among them).

* The Y walk is the fused FP32 walk: same labels (occupancy and work
  counters equal), Y sample 0 zero, and the reduce of its Y equals the fused
  walk's sums to FP32 phase rounding.
* On a 12-pair column set FP32 Y + reduce agrees with the FP64 walk + reduce
  on the same seeds within the production bound |dS| <= 2e-3.
* run_simulation_reduced in FP32 at the universal grid (1245 pairs x 25 b,
  128,000 steps of 1 us) on free water: every column against
  exp(-b D0) with the reference's free-water tolerance
  (tests/physics_audit/test_finite_delta_reference.py:35-61) and against the
  FP64 walk on the same seeds.
"""

from __future__ import annotations

import sys
from dataclasses import replace

import numpy as np
import pytest

from custom_madi_amd import engine, signal, walker_gpu
from custom_madi_amd.config import SimConfig
from custom_madi_amd.ensemble import DeviceGeometry, geometry_arrays
from tests import _geom

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class _Cfg:
    def __init__(self, steps, sph, G, sigma, ts):
        self.n_steps, self.steps_per_h, self.n_grid, self.sigma, self.ts = steps, sph, G, sigma, ts


@pytest.fixture(scope="module")
def cell_geom():
    ens = _geom.cellular(L=150.0, seed=9)
    return ens, DeviceGeometry(geometry_arrays(ens))


def _walk_setup():
    n, steps, sph, G = 65_536, 650, 50, 14
    sigma, ts, pp = np.sqrt(2 * 2.0 * 0.02), 0.02, 0.03
    return n, _Cfg(steps, sph, G, sigma, ts), pp


def test_y_walk_is_the_fused_walk(cell_geom):
    ens, dg = cell_geom
    n, cfg, pp = _walk_setup()
    delta, Deltas, bs = 2.0, (5.0, 8.0, 10.0), (0.0, 1000.0, 2500.0, 6000.0)
    pc = np.array([[engine.phase_coef(b, delta, D) for b in bs] for D in Deltas])
    spec = engine.PhaseSpec(j_delta=np.array([2, 2, 2]), j_Delta=np.array([5, 8, 10]), j_sum=np.array([7, 10, 12]),
                            pc=pc, dirs=engine.AXIS_DIRECTIONS)
    p = engine.walk_params(n, cfg, pp, 987)
    say("[ywalk] fused FP32 and Y FP32 walks, 65,536 walkers x 650 steps")
    fused = engine.walk_fused(dg, p, spec, precision=32)
    y = engine.walk_y_f32(dg, p)
    assert fused.n_escaped == 0 and fused.n_error == 0
    assert y.n_escaped == 0 and y.n_error == 0
    assert np.array_equal(y.occupancy, fused.occupancy)
    for k in ("classify", "safe_hit", "crossings", "rejects"):
        assert y.counters[k] == fused.counters[k], k
    Yh = y.Y.cpu().numpy()
    assert np.all(Yh[:, 0, :] == 0.0) and np.all(np.isfinite(Yh))
    assert np.all(np.abs(Yh[:, -1, :]) > 0.0)
    cols = [(k, b) for k in range(3) for b in range(4)]
    cs, sn = engine.reduce_columns(y.Y, [spec.j_delta[k] for k, _ in cols], [spec.j_Delta[k] for k, _ in cols],
                                   [spec.j_sum[k] for k, _ in cols], [spec.pc[k, b] for k, b in cols])
    fc = fused.cos.sum(axis=2).ravel()
    fs = fused.sin.sum(axis=2).ravel()
    # the fused walk forms dM and the phase in FP32 and uses the hardware
    # sin/cos; the reduce forms them in FP64 from the same FP32 Y samples
    say(f"[ywalk] max |cos sum diff| {np.abs(cs - fc).max():.3e} of {3 * n}")
    assert np.abs(cs - fc).max() <= 1e-5 * 3 * n
    assert np.abs(sn - fs).max() <= 1e-5 * 3 * n


def test_fp32_y_and_fp64_walks_agree_on_a_wide_column_set(cell_geom):
    ens, dg = cell_geom
    n, cfg, pp = _walk_setup()
    pairs = [(1.0, D) for D in (2.0, 3.0, 4.0, 5.0, 6.0, 7.0)] + [(2.0, D) for D in (3.0, 4.0, 5.0, 6.0, 8.0, 10.0)]
    bs = (0.0, 500.0, 1000.0, 2500.0, 6000.0)
    jd, jD, js, pcs = [], [], [], []
    for d, D in pairs:
        for b in bs:
            jd.append(int(d))
            jD.append(int(D))
            js.append(int(d + D))
            pcs.append(engine.phase_coef(b, d, D))
    assert max(js) < cfg.n_grid and len(pairs) > 8
    p = engine.walk_params(n, cfg, pp, 4242)
    say("[ywalk] 12-pair column set: FP32 Y + reduce against FP64 walk + reduce")
    y32 = engine.walk_y_f32(dg, p)
    c32, _ = engine.reduce_columns(y32.Y, jd, jD, js, pcs)
    assert y32.n_escaped == 0 and y32.n_error == 0
    del y32
    y64 = engine.walk_y(dg, p)
    assert not y64.escaped.any() and not y64.err.any()
    c64, _ = engine.reduce_columns(y64.Y, jd, jD, js, pcs)
    dS = np.abs(c32 - c64) / (3 * n)
    say(f"[ywalk] max |dS| FP32 vs FP64 = {dS.max():.3e}")
    assert dS.max() <= 2e-3


def test_universal_grid_free_water_fp32_drop_in():
    cfg = SimConfig(L=None, n_walkers=4096, n_ensembles=1, precision="fp32")
    cols = signal.build_columns(cfg)
    assert len(cols.phase_coef) == 31_125 and cfg.n_steps == 128_000
    say("[ywalk] universal grid, free water, run_simulation_reduced fp32 (4096 walkers x 128,000 steps)")
    r32 = walker_gpu.run_simulation_reduced(0.0, 0.0, 0.0, cols.j_delta, cols.j_Delta, cols.j_sum, cols.phase_coef,
                                            cfg=cfg, seed=11, verbose=False)
    assert r32.n_walkers == 4096 and r32.n_escaped == 0
    S = r32.cos_sum / r32.n_eff
    b = np.tile(np.asarray(cols.b_values, dtype=float), cols.n_pairs)
    expected = np.exp(-b * cfg.D0 * 1e-3)
    se = np.sqrt(np.maximum(1.0 - S ** 2, 0.0) / r32.n_eff)
    worst = np.max(np.abs(S - expected) - (5.0 * se + 0.012))
    say(f"[ywalk] worst column margin {worst:.3e}")
    assert worst <= 0.0
    say("[ywalk] the same entry in FP64")
    r64 = walker_gpu.run_simulation_reduced(0.0, 0.0, 0.0, cols.j_delta, cols.j_Delta, cols.j_sum, cols.phase_coef,
                                            cfg=replace(cfg, precision="fp64"), seed=11, verbose=False)
    dS = np.abs(S - r64.cos_sum / r64.n_eff)
    say(f"[ywalk] max |dS| FP32 vs FP64 over 31,125 columns = {dS.max():.3e}")
    assert dS.max() <= 2e-3
