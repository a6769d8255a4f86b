"""The FP32 walk's labels against the exact FP64 KD classifier, on the GPU.

madi_walk_fused_verify runs the production split walk and checks every label
it takes -- from the plane tier, from a certified exit, from the classifier
wave -- with the exact classifier of madi_classify on the same represented
point.  Every check must agree (the reference's own cache proof is the same
kind of replay: scripts/validate_exact_classifier_cache.py:73-140), and the
verified walk must give the production walk's results bit for bit.

Geometries: the C1 entry at full size (1,000 cells, 1e6 walkers x 2,300
steps); entries at v_i 0.5 and 0.95 whose annuli are short next to a step,
so the exit certificate is compiled in and used; the sweep's v_i = 0.994
box at dt 0.01 ms; and the C1 geometry with the tier's cap at 0.25 um and
20 um beyond the largest annulus (radii that change which steps need a full
classification).
"""

from __future__ import annotations

import sys

import numpy as np
import pytest

from custom_madi_amd import engine, workloads
from custom_madi_amd.config import SimConfig
from custom_madi_amd.ensemble import DeviceGeometry, create_periodic_ensemble, geometry_arrays
from custom_madi_amd.walker_gpu import kio_to_pp

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _verify(dg, params, spec, what, expect_exits=False):
    ref = engine.walk_fused(dg, params, spec, precision=32)
    res, v = engine.walk_fused_verify(dg, params, spec)
    say(f"[verify] {what}: {v}")
    assert v["exact_failed"] == 0
    assert v["tier_mismatch"] == 0 and v["exit_mismatch"] == 0 and v["classifier_mismatch"] == 0
    steps = int(params.n_walkers) * (int(params.n_steps) + 1)
    # every step (and each walker's start) takes one label
    assert v["tier_checked"] + v["exit_checked"] + v["classifier_checked"] == steps
    assert v["tier_checked"] > 0 and v["classifier_checked"] > 0
    if expect_exits:
        assert v["exit_checked"] > 0
    assert np.array_equal(res.cos, ref.cos) and np.array_equal(res.sin, ref.sin)
    assert np.array_equal(res.occupancy, ref.occupancy)
    assert res.n_escaped == 0 and res.n_error == 0
    return v


def test_c1_labels_at_full_size():
    cfg, ens, pp, spec = workloads.c1_entry()
    dg = DeviceGeometry(geometry_arrays(ens))
    _verify(dg, engine.walk_params(cfg.n_walkers, cfg, pp, seed=0), spec, "c1 1e6 walkers")


@pytest.mark.parametrize("cap_um", [0.25, 20.0])
def test_c1_geometry_with_other_tier_caps(cap_um):
    cfg, ens, pp, spec = workloads.c1_entry()
    dg = DeviceGeometry(geometry_arrays(ens))
    _verify(dg, engine.walk_params(200_000, cfg, pp, seed=5, tier_cap_um=cap_um), spec, f"c1 cap {cap_um} um")


@pytest.mark.parametrize("V_pL, vi, cells, kio, ts", [
    (0.1, 0.5, 1500, 40.0, 0.02),     # alpha* 0.62 um: short annulus next to a 0.28 um step
    (0.3, 0.95, 800, 90.0, 0.02),     # alpha* 0.06 um: thin gaps
    (0.0192, 0.994, 600, 101.0, 0.01),  # the sweep's tightest packing at its small-volume time step
])
def test_short_annulus_entries_use_and_pass_the_exit_certificate(V_pL, vi, cells, kio, ts):
    rho = vi / (V_pL * 1e-6)
    cfg = SimConfig(D0=2.0, ts=ts, T_max_ms=46.0, h_ms=1.0, n_walkers=1, small_deltas=[6.0],
                    big_deltas=[15.0, 25.0, 30.0, 40.0], b_values=[0.0, 1000.0, 2500.0, 4000.0, 6000.0],
                    periodic=True, precision="fp32")
    ens = create_periodic_ensemble(cells, rho, vi, cfg, seed=17)
    sigma = float(np.sqrt(2.0 * cfg.D0 * cfg.ts))
    assert ens.annulus.max() < 2.5 * sigma  # the certificate is compiled in for this entry
    dg = DeviceGeometry(geometry_arrays(ens))
    p = engine.walk_params(200_000, cfg, kio_to_pp(kio, ens.mean_AV, cfg), seed=9)
    _verify(dg, p, workloads.acquisition(), f"V {V_pL} pL v_i {vi}", expect_exits=True)
