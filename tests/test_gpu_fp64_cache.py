"""The FP64 walk's graph labels against the exact classifier, on the GPU.

classify_mode 5 (the reference's "exact_cached" classifier mode,
madi/walker_gpu.py:337-518) labels endpoints from the Delaunay graph in FP64
from the walker's own cell, takes labels from the certified ball of the last
one, and leaves decisions within rounding of a threshold (and cells whose
lists may be incomplete) to the exact KD classifier.  The reference proves
its cache equal to the exact classifier bit for bit
(scripts/validate_exact_classifier_cache.py:73-140); here the graph-labelled
walk must give the exact walk's Y, inside trace, escape and failure flags
bit for bit (and the fused walk its sums and occupancy), with most endpoints
labelled without the KD classifier: on a non-periodic Eq. S2-style box
(walkers escape; hull cells defer to the KD classifier), a periodic box, a
dense box with thin annuli, free water, and the reference's golden case
through run_walk_Y(classifier="exact_cached").
"""

from __future__ import annotations

import json
import sys

import numpy as np
import pytest

from custom_madi_amd import engine, walker_gpu
from custom_madi_amd.config import SimConfig
from custom_madi_amd.ensemble import DeviceGeometry, create_dummy_ensemble, geometry_arrays
from tests import _geom

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class _Cfg:
    def __init__(self, n_steps, steps_per_h, n_grid, sigma, ts):
        self.n_steps, self.steps_per_h, self.n_grid, self.sigma, self.ts = n_steps, steps_per_h, n_grid, sigma, ts


def _pair(dg, n, cfg, pp, seed):
    exact = engine.walk_y(dg, engine.walk_params(n, cfg, pp, seed))
    cached = engine.walk_y(dg, engine.walk_params(n, cfg, pp, seed, classify_mode=engine.MADI_CLASSIFY_FP64_CACHED))
    return exact, cached


def _assert_same(exact, cached, what):
    c = cached.counters
    say(f"[fp64 graph] {what}: scans {c['classify']} ball answers {c['safe_hit']} KD {c['fp64_recheck']} "
        f"moves {c['graph_iters']} (exact: {exact.counters['classify']} KD); kd nodes {c['kd_nodes']} vs "
        f"{exact.counters['kd_nodes']}")
    assert np.array_equal(cached.Y.cpu().numpy(), exact.Y.cpu().numpy())
    assert np.array_equal(cached.inside_trace, exact.inside_trace)
    assert np.array_equal(cached.escaped, exact.escaped) and np.array_equal(cached.err, exact.err)
    # every endpoint is a classification or an answer from the certified ball
    assert c["classify"] + c["safe_hit"] == exact.counters["classify"]
    assert c["crossings"] == exact.counters["crossings"] and c["rejects"] == exact.counters["rejects"]


def test_graph_labelled_walk_equals_exact_walk_non_periodic():
    ens = _geom.cellular(L=40.0, seed=5)
    dg = DeviceGeometry(geometry_arrays(ens))
    cfg = _Cfg(1500, 50, 31, float(np.sqrt(2 * 2.0 * 0.02)), 0.02)
    exact, cached = _pair(dg, 8192, cfg, 0.05, 11)
    _assert_same(exact, cached, "cellular L 40")
    assert exact.escaped.any()  # the box is small enough that some walkers leave it
    assert cached.counters["safe_hit"] > 0
    # the KD classifier runs for the starts and a small share of the endpoints
    assert cached.counters["fp64_recheck"] < 0.05 * exact.counters["classify"]


def test_cached_walk_equals_exact_walk_periodic_and_dense():
    for name, ens, sigma, ts in (
        ("periodic 64 cells", _geom.periodic(n_cells=64, seed=0), np.sqrt(2 * 2.0 * 0.05), 0.05),
        ("periodic dense v_i 0.95", _geom.periodic(n_cells=300, rho_per_uL=4.0e6, vi=0.95, seed=2),
         np.sqrt(2 * 2.0 * 0.01), 0.01),
    ):
        dg = DeviceGeometry(geometry_arrays(ens))
        cfg = _Cfg(1200, 40, 31, float(sigma), ts)
        exact, cached = _pair(dg, 8192, cfg, 0.2, 23)
        _assert_same(exact, cached, name)
        assert cached.counters["safe_hit"] > 0
        assert cached.counters["fp64_recheck"] < 0.05 * exact.counters["classify"]


def test_cached_fused_walk_equals_exact_fused_walk():
    ens = _geom.cellular(L=150.0, seed=9)
    dg = DeviceGeometry(geometry_arrays(ens))
    G = 14
    cfg = _Cfg(650, 50, G, float(np.sqrt(2 * 2.0 * 0.02)), 0.02)
    spec = engine.PhaseSpec(j_delta=np.array([1, 2]), j_Delta=np.array([5, 9]), j_sum=np.array([6, 11]),
                            pc=np.array([[0.0, 0.05, 0.2], [0.0, 0.1, 0.3]]), dirs=engine.AXIS_DIRECTIONS)
    a = engine.walk_fused(dg, engine.walk_params(4096, cfg, 0.03, 321), spec, precision=64)
    b = engine.walk_fused(dg, engine.walk_params(4096, cfg, 0.03, 321, classify_mode=engine.MADI_CLASSIFY_FP64_CACHED),
                          spec, precision=64)
    assert np.array_equal(a.cos, b.cos) and np.array_equal(a.sin, b.sin)
    assert np.array_equal(a.occupancy, b.occupancy)
    assert a.n_escaped == b.n_escaped == 0 and a.n_error == b.n_error == 0
    assert b.counters["safe_hit"] > 0


def test_free_water_walk_is_unchanged():
    cfg = SimConfig(D0=2.0, ts=0.02, T_max_ms=12.0, n_walkers=1024, n_ensembles=1, small_deltas=[2.0],
                    big_deltas=[5.0], b_values=[0.0, 1000.0])
    ens = create_dummy_ensemble(cfg)
    dg = DeviceGeometry(geometry_arrays(ens))
    exact, cached = _pair(dg, 1024, cfg, 0.5, 4)
    _assert_same(exact, cached, "free water")


def test_run_walk_Y_exact_cached_equals_the_reference_golden():
    """classifier="exact_cached" through the drop-in API: the reference's own
    CPU walk (tests/golden/walk_vectors.npz) within its golden tolerance, and
    bitwise the exact mode's Y."""
    ens, cfg, pp, g = _geom.reference_walk_golden()
    stream = walker_gpu.make_walk_random_stream(cfg, int(g["scalars"][13]), ensemble=ens)
    Yc, n_esc, tel = walker_gpu.run_walk_Y(ens, 0.0, cfg, verbose=False, pp=pp, return_telemetry=True,
                                           classifier="exact_cached", random_stream=stream)
    Ye = walker_gpu.run_walk_Y(ens, 0.0, cfg, verbose=False, pp=pp, random_stream=stream)[0]
    assert n_esc == 0 and np.array_equal(Yc, Ye)
    assert np.allclose(Yc, g["Y"], rtol=5e-11, atol=5e-12)
    assert tel["classifier_cache"]["tier1"] > 0
    assert tel["classifier_cache"]["requests"] == cfg.n_walkers * (cfg.n_steps + 1)
    say(f"[fp64 graph] golden case: {tel['classifier_cache']}")


def test_fp32_walks_reject_the_fp64_cache_mode():
    ens = _geom.periodic(n_cells=64, seed=0)
    dg = DeviceGeometry(geometry_arrays(ens))
    cfg = _Cfg(100, 50, 3, float(np.sqrt(2 * 2.0 * 0.02)), 0.02)
    spec = engine.PhaseSpec(j_delta=np.array([0]), j_Delta=np.array([1]), j_sum=np.array([2]),
                            pc=np.array([[0.0]]), dirs=engine.AXIS_DIRECTIONS)
    p = engine.walk_params(256, cfg, 0.1, 1, classify_mode=engine.MADI_CLASSIFY_FP64_CACHED)
    with pytest.raises(ValueError, match="FP64 walk's label cache"):
        engine.walk_fused(dg, p, spec, precision=32)
    cfg.classifier_mode, cfg.precision = "exact_cached", "fp32"
    assert engine.walk_params(256, cfg, 0.1, 1).classify_mode == 0  # FP32 never selects the FP64 cache
    cfg.precision = "fp64"
    assert engine.walk_params(256, cfg, 0.1, 1).classify_mode == engine.MADI_CLASSIFY_FP64_CACHED


@pytest.mark.parametrize("mode", ["exact", "exact_cached"])
def test_fp64_kio_sweep_in_one_launch_equals_walks_one_by_one(mode, monkeypatch):
    """The FP64 walks of a k_io sweep run several per launch
    (madi_walk_f64_multi) and give each walk's results bit for bit, across
    walker chunks and with the sweep split over launches."""
    from custom_madi_amd import signal
    ens = _geom.periodic(n_cells=64, seed=3)
    cfg = SimConfig(D0=2.0, ts=0.02, T_max_ms=12.0, n_walkers=2048, walker_chunk=1024, n_ensembles=1,
                    small_deltas=[1.0, 2.0, 3.0], big_deltas=[4.0, 5.0, 6.0, 7.0, 8.0],
                    b_values=[0.0, 1000.0, 2500.0], classifier_mode=mode)
    cols = signal.build_columns(cfg)
    args = (cols.j_delta, cols.j_Delta, cols.j_sum, cols.phase_coef)
    assert engine.spec_from_columns(*args) is None  # reduced from Y
    pps = [0.0, 0.01, 0.05, 0.3, 1.0]
    one = [walker_gpu._walk_and_reduce_one_ensemble(ens, pp, cfg, *args, 77) for pp in pps]
    # two walks per launch: the sweep takes three launches per walker chunk
    monkeypatch.setattr(walker_gpu, "_MULTI_WALK_Y_BYTES", 2 * 1024 * cfg.n_grid * 24)
    many = walker_gpu._walk_and_reduce_kio_sweep(ens, pps, cfg, *args, 77)
    for a, b in zip(one, many):
        assert np.array_equal(a.cos_sum, b.cos_sum) and np.array_equal(a.sin_sum, b.sin_sum)
        assert np.array_equal(a.occupancy_counts, b.occupancy_counts) and a.n_walkers == b.n_walkers
        assert a.classifier_cache_stats == b.classifier_cache_stats
        assert a.pp_values == b.pp_values and a.analytic_kio_eq5_values == b.analytic_kio_eq5_values
    say(f"[fp64 sweep] {mode}: {len(pps)} walks equal one by one and together")


@pytest.mark.parametrize("case", ["escapes K3", "ragged K7", "max K64"])
def test_walk_y_multi_equals_walk_y_per_walk(case):
    """madi_walk_f64_multi against madi_walk_f64 walk by walk, bit for bit:
    Y, inside traces, escape and failure flags and counters.  A
    non-periodic box whose walkers escape (a walk's escaped threads still
    make the walker's shared draws for its other walks), walker counts that
    leave the last block part-filled, step counts that are not a multiple of
    the walk count, and the largest walk count."""
    mode = {"escapes K3": engine.MADI_CLASSIFY_FP64_CACHED, "ragged K7": 0,
            "max K64": engine.MADI_CLASSIFY_FP64_CACHED}[case]
    if case == "escapes K3":
        ens, n, pps, cfg = _geom.cellular(L=40.0, seed=5), 1000, [0.0, 0.05, 1.0], \
            _Cfg(1500, 50, 31, float(np.sqrt(2 * 2.0 * 0.02)), 0.02)
    elif case == "ragged K7":
        ens, n, pps, cfg = _geom.periodic(n_cells=64, seed=1), 333, [0.0, 0.01, 0.02, 0.1, 0.3, 0.7, 1.0], \
            _Cfg(703, 50, 15, float(np.sqrt(2 * 2.0 * 0.05)), 0.05)
    else:
        ens, n, cfg = _geom.periodic(n_cells=64, seed=2), 130, _Cfg(400, 40, 11, float(np.sqrt(2 * 2.0 * 0.05)), 0.05)
        pps = list(np.linspace(0.0, 1.0, 64))
    dg = DeviceGeometry(geometry_arrays(ens))
    params = engine.walk_params(n, cfg, pps[0], 19, classify_mode=mode)
    many = engine.walk_y_multi(dg, params, pps)
    counted = engine.walk_y_multi(dg, params, pps, traces=False)  # occupancy counted on the device
    escaped_any = False
    for pp, b, c in zip(pps, many, counted):
        assert c.inside_trace is None
        assert np.array_equal(c.occupancy, np.sum(b.inside_trace, axis=0, dtype=np.int64))
        a = engine.walk_y(dg, engine.walk_params(n, cfg, pp, 19, classify_mode=mode))
        assert np.array_equal(a.Y.cpu().numpy(), b.Y.cpu().numpy())
        assert np.array_equal(a.inside_trace, b.inside_trace)
        assert np.array_equal(a.escaped, b.escaped) and np.array_equal(a.err, b.err)
        assert a.counters == b.counters
        escaped_any |= bool(a.escaped.any())
    if case == "escapes K3":
        assert escaped_any  # the box is small enough that walkers leave it
    say(f"[fp64 multi] {case}: {len(pps)} walks of {n} walkers equal one by one")


def test_multi_kio_ensembles_in_turn_equal_ensembles_built_first():
    """run_simulation_multi_kio_reduced builds an entry's Eq. S8 ensembles one
    at a time (the next one's CPU geometry on a worker thread while the
    current one walks); its results equal those of building every ensemble
    first, as the reference does, bit for bit."""
    from custom_madi_amd import signal
    # a short walk's Eq. S8 box is small: its realised v_i is within 0.03
    cfg = SimConfig(D0=2.0, ts=0.02, T_max_ms=12.0, n_walkers=1024, n_ensembles=3, small_deltas=[1.0, 2.0],
                    big_deltas=[4.0, 6.0], b_values=[0.0, 1000.0, 2500.0], classifier_mode="exact_cached",
                    geometry_validation_points=20_000, geometry_vi_tolerance=0.03)
    cols = signal.build_columns(cfg)
    args = (cols.j_delta, cols.j_Delta, cols.j_sum, cols.phase_coef)
    rho, V, kios, seed = 1.0e5, 6.3, [0.0, 5.0, 20.0, 60.0], 7
    got = walker_gpu.run_simulation_multi_kio_reduced(rho, V, kios, *args, cfg=cfg, seed=seed, verbose=False)
    want = {k: [] for k in kios}
    for e, ens in enumerate(walker_gpu._build_ensembles_for_entry(rho, V, cfg, seed, False)):
        pps = [0.0 if k == 0.0 else walker_gpu._checked_pp(walker_gpu.kio_to_pp(k, ens.mean_AV, cfg)) for k in kios]
        for k, r in zip(kios, walker_gpu._walk_and_reduce_kio_sweep(ens, pps, cfg, *args,
                                                                     walker_gpu._walk_seed(seed, e))):
            want[k].append(r)
    for k in kios:
        w = walker_gpu._merge_results(want[k])
        g = got[k]
        assert np.array_equal(g.cos_sum, w.cos_sum) and np.array_equal(g.sin_sum, w.sin_sum)
        assert np.array_equal(g.occupancy_counts, w.occupancy_counts) and g.n_walkers == w.n_walkers == 3 * 1024
        assert g.pp_values == w.pp_values
        assert json.dumps(g.geometry_stats, sort_keys=True) == json.dumps(w.geometry_stats, sort_keys=True)
    say(f"[fp64 entry] {len(kios)} k_io x 3 ensembles built in turn equal those built first")
