"""The drop-in walk API and its failure contracts on the HIP path.

* run_walk_Y with make_walk_random_stream (madi/walker_gpu.py:115-127,
  1093-1133) against the oracle's walk on the same injected stream: Y within
  the reference's golden tolerance (rtol 5e-11, atol 5e-12,
  scripts/run_gpu_golden_check.py) and its telemetry.
* The reference's free-water check (tests/physics_audit/
  test_finite_delta_reference.py:35-61) through run_walk_Y on the device,
  and through run_simulation_reduced (the fused FP64 and FP32 walks).
* A walker leaving Omega_sim raises RuntimeError "fatal boundary condition"
  (tests/physics_audit/test_p0_remediation.py:66-84) through run_walk_Y and
  run_simulation_reduced, FP64 and FP32.
* Coincident seeds raise the classifier-failure RuntimeError
  (madi/walker_gpu.py:938-940).
* A walk split into walker chunks gives the sums of its chunks, counters
  included.
* In the batched and served forms an entry whose walkers escape fails alone:
  the other entries equal their solo runs bit for bit, and the failed one is
  kept out of the results and the library stream.
"""

from __future__ import annotations

import sys
from dataclasses import replace

import numpy as np
import pytest

from oracle import oracle as O
from custom_madi_amd import batch, engine, serve, signal, walker_gpu
from custom_madi_amd.config import GAMMA_RAD, SimConfig
from custom_madi_amd.ensemble import (
    DeviceGeometry, build_kdtree, create_dummy_ensemble, geometry_arrays,
)
from tests import _geom

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _cfg(**kw):
    base = dict(D0=2.0, ts=0.02, T_max_ms=12.0, n_walkers=256, n_ensembles=1, small_deltas=[2.0],
                big_deltas=[5.0, 8.0], b_values=[0.0, 1000.0, 2500.0])
    base.update(kw)
    return SimConfig(**base)


def test_run_walk_Y_with_the_reference_stream_matches_the_oracle():
    ens = _geom.cellular(L=50.0, seed=3)
    cfg = _cfg()
    stream = walker_gpu.make_walk_random_stream(cfg, 20_260_901, ensemble=ens)
    say("[dropin] run_walk_Y, injected stream")
    Y, n_esc, tel = walker_gpu.run_walk_Y(ens, 30.0, cfg, seed=7, verbose=False, return_telemetry=True,
                                          random_stream=stream)
    pp = walker_gpu.kio_to_pp(30.0, ens.mean_AV, cfg)
    og = O.Geometry.from_arrays(geometry_arrays(ens))
    oY, itr, esc, err = O.walk(og, cfg.n_walkers, cfg.n_steps, cfg.steps_per_h, cfg.n_grid, cfg.sigma, cfg.ts, pp,
                               ens.source_lo, ens.source_hi,
                               stream=(stream.initial_positions, stream.increments, stream.acceptance_uniforms))
    assert not err.any() and not esc.any() and n_esc == 0
    assert np.allclose(Y, oY, rtol=5e-11, atol=5e-12)
    assert tel["pp"] == pp
    assert np.array_equal(tel["occupancy_counts"], itr.sum(axis=0))
    assert np.allclose(tel["occupancy_fraction"], itr.mean(axis=0))
    assert tel["classifier_cache"]["requests"] >= cfg.n_walkers * cfg.n_steps


def test_run_walk_Y_equals_the_reference_cpu_walk_golden():
    """The reference's own _walk_cpu output (tests/golden/walk_vectors.npz,
    made by tests/golden/make_walk_vectors.py) on its seeded Eq. S2 population
    and injected stream: the HIP FP64 walk through run_walk_Y within the
    reference's golden tolerance, its occupancy equal to the reference's inside
    trace; madi_reduce_f64 on the reference's Y against its _reduce_cpu sums."""
    ens, cfg, pp, g = _geom.reference_walk_golden()
    stream = walker_gpu.make_walk_random_stream(cfg, int(g["scalars"][13]), ensemble=ens)
    assert _geom.stream_digest(stream) == str(g["stream_sha256"])
    say("[dropin] run_walk_Y against the reference CPU walk golden")
    Y, n_esc, tel = walker_gpu.run_walk_Y(ens, 0.0, cfg, verbose=False, pp=pp, return_telemetry=True,
                                          random_stream=stream)
    assert n_esc == 0
    assert np.allclose(Y, g["Y"], rtol=5e-11, atol=5e-12)
    assert np.array_equal(tel["occupancy_counts"], g["inside_trace"].astype(np.int64).sum(axis=0))
    from custom_madi_amd import _hip
    Yd = _hip.to_device(np.ascontiguousarray(g["Y"], dtype=np.float64))
    cs, sn = engine.reduce_columns(Yd, g["cols_jd"], g["cols_jD"], g["cols_js"], g["cols_pc"])
    assert np.allclose(cs, g["cos"], rtol=1e-11, atol=1e-11) and np.allclose(sn, g["sin"], rtol=1e-11, atol=1e-11)


def _free_water_cfg():
    # the reference's free-water check: S8 box of the acellular control,
    # 512 walkers, 70 ms, microsecond steps
    timings = [(1.0, 20.0), (4.0, 20.0), (8.0, 20.0), (12.0, 20.0), (7.0, 25.0), (20.0, 50.0)]
    cfg = SimConfig(L=None, n_walkers=512, n_ensembles=1, T_max_ms=70.0,
                    small_deltas=sorted({d for d, _ in timings}), big_deltas=sorted({D for _, D in timings}),
                    b_values=[0.0, 500.0, 1000.0, 2000.0, 4000.0, 6000.0])
    return cfg, timings


def test_free_water_through_run_walk_Y_matches_stejskal_tanner():
    cfg, timings = _free_water_cfg()
    ens = create_dummy_ensemble(cfg)
    say("[dropin] free water, run_walk_Y (70,000 steps)")
    Y, escaped = walker_gpu.run_walk_Y(ens, 0.0, cfg, pp=0.0, seed=20_260_901, verbose=False)
    assert escaped == 0
    for delta, Delta in timings:
        jd, jD, js = (int(round(t / cfg.h_ms)) for t in (delta, Delta, delta + Delta))
        dM = Y[:, jd, :] + Y[:, jD, :] - Y[:, js, :]
        for b in cfg.b_values:
            phase = GAMMA_RAD * signal.G_from_b(b, delta, Delta) * dM * 1.0e-9
            x = np.cos(phase).ravel()
            se = np.std(x, ddof=1) / np.sqrt(x.size)
            assert abs(x.mean() - np.exp(-b * cfg.D0 * 1e-3)) <= 5.0 * se + 0.012, (delta, Delta, b)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_free_water_through_run_simulation_reduced(precision):
    cfg, timings = _free_water_cfg()
    cfg = replace(cfg, precision=precision, n_walkers=4096, ts=0.01, small_deltas=[8.0], big_deltas=[20.0])
    cols = signal.build_columns(cfg)
    say(f"[dropin] free water, run_simulation_reduced {precision}")
    res = walker_gpu.run_simulation_reduced(0.0, 0.0, 0.0, cols.j_delta, cols.j_Delta, cols.j_sum, cols.phase_coef,
                                            cfg=cfg, seed=5, verbose=False)
    assert res.n_walkers == 4096 and res.n_escaped == 0
    S = res.cos_sum / res.n_eff
    expected = np.exp(-np.asarray(cols.b_values) * cfg.D0 * 1e-3)
    se = np.sqrt(np.maximum(1.0 - S ** 2, 0.0) / res.n_eff)
    assert np.all(np.abs(S - expected) <= 5.0 * se + 0.012)


def test_an_escape_through_run_walk_Y_is_fatal():
    # tests/physics_audit/test_p0_remediation.py:66-84 on the device
    cfg = SimConfig(L=2.0, n_walkers=1, n_ensembles=1, T_max_ms=2.0, small_deltas=[1.0], big_deltas=[1.0],
                    b_values=[0.0])
    steps = cfg.n_steps
    increments = np.zeros((steps, 1, 3), dtype=np.float64)
    increments[0, 0, 0] = 2.0
    stream = walker_gpu.WalkRandomStream(initial_positions=np.asarray([[1.0, 1.0, 1.0]]), increments=increments,
                                         acceptance_uniforms=np.zeros((steps, 1), dtype=np.float64))
    say("[dropin] forced escape, run_walk_Y")
    with pytest.raises(RuntimeError, match="fatal boundary condition"):
        walker_gpu.run_walk_Y(create_dummy_ensemble(cfg), 0.0, cfg, pp=0.0, verbose=False, random_stream=stream)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_an_escape_through_run_simulation_reduced_is_fatal(precision):
    # a 3 um box: walkers starting in its 1.2 um source cube leave it long before 12 ms
    cfg = _cfg(L=3.0, precision=precision)
    cols = signal.build_columns(cfg)
    say(f"[dropin] escaping walk, run_simulation_reduced {precision}")
    with pytest.raises(RuntimeError, match="fatal boundary condition"):
        walker_gpu.run_simulation_reduced(0.0, 0.0, 0.0, cols.j_delta, cols.j_Delta, cols.j_sum, cols.phase_coef,
                                          cfg=cfg, seed=1, verbose=False)


def test_an_escape_in_a_kio_sweep_is_fatal():
    """The FP64 k_io sweep (several walks per launch, columns reduced from Y)
    keeps the reference's fatal escape (madi/walker_gpu.py:1023-1028)."""
    cfg = _cfg(L=3.0, small_deltas=[1.0, 2.0, 3.0], big_deltas=[4.0, 5.0, 6.0, 7.0, 8.0])
    cols = signal.build_columns(cfg)
    args = (cols.j_delta, cols.j_Delta, cols.j_sum, cols.phase_coef)
    assert engine.spec_from_columns(*args) is None  # the sweep path, not the fused walk
    say("[dropin] escaping walks, run_simulation_multi_kio_reduced k_io sweep")
    with pytest.raises(RuntimeError, match="fatal boundary condition"):
        walker_gpu.run_simulation_multi_kio_reduced(0.0, 0.0, [0.0, 20.0, 60.0], *args, cfg=cfg, seed=1,
                                                    verbose=False)


def test_coincident_seeds_are_a_classifier_failure_of_the_exact_walk():
    ens = _geom.cellular(L=50.0, seed=3)
    mid = np.all((ens.seeds > ens.source_lo + 1.0) & (ens.seeds < ens.source_hi - 1.0), axis=1)
    i = int(np.flatnonzero(mid)[0])
    j = int(np.argsort(np.linalg.norm(ens.seeds - ens.seeds[i], axis=1))[1])  # i's nearest neighbour
    ens.seeds[j] = ens.seeds[i]
    (ens.kd_node_seed, ens.kd_node_axis, ens.kd_node_left, ens.kd_node_right,
     ens.kd_node_parent) = build_kdtree(ens.seeds)
    cfg = _cfg(n_walkers=64)
    rng = np.random.default_rng(0)
    stream = walker_gpu.WalkRandomStream(
        initial_positions=ens.seeds[i] + rng.uniform(-0.05, 0.05, size=(64, 3)),
        increments=rng.normal(0.0, cfg.sigma, size=(cfg.n_steps, 64, 3)),
        acceptance_uniforms=rng.uniform(0.0, 1.0, size=(cfg.n_steps, 64)))
    say("[dropin] coincident seeds, run_walk_Y")
    with pytest.raises(RuntimeError, match="classification failed"):
        walker_gpu.run_walk_Y(ens, 20.0, cfg, verbose=False, random_stream=stream)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_walker_chunks_sum_to_the_whole_walk(precision):
    ens = _geom.periodic(n_cells=200, rho_per_uL=5.0e5, seed=4)
    cfg = _cfg(n_walkers=20_000, periodic=True, precision=precision, L=ens.L)
    cols = signal.build_columns(cfg)
    args = (ens, 0.02, cfg, cols.j_delta, cols.j_Delta, cols.j_sum, cols.phase_coef)
    whole = walker_gpu._walk_and_reduce_one_ensemble(*args[:2], replace(cfg, walker_chunk=10_000), *args[3:], 11)
    a = walker_gpu._walk_and_reduce_one_ensemble(*args[:2], replace(cfg, n_walkers=10_000), *args[3:], 11)
    b = walker_gpu._walk_and_reduce_one_ensemble(*args[:2], replace(cfg, n_walkers=10_000), *args[3:], 11 + 10_000)
    assert whole.n_walkers == 20_000
    assert np.allclose(whole.cos_sum, a.cos_sum + b.cos_sum, rtol=1e-12, atol=1e-9)
    assert np.array_equal(whole.occupancy_counts, a.occupancy_counts + b.occupancy_counts)
    wc, ac, bc = (r.classifier_cache_stats[0] for r in (whole, a, b))
    for k in ("device_classify", "device_safe_hit", "device_facets", "device_crossings", "device_rejects"):
        assert wc[k] == ac[k] + bc[k], k


def _escaping_entry(key):
    """A non-periodic 12 um box: its walkers leave Omega_sim within 46 ms."""
    ens = _geom.cellular(L=12.0, rho_per_uL=5.0e5, seed=8, margin_spacings=1.0)
    req = batch.EntryRequest(key=key, kio=10.0, rho=5.0e5, V=1.2, n_cells=len(ens.seeds), n_walkers=3000, seed=0,
                             walk_seed=99)
    cfg = req.config()
    return batch.PreparedEntry(req, geometry_arrays(ens, with_kd=False),
                               walker_gpu.kio_to_pp(10.0, ens.mean_AV, cfg), ens.mean_AV, ens.alpha_star, ens.L, 0.0)


def test_an_escaping_entry_fails_alone_in_batched_and_served_runs(tmp_path):
    spec = batch.acquisition_spec(n_dirs=6, delta=6.0, Deltas=(15.0, 25.0), b_values=(0.0, 1000.0, 4000.0))
    good = [batch.prepare_entry(r) for r in batch.library_requests(4, seed=31, walkers=(4000, 9000),
                                                                     cells=(60, 300))]
    bad = _escaping_entry(("escapes",))
    solo = {p.req.key: batch.run_batched([p], spec)[0] for p in good}
    say("[dropin] batch with an escaping entry")
    failed = []
    mixed = batch.run_batched([good[0], bad, *good[1:]], spec, failed=failed)
    assert [f.key for f in failed] == [("escapes",)]
    assert failed[0].n_escaped > 0 and "fatal boundary condition" in failed[0].error
    assert sorted(r.key for r in mixed) == sorted(solo)
    for r in mixed:
        s = solo[r.key]
        assert np.array_equal(r.cos, s.cos) and np.array_equal(r.sin, s.sin)
        assert np.array_equal(r.occupancy, s.occupancy)
    with pytest.raises(RuntimeError, match="fatal boundary condition"):
        batch.run_batched([bad, good[0]], spec)

    say("[dropin] served requests with an escaping entry")
    lookup = {p.req.key: p for p in good + [bad]}
    sch = batch.DeviceScheduler(spec)
    out = tmp_path / "served.jsonl"
    stream = serve.LibraryStream(str(out))
    srv = serve.Server(sch, stream=stream, prepare=lambda r: lookup[r.key])
    reqs = [good[0].req, bad.req, *[p.req for p in good[1:]]]
    rep = serve.serve_poisson(srv, reqs, np.linspace(0.0, 0.2, len(reqs)))
    sch.close()
    stream.close()
    assert rep["finished"] == len(good) and rep["failed"] == 1
    assert [r.key for r in srv.failed] == [("escapes",)]
    for r in srv.results:
        assert np.array_equal(r.cos, solo[r.key].cos)
    assert len(out.read_text().splitlines()) == len(good)
