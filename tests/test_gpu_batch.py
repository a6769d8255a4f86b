"""Batched / served walks and the drop-in API on an MI355X.

An entry's result must be bitwise identical whether it runs alone (one
launch), in a batch, or in a different mix of entries; the batched FP32
walk must agree with the FP64 exact-classifier walk within |dS| <= 2e-3 on
the same seeds; the drop-in library API must build, save and load.
"""

from __future__ import annotations

import sys

import numpy as np
import pytest

from custom_madi_amd import batch, engine
from custom_madi_amd.config import SimConfig
from custom_madi_amd.ensemble import DeviceGeometry, geometry_arrays

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def small_requests():
    return batch.library_requests(5, seed=11, walkers=(3000, 9000), cells=(60, 400), T_max_ms=12.0)


@pytest.fixture(scope="module")
def spec():
    return batch.acquisition_spec(n_dirs=6, delta=2.0, Deltas=(5.0, 8.0, 10.0), b_values=(0.0, 1000.0, 4000.0))


def _same(a, b):
    return (np.array_equal(a.cos, b.cos) and np.array_equal(a.sin, b.sin)
            and np.array_equal(a.occupancy, b.occupancy))


def test_batched_equals_plain_bitwise(small_requests, spec):
    preps = [batch.prepare_entry(r) for r in small_requests]
    plain = batch.run_plain(small_requests, spec)
    mixed = batch.run_batched(preps, spec, max_tiles_per_generation=7)
    by_key = {r.key: r for r in mixed}
    for r in plain:
        assert r.n_escaped == 0 and r.n_error == 0
        assert _same(r, by_key[r.key]), r.key


def test_entry_alone_equals_entry_in_mixes(small_requests, spec):
    preps = [batch.prepare_entry(r) for r in small_requests]
    alone = batch.run_batched(preps[2:3], spec)[0]
    mix1 = {r.key: r for r in batch.run_batched(preps, spec)}
    mix2 = {r.key: r for r in batch.run_batched(preps[::-1], spec, max_tiles_per_generation=3, grid_per_cu=1)}
    assert _same(alone, mix1[alone.key])
    assert _same(alone, mix2[alone.key])


def test_batched_fp32_matches_fp64_exact_walk(spec):
    req = batch.library_requests(1, seed=5, walkers=(65_536, 65_536), cells=(300, 300), T_max_ms=12.0)[0]
    prep = batch.prepare_entry(req)
    res32 = batch.run_batched([prep], spec)[0]
    # the same geometry with the exact classifier's KD tree for the FP64 walk
    from custom_madi_amd.ensemble import create_periodic_ensemble
    ens = create_periodic_ensemble(req.n_cells, req.rho, req.rho * req.V * 1e-6, req.config(), seed=req.seed)
    dg = DeviceGeometry(geometry_arrays(ens))
    p = engine.walk_params(req.n_walkers, req.config(), prep.pp, req.walk_seed)
    # the FP64 run adds b = 16000 (4 x 4000): Var cos(phi) = (1 + S(4b)) / 2 - S(b)^2
    specv = batch.acquisition_spec(n_dirs=6, delta=2.0, Deltas=(5.0, 8.0, 10.0),
                                   b_values=(0.0, 1000.0, 4000.0, 16000.0))
    res64 = engine.walk_fused(dg, p, specv, precision=64)
    from custom_madi_amd import workloads
    S64, se = workloads.signal_se(res64.cos, req.n_walkers, np.array([0, 1, 2]), np.array([0, 2, 3]))
    dS = np.abs(res32.cos / req.n_walkers - S64)
    say(f"batched fp32 vs fp64: max |dS| = {dS.max():.3e}, max |dS| / SE = "
        f"{(dS / np.maximum(se, 1e-300)).max():.3f}")
    assert dS.max() <= 2e-3
    assert np.all(dS <= 3.0 * se)
    assert res64.n_escaped == 0


def test_scheduler_respects_hbm_reserve(small_requests, spec):
    import torch
    prep = batch.prepare_entry(small_requests[0])
    free, _ = torch.cuda.mem_get_info()
    sch = batch.DeviceScheduler(spec, hbm_reserve_bytes=free + (1 << 30))
    assert not sch.admit(prep)
    sch = batch.DeviceScheduler(spec)
    assert sch.admit(prep)
    out = sch.drain()
    assert len(out) == 1 and out[0].n_escaped == 0


def test_serving_finishes_every_request(spec):
    from custom_madi_amd import serve
    reqs = batch.library_requests(6, seed=21, walkers=(2000, 6000), cells=(60, 200), T_max_ms=12.0)
    sch = batch.DeviceScheduler(spec, max_tiles_per_generation=8)
    srv = serve.Server(sch)
    rep = serve.serve_poisson(srv, reqs, np.linspace(0.0, 0.3, len(reqs)))
    assert rep["finished"] == len(reqs)
    assert rep["latency_s_p99"] >= rep["latency_s_p50"] > 0
    ref = {r.key: r for r in batch.run_batched([batch.prepare_entry(r) for r in reqs], spec)}
    for r in srv.results:
        assert _same(r, ref[r.key])


def test_dropin_library_build_v5_round_trip(tmp_path):
    """build_library_from_triplets on reference-semantics geometry (S8 box, Omega_pop,
    exact FP64 walk, v5 diagnostics on the declared subset grid), saved and loaded."""
    from custom_madi_amd import library, signal
    from custom_madi_amd.config import ENSEMBLE_MEAN_SUBSET_B_VALUES_S_MM2
    cfg = SimConfig(n_walkers=256, n_ensembles=2, T_max_ms=110.0, D0=2.0, ts=0.05,
                    small_deltas=[5.0, 7.0, 10.0, 12.0, 15.0, 20.0, 25.0, 30.0],
                    big_deltas=[15.0, 25.0, 30.0, 36.0, 40.0, 50.0, 60.0, 80.0],
                    b_values=list(ENSEMBLE_MEAN_SUBSET_B_VALUES_S_MM2), geometry_validation_points=20_000,
                    geometry_vi_tolerance=0.03)
    cols = signal.build_columns(cfg)
    out = tmp_path / "lib.npz"
    lib = library.build_library_from_triplets([(20.0, 1.0e5, 6.0), (0.0, 0.0, 0.0)], cfg=cfg, seed=7,
                                              save_path=str(out), verbose=False)
    assert len(lib) == 2
    back = library.load_library(str(out))
    meta = library.load_library_meta(str(out))
    assert meta["library_schema"] == "madi-library-v5" and meta["has_v5_diagnostics"]
    assert meta["ensemble_means_subset_shape"] == (2, 2, library.ENSEMBLE_MEAN_SUBSET_N_COLUMNS)
    cell = [e for e in back if not e.is_free_water][0]
    free = [e for e in back if e.is_free_water][0]
    assert cell.vector.shape == (cols.n_pairs * cols.n_b,)
    assert np.all(cell.vector[::cols.n_b] == 1.0)      # b = 0 columns
    assert 0.55 < cell.vi < 0.65
    assert cell.metadata["boundary"]["n_escaped"] == 0
    # free water: each column is a mean of cos(phase) over 3 axes x walkers x
    # ensembles, with phase ~ N(0, 2 b D0); its standard error follows from
    # E[cos^2] = (1 + S^4) / 2 with S = exp(-b D0)
    b = np.tile(np.asarray(cfg.b_values), cols.n_pairs)
    expect = np.exp(-b * cfg.D0 * 1e-3)
    n_eff = 3 * cfg.n_walkers * cfg.n_ensembles
    se = np.sqrt(np.maximum((1.0 + expect ** 4) / 2.0 - expect ** 2, 0.0) / n_eff)
    assert np.all(np.abs(free.vector - expect) <= 5.0 * se + 0.012)


def test_dropin_fp64_fused_equals_y_reduce_path():
    from custom_madi_amd import walker_gpu
    from tests import _geom
    ens = _geom.cellular(L=150.0, seed=9)
    cfg = SimConfig(L=150.0, n_walkers=3000, T_max_ms=6.0, D0=2.0, ts=0.01, small_deltas=[1.0],
                    big_deltas=[3.0, 4.0], b_values=[0.0, 1000.0, 4000.0])
    from custom_madi_amd import signal
    cols = signal.build_columns(cfg)
    fused = walker_gpu._walk_and_reduce_one_ensemble(ens, 0.05, cfg, cols.j_delta, cols.j_Delta, cols.j_sum,
                                                     cols.phase_coef, 1234)
    # a column set too large to fuse takes the Y + reduce path; compare the shared columns
    many = np.concatenate([cols.phase_coef, np.linspace(0.001, 0.02, 40)])
    jd = np.concatenate([cols.j_delta, np.full(40, 1)])
    jD = np.concatenate([cols.j_Delta, np.full(40, 3)])
    js = np.concatenate([cols.j_sum, np.full(40, 4)])
    assert engine.spec_from_columns(jd, jD, js, many) is None
    unfused = walker_gpu._walk_and_reduce_one_ensemble(ens, 0.05, cfg, jd, jD, js, many, 1234)
    n = len(cols.phase_coef)
    assert np.allclose(fused.cos_sum, unfused.cos_sum[:n], rtol=1e-10, atol=1e-8)
    assert np.array_equal(fused.occupancy_counts, unfused.occupancy_counts)


def test_tuner_choices_give_identical_results(spec):
    """The classification mode (per-lane state machine with 1 or 2 rounds per
    iteration, state machine with lane refill, walker waves with classifier
    waves)
    changes only where and when a proposal is classified and which lane walks
    a walker, never the result: every choice must give the same sums bit for
    bit, in the single-entry and the persistent batch kernels."""
    req = batch.library_requests(1, seed=31, walkers=(20_000, 20_000), cells=(200, 200), T_max_ms=12.0)[0]
    prep = batch.prepare_entry(req)
    geom = DeviceGeometry(prep.arrays)
    ref = None
    for rounds, cw in ((1, 1), (2, 1), (1, 3), (2, 3), (1, 4), (2, 4)):
        if True:
            say(f"single entry: scan_rounds={rounds} classify_mode={cw}")
            p = engine.walk_params(req.n_walkers, req.config(), prep.pp, req.walk_seed, scan_rounds=rounds,
                                   classify_mode=cw)
            res = engine.walk_fused(geom, p, spec, precision=32)
            assert res.n_escaped == 0 and res.n_error == 0
            if ref is None:
                ref = res
            else:
                assert np.array_equal(res.cos, ref.cos) and np.array_equal(res.sin, ref.sin)
                assert np.array_equal(res.occupancy, ref.occupancy)
    for cw in (1, 3, 4):
        say(f"batch: classify_mode={cw}")
        out = batch.run_batched([prep], spec, classify_mode=cw)[0]
        assert np.array_equal(out.cos, ref.cos) and np.array_equal(out.occupancy, ref.occupancy)


def test_exit_certificate_changes_no_result(spec):
    """The streamed walks (modes 3 and 4) answer an inside walker's proposal
    past its cell without classifying it when the draw rejects both crossing
    kinds (tier_exit); the per-tile walk (mode 1) classifies every such
    proposal.  With annuli far below a step deviation (v_i = 0.95) the
    certificate fires and the sums still equal mode 1's bit for bit, single
    entry and batched; with annuli above 2.5 deviations (v_i = 0.5) it stays
    off."""
    for vi, fires in ((0.95, True), (0.5, False)):
        req = batch.library_requests(1, seed=41, walkers=(20_000, 20_000), cells=(200, 200), V=(2.0, 2.0),
                                     vi=(vi, vi), T_max_ms=12.0)[0]
        prep = batch.prepare_entry(req)
        geom = DeviceGeometry(prep.arrays)
        ref = None
        for cw in (1, 3, 4):
            p = engine.walk_params(req.n_walkers, req.config(), prep.pp, req.walk_seed, classify_mode=cw)
            res = engine.walk_fused(geom, p, spec, precision=32)
            c = res.counters
            say(f"v_i {vi} classify_mode {cw}: {c['exits']} certified exits, {c['classify']} classifications")
            assert res.n_escaped == 0 and res.n_error == 0
            if ref is None:
                assert c["exits"] == 0
                ref = res
                continue
            assert (c["exits"] > 0) == fires
            assert np.array_equal(res.cos, ref.cos) and np.array_equal(res.sin, ref.sin)
            assert np.array_equal(res.occupancy, ref.occupancy)
        out = batch.run_batched([prep], spec, classify_mode=4)[0]
        assert np.array_equal(out.cos, ref.cos) and np.array_equal(out.occupancy, ref.occupancy)


def test_autotuner_picks_a_candidate_and_caches_it(spec, tmp_path, monkeypatch):
    from custom_madi_amd import tune
    monkeypatch.setenv("MADI_TUNE_CACHE", str(tmp_path / "tune.json"))
    preps = tune.sample_entries(n=2, walkers=(5_000, 9_000))
    cands = [(1, 2, 3), (3, 1, 4)]
    choice = tune.autotune(spec, preps=preps, candidates=cands, use_cache=False, log=say)
    assert (choice.classify_mode, choice.scan_rounds, choice.grid_per_cu) in cands
    assert choice.trial_s > 0
    again = tune.autotune(spec, preps=preps, candidates=[(1, 1, 2)])   # served from the cache
    assert (again.classify_mode, again.scan_rounds, again.grid_per_cu) == \
        (choice.classify_mode, choice.scan_rounds, choice.grid_per_cu)


def test_graph_replays_equal_eager_launches(small_requests, spec):
    """Generations replayed from captured graphs (several replays of each,
    with eager finalisation work in between) give the eager launches'
    results bit for bit."""
    preps = [batch.prepare_entry(r) for r in small_requests]
    eager = {r.key: r for r in batch.run_batched(preps, spec, max_tiles_per_generation=5, use_graphs=False)}
    sch = batch.DeviceScheduler(spec, max_tiles_per_generation=5)
    assert len(sch.graphs) == sch.max_inflight
    for p in preps:
        assert sch.admit(p)
    out = sch.drain()
    say(f"graph launches: {sch.graph_launches} over {sch.gen} generations")
    assert sch.graph_launches == sch.gen and sch.gen >= 2 * sch.max_inflight
    sch.close()
    assert len(out) == len(preps)
    for r in out:
        assert _same(r, eager[r.key]), r.key


def test_entries_retire_while_others_are_admitted(spec):
    """One-tile generations with entries admitted between them: every entry's
    geometry is uploaded while earlier entries finish and release theirs.  A
    geometry released before the generation that reads it has completed could
    be overwritten by the next upload; every result must equal the batched
    run's bit for bit."""
    from custom_madi_amd import serve
    reqs = batch.library_requests(12, seed=57, walkers=(700, 2500), cells=(60, 160), T_max_ms=12.0)
    sch = batch.DeviceScheduler(spec, max_tiles_per_generation=1)
    srv = serve.Server(sch)
    rep = serve.serve_poisson(srv, reqs, np.linspace(0.0, 0.2, len(reqs)))
    assert rep["finished"] == len(reqs) and sch.gen >= sum(-(-r.n_walkers // 256) for r in reqs)
    ref = {r.key: r for r in batch.run_batched([batch.prepare_entry(r) for r in reqs], spec)}
    for r in srv.results:
        assert r.n_escaped == 0 and r.n_error == 0
        assert _same(r, ref[r.key])


def test_walks_longer_than_the_lds_grid_run_the_streamed_walk_with_the_same_results():
    """A 300 ms walk on a 1 ms grid has 301 samples, more than the split walk
    keeps in LDS (256): the default launch and a graph replay run the streamed
    walk instead, with the per-tile walk's results bit for bit."""
    req = batch.library_requests(1, seed=41, walkers=(3000, 3000), cells=(100, 100), T_max_ms=300.0, ts=0.1)[0]
    assert req.config().n_grid == 301
    spec = batch.acquisition_spec(n_dirs=3, delta=6.0, Deltas=(15.0, 250.0), b_values=(0.0, 1000.0))
    prep = batch.prepare_entry(req)
    lookup = lambda r: prep  # noqa: E731
    say("long grid: default (split requested) and per-tile plain runs, then a graph-replayed batch")
    default = batch.run_plain([req], spec, prepare=lookup)[0]
    tile = batch.run_plain([req], spec, prepare=lookup, classify_mode=1)[0]
    batched = batch.run_batched([prep], spec)[0]
    assert default.occupancy.shape == (301,)
    assert _same(default, tile) and _same(default, batched)


class _WalkCfg:
    def __init__(self, steps, sph, G, sigma, ts):
        self.n_steps, self.steps_per_h, self.n_grid, self.sigma, self.ts = steps, sph, G, sigma, ts


def test_geometry_release_is_refused_while_a_kernel_reads_it():
    """madi_geom_destroy_async frees nothing (MADI_ERR_BUSY) while the last
    kernel launched on the geometry is still queued or running, and frees it
    once that kernel has finished.  The walk and the free are on one stream,
    so the free is ordered behind the walk either way."""
    import ctypes as C
    import torch
    from custom_madi_amd import _hip
    from tests import _geom
    dg = DeviceGeometry(geometry_arrays(_geom.cellular(L=150.0, seed=9)))
    n, cfg = 65_536, _WalkCfg(2000, 50, 41, np.sqrt(2 * 2.0 * 0.02), 0.02)
    p = engine.walk_params(n, cfg, 0.03, 5)
    Y = torch.empty((n, cfg.n_grid, 3), dtype=torch.float64, device="cuda")
    itr = torch.empty((n, cfg.n_grid), dtype=torch.int8, device="cuda")
    esc = torch.empty(n, dtype=torch.int8, device="cuda")
    err = torch.empty(n, dtype=torch.int8, device="cuda")
    lib = _hip.load_library()
    s = _hip.stream_ptr()
    torch.cuda.synchronize()
    say("[batch] FP64 walk launched, release requested before it finishes")
    _hip.check(lib.madi_walk_f64(dg.handle, C.byref(p), None, None, None, _hip.dptr(Y), _hip.dptr(itr),
                                 _hip.dptr(esc), _hip.dptr(err), None, s), "madi_walk_f64")
    with pytest.raises(RuntimeError, match="has not finished"):
        dg.release(s)
    assert dg.handle is not None and dg.handle.value
    torch.cuda.synchronize()
    dg.release(s)
    assert dg.handle is None
    torch.cuda.synchronize()
