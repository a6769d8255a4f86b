"""The benchmark configurations at their own sizes on an MI355X.

* C1 (1,000 cells, 1e6 walkers x 2,300 steps, 480 signals): the FP32
  production walk against the FP64 exact-classifier walk (the oracle run on
  the device) on the same geometry and seeds -- per (Delta, b) |dS| <= 2e-3
  and every signal within 3 Monte Carlo standard errors (the reference's
  sample SE, tests/physics_audit/test_finite_delta_reference.py:24-33),
  occupancy within 3 binomial standard errors.
* C2 (256 uneven entries): the batched form equals the plain
  one-entry-per-launch form bit for bit for every entry; four seeded entries
  against the FP64 walk as for C1.
* C3 (Poisson arrivals at 20 /s, a 10 s window): every request finishes,
  entries overlap on the device, and each served result equals its plain
  one-entry run bit for bit.
* C4 (full-range sweep): k_io = 0, v_i = 0.994 at 0.019 pL and the
  smallest volume 0.01 pL (both dt 0.01 ms), the largest volume, the box
  with the most cells and two seeded entries, batched, each against the
  FP64 walk.

The standard error of a signal S(b) = mean cos(phi) comes from the signal at
4b, the mean of cos(2 phi): Var(cos phi) = (1 + S(4b)) / 2 - S(b)^2
(workloads.acquisition_with_variance).
"""

from __future__ import annotations

import sys

import numpy as np
import pytest

from custom_madi_amd import batch, engine, serve, workloads
from custom_madi_amd.ensemble import DeviceGeometry, create_periodic_ensemble, geometry_arrays

pytestmark = pytest.mark.gpu

DS_MAX = 2e-3


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _check_against_fp64(S32, cos64, n, idx, idx4, what):
    """|dS| <= 2e-3 per signal and <= 3 SE (SE of the FP64 estimate)."""
    S64, se = workloads.signal_se(cos64, n, idx, idx4)
    dS = np.abs(S32 - S64)
    ratio = np.where(se > 0, dS / np.where(se > 0, se, 1.0), np.where(dS > 0, np.inf, 0.0))
    say(f"{what}: max |dS| {dS.max():.3e}, max |dS|/SE {ratio.max():.3f}, per (Delta, b) max "
        f"{np.round(dS.max(axis=2), 6).tolist()}")
    assert dS.max() <= DS_MAX, what
    assert np.all(dS <= 3.0 * se), what
    return dS


def _fp64_run(req, prep_pp, specv):
    ens = create_periodic_ensemble(req.n_cells, req.rho, req.rho * req.V * 1e-6, req.config(), seed=req.seed)
    dg = DeviceGeometry(geometry_arrays(ens))
    p = engine.walk_params(req.n_walkers, req.config(), prep_pp, req.walk_seed)
    return engine.walk_fused(dg, p, specv, precision=64)


def test_c1_fp32_walk_matches_the_fp64_walk_at_full_size():
    cfg, ens, pp, spec = workloads.c1_entry()
    N = cfg.n_walkers
    dg = DeviceGeometry(geometry_arrays(ens))
    specv, idx, idx4 = workloads.acquisition_with_variance()
    p = engine.walk_params(N, cfg, pp, seed=0)
    say(f"[c1] FP32 walk, {N} walkers x {cfg.n_steps} steps, 480 signals")
    r480 = engine.walk_fused(dg, p, spec, precision=32)
    r32 = engine.walk_fused(dg, p, specv, precision=32)
    # the variance columns leave the acquisition's own columns untouched
    assert np.array_equal(r32.cos[:, idx, :], r480.cos) and np.array_equal(r32.sin[:, idx, :], r480.sin)
    assert r480.n_escaped == 0 and r480.n_error == 0
    say("[c1] FP64 walk on the same seeds")
    r64 = engine.walk_fused(dg, p, specv, precision=64)
    assert r64.n_escaped == 0 and r64.n_error == 0
    _check_against_fp64(r480.cos / N, r64.cos, N, idx, idx4, "c1")
    f32, f64 = r480.occupancy / N, r64.occupancy / N
    occ_se = np.sqrt(np.maximum(f64 * (1.0 - f64), 0.0) / N)
    say(f"[c1] occupancy max |d| {np.abs(f32 - f64).max():.2e} (binomial SE {occ_se.max():.2e})")
    assert np.all(np.abs(f32 - f64) <= 3.0 * occ_se + 1.0 / N)


@pytest.fixture(scope="module")
def c2():
    reqs = workloads.c2_requests()
    say(f"[c2] preparing {len(reqs)} entries")
    preps = [batch.prepare_entry(r) for r in reqs]
    return reqs, preps


def test_c2_every_batched_entry_equals_its_plain_run(c2):
    reqs, preps = c2
    spec = workloads.acquisition()
    lookup = {p.req.key: p for p in preps}
    say("[c2] plain form, one entry per launch")
    plain = batch.run_plain(reqs, spec, prepare=lambda r: lookup[r.key])
    say("[c2] batched form")
    batched = {r.key: r for r in batch.run_batched(preps, spec)}
    assert len(plain) == len(batched) == 256
    same = sum(int(np.array_equal(r.cos, batched[r.key].cos) and np.array_equal(r.sin, batched[r.key].sin)
                   and np.array_equal(r.occupancy, batched[r.key].occupancy)) for r in plain)
    say(f"[c2] bitwise equal: {same}/{len(plain)}")
    assert same == len(plain)

    specv, idx, idx4 = workloads.acquisition_with_variance()
    for i in np.sort(np.random.default_rng(2).choice(len(reqs), size=4, replace=False)):
        r = reqs[int(i)]
        say(f"[c2] FP64 check of entry {r.key}: {r.n_cells} cells, {r.n_walkers} walkers")
        r64 = _fp64_run(r, lookup[r.key].pp, specv)
        assert r64.n_escaped == 0 and r64.n_error == 0
        _check_against_fp64(batched[r.key].cos / r.n_walkers, r64.cos, r.n_walkers, idx, idx4, f"c2 {r.key}")


def test_c3_served_arrivals_equal_their_plain_runs():
    reqs, arrivals = workloads.c3_requests(20.0, duration_s=10.0)
    say(f"[c3] preparing {len(reqs)} requests")
    lookup = {r.key: batch.prepare_entry(r) for r in reqs}
    spec = workloads.acquisition()
    sch = batch.DeviceScheduler(spec)
    srv = serve.Server(sch, prepare=lambda r: lookup[r.key])
    say(f"[c3] serving {len(reqs)} Poisson arrivals over 10 s")
    rep = serve.serve_poisson(srv, reqs, arrivals)
    sch.close()
    say(f"[c3] finished {rep['finished']}/{rep['requests']}, p50 {rep['latency_s_p50']:.3f} s, "
        f"p99 {rep['latency_s_p99']:.3f} s, {sch.gen} generations")
    assert rep["finished"] == len(reqs) and rep["failed"] == 0
    # continuous batching: entries were on the device together
    ev = sorted([(r.admitted, 1) for r in srv.results] + [(r.finished, -1) for r in srv.results])
    live = np.cumsum([d for _, d in ev])
    say(f"[c3] most entries on the device at once: {live.max()}")
    assert live.max() > 1
    served = {r.key: r for r in srv.results}
    plain = batch.run_plain(reqs, spec, prepare=lambda r: lookup[r.key])
    for r in plain:
        s = served[r.key]
        assert np.array_equal(r.cos, s.cos) and np.array_equal(r.sin, s.sin), r.key
        assert np.array_equal(r.occupancy, s.occupancy), r.key


def _c4_pick(reqs):
    """Seeded C4 sample holding the sweep's extremes: k_io = 0; v_i = 0.994 at
    its smallest volume (0.019 pL, dt 0.01 ms); the smallest volume, 0.01 pL
    (v_i 0.5: higher v_i there exceeds the sweep's rho range); the largest
    volume; the box with the most cells."""
    vi = np.array([r.rho * r.V * 1e-6 for r in reqs])
    V = np.array([r.V for r in reqs])
    kio = np.array([r.kio for r in reqs])
    cells = np.array([r.n_cells for r in reqs])
    picks = [
        int(np.flatnonzero((kio == 0.0) & (np.abs(vi - 0.747) < 1e-3))[0]),
        int(np.flatnonzero((np.abs(vi - 0.994) < 1e-3) & (kio > 100.0))[0]),
        int(np.flatnonzero((V == V.min()) & (kio > 60.0))[0]),
        int(np.flatnonzero(V == V.max())[-1]),
        int(np.argmax(cells * (kio > 50.0))),
    ]
    rng = np.random.default_rng(4)
    picks += [int(i) for i in rng.choice(len(reqs), size=2, replace=False)]
    return [reqs[i] for i in dict.fromkeys(picks)]


def test_c4_sweep_extremes_match_the_fp64_walk():
    reqs = _c4_pick(workloads.c4_requests())
    spec = workloads.acquisition()
    specv, idx, idx4 = workloads.acquisition_with_variance()
    for r in reqs:
        say(f"[c4] entry {r.key}: v_i {r.rho * r.V * 1e-6:.3f}, {r.n_cells} cells, dt {r.ts} ms, "
            f"{r.n_walkers} walkers")
    assert any(r.kio == 0.0 for r in reqs) and any(r.ts == 0.01 for r in reqs)
    preps = [batch.prepare_entry(r) for r in reqs]
    res = {x.key: x for x in batch.run_batched(preps, spec)}
    assert len(res) == len(reqs)
    for r, prep in zip(reqs, preps):
        r64 = _fp64_run(r, prep.pp, specv)
        assert r64.n_escaped == 0 and r64.n_error == 0
        _check_against_fp64(res[r.key].cos / r.n_walkers, r64.cos, r.n_walkers, idx, idx4, f"c4 {r.key}")
