"""The oracle walk and reduce against the reference's own CPU walk.

tests/golden/walk_vectors.npz holds the reference's _walk_cpu Y, inside
trace and escape flags (madi/walker_gpu.py:698-840) and its _reduce_cpu
column sums (madi/walker_gpu.py:954-968) for 32 walkers x 160 steps on the
reference's seeded Eq. S2 population, driven by make_walk_random_stream
(madi/walker_gpu.py:115-127).  Here the same stream is regenerated with this
package's make_walk_random_stream (its SHA-256 must equal the reference's),
the oracle walks it, and its Y must agree within the reference's CPU/GPU
golden tolerance (rtol 5e-11, atol 5e-12, scripts/run_gpu_golden_check.py),
its inside trace exactly.
"""

from __future__ import annotations

import numpy as np

from custom_madi_amd import walker_gpu
from custom_madi_amd.ensemble import geometry_arrays
from oracle import oracle as O
from tests import _geom


def test_stream_and_geometry_are_the_reference_s():
    ens, cfg, pp, g = _geom.reference_walk_golden()
    stream = walker_gpu.make_walk_random_stream(cfg, int(g["scalars"][13]), ensemble=ens)
    assert _geom.stream_digest(stream) == str(g["stream_sha256"])
    assert cfg.n_steps == int(g["scalars"][9]) and cfg.steps_per_h == int(g["scalars"][10])
    assert cfg.n_grid == int(g["scalars"][11]) and cfg.sigma == g["scalars"][8]


def test_oracle_walk_equals_the_reference_cpu_walk():
    ens, cfg, pp, g = _geom.reference_walk_golden()
    stream = walker_gpu.make_walk_random_stream(cfg, int(g["scalars"][13]), ensemble=ens)
    og = O.Geometry.from_arrays(geometry_arrays(ens))
    Y, itr, esc, err = O.walk(og, cfg.n_walkers, cfg.n_steps, cfg.steps_per_h, cfg.n_grid, cfg.sigma, cfg.ts, pp,
                              ens.source_lo, ens.source_hi,
                              stream=(stream.initial_positions, stream.increments, stream.acceptance_uniforms))
    assert not err.any()
    assert np.array_equal(esc.astype(bool), g["escaped"].astype(bool))
    assert np.array_equal(itr, g["inside_trace"])
    assert np.allclose(Y, g["Y"], rtol=5e-11, atol=5e-12)
    # the case crosses membranes: walkers change between inside and outside
    assert np.count_nonzero(np.diff(g["inside_trace"].astype(int), axis=1)) >= 10
    cs, sn = O.reduce(g["Y"], g["cols_jd"], g["cols_jD"], g["cols_js"], g["cols_pc"])
    assert np.allclose(cs, g["cos"], rtol=1e-12, atol=1e-12) and np.allclose(sn, g["sin"], rtol=1e-12, atol=1e-12)


def test_cpu_baseline_restatement_equals_the_reference_cpu_walk():
    """oracle.walk_cpu_reference (the CPU algorithm the benchmark times as the
    reference's CPU fallback) gives the reference's _walk_cpu Y and inside
    trace bit for bit on the golden case."""
    ens, cfg, pp, g = _geom.reference_walk_golden()
    stream = walker_gpu.make_walk_random_stream(cfg, int(g["scalars"][13]), ensemble=ens)
    Y, itr, esc = O.walk_cpu_reference(ens.seeds, ens.annulus, ens.L, ens.source_lo, ens.source_hi, cfg.n_walkers,
                                       cfg.n_steps, cfg.steps_per_h, cfg.n_grid, cfg.sigma, cfg.ts, pp,
                                       stream=(stream.initial_positions, stream.increments,
                                               stream.acceptance_uniforms))
    assert np.array_equal(Y, g["Y"]) and np.array_equal(itr, g["inside_trace"])
    assert np.array_equal(esc.astype(bool), g["escaped"].astype(bool))
