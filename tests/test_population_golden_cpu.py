"""The population, annulus and KD build against the reference's own outputs.

tests/golden/population_vectors.npz holds, for three seeded cases, what the
reference's _make_population / alpha_star_from_vi / _build_gpu_kdtree
(madi/ensemble.py:405-638) return (written by
tests/golden/make_population_vectors.py).  The same generator seeds must give
bit-identical seeds (the numpy PCG64 call order: Poisson count, then uniform
coordinates), nearest-neighbour distances, certificate, annulus and KD
arrays here, and leave the generator in the same state.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from custom_madi_amd import ensemble as E
from custom_madi_amd.config import SimConfig

GOLD = Path(__file__).resolve().parent / "golden" / "population_vectors.npz"


@pytest.mark.parametrize("case", [0, 1, 2])
def test_population_equals_the_reference(case):
    g = np.load(GOLD)
    rho, side, seed, vi = g[f"case{case}_params"]
    cfg = SimConfig()
    rng = np.random.default_rng(int(seed))
    seeds, nn, cert = E._make_population(E._rho_um3(rho), float(side), cfg, rng)
    assert np.array_equal(seeds, g[f"case{case}_seeds"])
    assert np.array_equal(nn, g[f"case{case}_nn"])
    c = g[f"case{case}_cert"]
    got = [cert.pop_margin_um, cert.s10_distance_upper_um, float(cert.s10_pass), float(cert.n_relevant_seeds),
           cert.s11_min_clearance_margin_um, float(cert.s11_pass), cert.certification_spacing_um,
           float(cert.expansion_index)]
    assert np.array_equal(np.array(got), c)
    annulus = np.minimum(E.alpha_star_from_vi(float(vi), float(rho)), cfg.kappa * nn / 2.0)
    assert np.array_equal(annulus, g[f"case{case}_annulus"])
    for name, arr in zip(("seed", "axis", "left", "right", "parent"), E.build_kdtree(seeds)):
        assert np.array_equal(arr, g[f"case{case}_kd_{name}"]), name
    assert np.array_equal(rng.uniform(0.0, 1.0, size=4), g[f"case{case}_next_uniform"])
