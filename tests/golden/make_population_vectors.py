"""Writes population_vectors.npz: the reference's Omega_pop sampling, annulus
and KD build (madi/ensemble.py:525-638) for three seeded cases.

The fixture pins what no closed form or invariant does: for a given generator
seed the reference draws a Poisson count and then uniform seeds
(madi/ensemble.py:600-604), measures nearest-neighbour distances and the Eq.
S10 grid bound with cKDTree, expands the margin until S10-S11 certify, caps
the annulus at min(alpha*, kappa d_nn / 2) and builds the balanced KD tree.
Run it in a container that holds the reference tree at /root/reference (it
imports the reference's CPU code from there); the tests read only the .npz.
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, "/root/reference")
from madi import ensemble as RE  # noqa: E402
from madi.config import SimConfig  # noqa: E402

OUT = Path(__file__).resolve().parent / "population_vectors.npz"
# (rho cells/uL, Omega_sim side um, generator seed, v_i for alpha*)
CASES = [(5.0e5, 30.0, 11, 0.6), (2.0e6, 20.0, 12, 0.8), (8.0e4, 60.0, 13, 0.5)]


def main() -> None:
    cfg = SimConfig()
    out = {}
    for i, (rho, side, seed, vi) in enumerate(CASES):
        rng = np.random.default_rng(seed)
        seeds, _, nn, cert = RE._make_population(RE._rho_um3(rho), side, cfg, rng)
        alpha = RE.alpha_star_from_vi(vi, rho)
        annulus = np.minimum(alpha, cfg.kappa * nn / 2.0)
        nodes = RE._build_gpu_kdtree(seeds)
        out[f"case{i}_params"] = np.array([rho, side, seed, vi])
        out[f"case{i}_seeds"] = seeds
        out[f"case{i}_nn"] = nn
        out[f"case{i}_annulus"] = annulus
        out[f"case{i}_cert"] = np.array([cert.pop_margin_um, cert.s10_distance_upper_um, float(cert.s10_pass),
                                         float(cert.n_relevant_seeds), cert.s11_min_clearance_margin_um,
                                         float(cert.s11_pass), cert.certification_spacing_um,
                                         float(cert.expansion_index)])
        for name, arr in zip(("seed", "axis", "left", "right", "parent"), nodes):
            out[f"case{i}_kd_{name}"] = np.asarray(arr)
        # the generator's next draw after the population (the v_i measurement
        # draws from this state)
        out[f"case{i}_next_uniform"] = rng.uniform(0.0, 1.0, size=4)
        print(f"case {i}: {len(seeds)} seeds, margin {cert.pop_margin_um:.3f}, expansion {cert.expansion_index}")
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
