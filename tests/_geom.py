"""Small seeded geometries for tests (no GPU needed to build them)."""

from __future__ import annotations

import numpy as np
from scipy.spatial import cKDTree

from custom_madi_amd.ensemble import (
    Ensemble, GeometryStats, PopulationCertificate, alpha_star_from_vi, build_kdtree,
    periodic_nearest_distances,
)


def _stats(side):
    cert = PopulationCertificate(0.0, 0.0, True, 0, float("inf"), True, 0.0, 0)
    return GeometryStats(0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0, 0, 0, side, side,
                         cert, "", {})


def cellular(L=60.0, rho_per_uL=5.0e5, vi=0.6, seed=1, margin_spacings=3.0, alpha=None, mean_av=0.6):
    """Non-periodic SI-style box: seeds in [-m, L+m)^3, annulus min(alpha*, 0.45 d_nn)."""
    rng = np.random.default_rng(seed)
    rho = rho_per_uL * 1e-9
    m = margin_spacings * rho ** (-1.0 / 3.0)
    n = rng.poisson(rho * (L + 2 * m) ** 3)
    seeds = rng.uniform(-m, L + m, size=(n, 3))
    nn = cKDTree(seeds).query(seeds, k=2)[0][:, 1]
    a_star = alpha_star_from_vi(vi, rho_per_uL) if alpha is None else alpha
    ann = np.minimum(a_star, 0.45 * nn)
    nodes = build_kdtree(seeds)
    src = 0.4 * L
    return Ensemble(seeds=seeds, annulus=ann, rho=rho_per_uL, V=vi / rho / 1e3, vi=vi, alpha_star=a_star, L=L,
                    source_lo=0.5 * (L - src), source_hi=0.5 * (L + src), mean_AV=mean_av, geometry=_stats(L),
                    rho_requested=rho_per_uL, V_requested=vi / rho / 1e3, kd_node_seed=nodes[0],
                    kd_node_axis=nodes[1], kd_node_left=nodes[2], kd_node_right=nodes[3], kd_node_parent=nodes[4])


def periodic(n_cells=64, rho_per_uL=6.0e5, vi=0.6, seed=0, mean_av=0.6):
    """Periodic cube of n_cells uniform seeds (the synthetic benchmark boxes)."""
    rho = rho_per_uL * 1e-9
    L = (n_cells / rho) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    seeds = rng.uniform(0.0, L, size=(n_cells, 3))
    a_star = alpha_star_from_vi(vi, rho_per_uL)
    ann = np.minimum(a_star, 0.45 * periodic_nearest_distances(seeds, L))
    nodes = build_kdtree(seeds)
    return Ensemble(seeds=seeds, annulus=ann, rho=rho_per_uL, V=vi / rho / 1e3, vi=vi, alpha_star=a_star, L=L,
                    source_lo=0.0, source_hi=L, mean_AV=mean_av, geometry=_stats(L), rho_requested=rho_per_uL,
                    V_requested=vi / rho / 1e3, kd_node_seed=nodes[0], kd_node_axis=nodes[1],
                    kd_node_left=nodes[2], kd_node_right=nodes[3], kd_node_parent=nodes[4], periodic=True)


def lattice(n=4, h=2.0, jitter=1e-3, annulus=0.9, seed=0):
    """Periodic cube of n^3 seeds on a slightly jittered cubic lattice of spacing h (Voronoi cells are near-cubes
    of side h); every annulus is `annulus` until the caller changes it."""
    rng = np.random.default_rng(seed)
    L = n * h
    ijk = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), axis=-1).reshape(-1, 3)
    seeds = (ijk + 0.5) * h + rng.uniform(-jitter, jitter, size=(n ** 3, 3))
    ann = np.full(n ** 3, float(annulus))
    nodes = build_kdtree(seeds)
    rho = n ** 3 / L ** 3 * 1e9
    return Ensemble(seeds=seeds, annulus=ann, rho=rho, V=h ** 3 * 1e-3, vi=0.5, alpha_star=float(annulus), L=L,
                    source_lo=0.0, source_hi=L, mean_AV=0.6, geometry=_stats(L), rho_requested=rho,
                    V_requested=h ** 3 * 1e-3, kd_node_seed=nodes[0], kd_node_axis=nodes[1],
                    kd_node_left=nodes[2], kd_node_right=nodes[3], kd_node_parent=nodes[4], periodic=True)


def reference_walk_golden():
    """(ensemble, cfg, p_p, fixture) of tests/golden/walk_vectors.npz: the
    reference's Eq. S2 population for generator seed 21 and its own CPU walk
    of 32 walkers on it (tests/golden/make_walk_vectors.py)."""
    from pathlib import Path

    from custom_madi_amd.config import SimConfig

    g = np.load(Path(__file__).resolve().parent / "golden" / "walk_vectors.npz")
    (side, lo, hi, pp, D0, ts, T_max, h, _sigma, _n_steps, _sph, _n_grid, n_walkers, _stream_seed, rho,
     vi) = g["scalars"]
    cfg = SimConfig(D0=float(D0), ts=float(ts), T_max_ms=float(T_max), h_ms=float(h), n_walkers=int(n_walkers),
                    n_ensembles=1, small_deltas=[0.4], big_deltas=[1.2], b_values=[0.0, 1000.0])
    seeds, ann = g["seeds"], g["annulus"]
    nodes = (g["kd_seed"], g["kd_axis"], g["kd_left"], g["kd_right"], g["kd_parent"])
    a_star = alpha_star_from_vi(float(vi), float(rho))
    ens = Ensemble(seeds=seeds, annulus=ann, rho=float(rho), V=float("nan"), vi=float("nan"), alpha_star=a_star,
                   L=float(side), source_lo=float(lo), source_hi=float(hi), mean_AV=1.0, geometry=_stats(float(side)),
                   rho_requested=float(rho), V_requested=float("nan"), kd_node_seed=nodes[0],
                   kd_node_axis=nodes[1], kd_node_left=nodes[2], kd_node_right=nodes[3], kd_node_parent=nodes[4])
    return ens, cfg, float(pp), g


def stream_digest(stream) -> str:
    import hashlib
    h = hashlib.sha256()
    for a in (stream.initial_positions, stream.increments, stream.acceptance_uniforms):
        h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    return h.hexdigest()
