"""Library building and the .npz v5 on-disk format (build side of madi/library.py).

``build_library_from_triplets`` / ``build_library`` keep the reference's
signatures (madi/library.py:445-670) and file schema (madi/library.py:872-1120):
entries keyed by rounded (k_io, rho, V), grouped by (rho, V) so geometry is
reused across the k_io sweep, one checkpoint per group, and the v5
Monte-Carlo diagnostics (imaginary signal, between-ensemble variance, the
declared 200-column per-ensemble subset).  The walks run in the HIP kernels
through ``signal.compute_signals_multi_kio``.  The MAP matchers that read a
library live in ``match.py``.
"""

from __future__ import annotations

import json
import os
import time
from collections import defaultdict
from dataclasses import dataclass, field
from typing import List, Mapping, Optional, Tuple

import numpy as np

from . import signal as sig
from .config import (ENSEMBLE_MEAN_SUBSET_B_VALUES_S_MM2, ENSEMBLE_MEAN_SUBSET_DELTA_PAIRS_MS, SimConfig)

MADI_LIBRARY_SCHEMA_V5 = "madi-library-v5"
MADI_LIBRARY_SCHEMA_V4 = "madi-library-v4"
MADI_LIBRARY_ENTRY_SCHEMA_V5 = "madi-library-entry-v5"
ENSEMBLE_MEAN_SUBSET_N_COLUMNS = len(ENSEMBLE_MEAN_SUBSET_DELTA_PAIRS_MS) * len(ENSEMBLE_MEAN_SUBSET_B_VALUES_S_MM2)

# Legacy fixed-delta acquisition, used only to describe pre-(delta,Delta,b) files.
DELTA_SMALL = 20.0
DELTAS_BIG = [50.0]
BVALS_UNIQUE = np.array([500, 1000, 1500, 2000, 2500])


@dataclass
class LibraryEntry:
    """One library entry (fields of madi/library.py:76-103)."""

    kio: float
    rho: float
    V: float
    vector: np.ndarray
    signal_imag: Optional[np.ndarray] = None
    signal_variance: Optional[np.ndarray] = None
    ensemble_means_subset: Optional[np.ndarray] = None
    vi: Optional[float] = None
    weight: Optional[float] = None
    kio_nominal: Optional[float] = None
    rho_nominal: Optional[float] = None
    V_nominal: Optional[float] = None
    pp: Optional[float] = None
    kio_analytic_eq5: Optional[float] = None
    is_free_water: bool = False
    metadata: dict = field(default_factory=dict)

    @property
    def realised_vi(self) -> float:
        if self.vi is not None and np.isfinite(self.vi):
            return float(self.vi)
        return float(self.rho) * float(self.V) * 1e-6


DEFAULT_KIOS = [2, 5, 8, 12, 18, 25, 35, 50, 75, 100]
DEFAULT_RHOS = [100_000, 200_000, 400_000, 600_000, 800_000]
DEFAULT_VS = [0.5, 1.0, 2.0, 3.0, 4.0, 5.0]
VI_HARD_MAX = 0.99


def _cell_widths(values, lower_bound=None, upper_bound=None) -> np.ndarray:
    """Midpoint-cell widths of a strictly increasing grid (madi/library.py:186-201)."""
    v = np.asarray(values, dtype=float)
    if v.ndim != 1 or len(v) < 2 or np.any(np.diff(v) <= 0.0):
        raise ValueError("grid axis must be strictly increasing with at least two values")
    lo = float(lower_bound) if lower_bound is not None else v[0] - 0.5 * (v[1] - v[0])
    hi = float(upper_bound) if upper_bound is not None else v[-1] + 0.5 * (v[-1] - v[-2])
    if lo > v[0] or hi < v[-1]:
        raise ValueError("quadrature bounds must enclose every grid node")
    edges = np.concatenate(([lo], 0.5 * (v[:-1] + v[1:]), [hi]))
    return np.diff(edges)


def _entry_key(kio, rho, V):
    if not np.isfinite(kio) or (rho == 0.0 and V == 0.0):
        return ("free_water", 0.0, 0.0)
    return (round(float(kio), 4), round(float(rho), 1), round(float(V), 6))


def _existing_keys(library):
    return {_entry_key(e.kio_nominal if e.kio_nominal is not None else e.kio,
                       e.rho_nominal if e.rho_nominal is not None else e.rho,
                       e.V_nominal if e.V_nominal is not None else e.V) for e in library}


@dataclass(frozen=True)
class LogGridDefinition:
    """Masked log grid plus a free-water atom (madi/library.py:109-183)."""

    kios: np.ndarray
    rhos: np.ndarray
    Vs: np.ndarray
    vi_min: float = 0.40
    vi_max: float = 0.99
    free_water_weight: float = 1.0

    def triplets_and_weights(self):
        if np.any(self.rhos <= 0.0) or np.any(self.Vs <= 0.0):
            raise ValueError("log-coordinate rho and V grid values must be positive")
        if np.any(self.kios < 0.0):
            raise ValueError("k_io grid values must be non-negative")
        lx, ly = np.log(self.rhos), np.log(self.Vs)
        dx = _cell_widths(lx, float(lx[0]), float(lx[-1]))
        dy = _cell_widths(ly, float(ly[0]), float(ly[-1]))
        dk = _cell_widths(self.kios, 0.0, float(self.kios[-1]))
        triplets = [(0.0, 0.0, 0.0)]
        weights = {_entry_key(0.0, 0.0, 0.0): float(self.free_water_weight)}
        for ik, k in enumerate(self.kios):
            for ir, rho in enumerate(self.rhos):
                for iv, V in enumerate(self.Vs):
                    vi = float(rho * V * 1e-6)
                    if self.vi_min <= vi <= self.vi_max:
                        t = (float(k), float(rho), float(V))
                        triplets.append(t)
                        weights[_entry_key(*t)] = float(rho * V * dx[ir] * dy[iv] * dk[ik])
        return triplets, weights

    def metadata(self) -> dict:
        return {
            "coordinate_transform": {"rho": "uniform_log", "V": "uniform_log", "kio": "piecewise_uniform_linear"},
            "vi_min": float(self.vi_min), "vi_max": float(self.vi_max), "rho_values": self.rhos.tolist(),
            "V_values": self.Vs.tolist(), "kio_values": self.kios.tolist(),
            "free_water_weight": float(self.free_water_weight),
            "weight_expression": "rho * V * dlogrho * dlogV * dkio",
        }


def make_remediation_log_grid(*, n_rho=64, n_V=64, kios=None, rho_min=1.0e4, rho_max=1.0e7, V_min=0.01,
                              V_max=200.0, vi_min=0.40, vi_max=0.99) -> LogGridDefinition:
    """Production log grid (madi/library.py:204-228)."""
    if not (0.0 < rho_min < rho_max):
        raise ValueError("remediation rho bounds must satisfy 0 < min < max")
    if not (0.0 < V_min < V_max):
        raise ValueError("remediation V bounds must satisfy 0 < min < max")
    if kios is None:
        kios = np.r_[0.0, np.arange(1.0, 31.0, 1.0), np.arange(35.0, 131.0, 5.0)]
    return LogGridDefinition(np.asarray(kios, dtype=float), np.geomspace(rho_min, rho_max, int(n_rho)),
                             np.geomspace(V_min, V_max, int(n_V)), float(vi_min), float(vi_max))


def _filter_valid(triplets, vi_min=0.0, vi_max=VI_HARD_MAX):
    hi = min(vi_max, VI_HARD_MAX)
    out = []
    for k, r, v in triplets:
        if r == 0.0 and v == 0.0:
            out.append((k, r, v))
        elif vi_min <= (r / 1e9) * (v * 1e3) <= hi:
            out.append((k, r, v))
    return out


def _summarise_realised_geometry(per_ensemble: list, cfg: SimConfig) -> dict:
    """Aggregate realised geometry of an entry's ensembles (madi/library.py:269-309)."""
    if not per_ensemble:
        return {"vi": 0.0, "rho_per_uL": 0.0, "mean_volume_pL": 0.0, "governing_mean_A_over_V_um_inv": 0.0,
                "n_seeds_sim": 0}
    cells = np.asarray([int(x["n_seeds_sim"]) for x in per_ensemble], dtype=float)
    vis = np.asarray([float(x["realised_vi"]) for x in per_ensemble])
    vols = np.asarray([float(x["sim_side_um"]) ** 3 for x in per_ensemble])
    tc, tv = float(cells.sum()), float(vols.sum())
    ti = float(np.sum(vis * vols))
    rho = tc / tv * 1e9 if tv else 0.0
    V_pL = ti / tc / 1e3 if tc else 0.0
    w = vols / max(tv, 1.0)

    def vw(name):
        return float(np.sum(w * np.asarray([float(x[name]) for x in per_ensemble])))

    return {
        "vi": float(ti / tv) if tv else 0.0,
        "vi_between_ensemble_sd": float(np.std(vis, ddof=1)) if len(vis) > 1 else 0.0,
        "rho_per_uL": float(rho), "mean_volume_pL": float(V_pL), "mean_volume_um3": float(V_pL * 1e3),
        "governing_mean_A_over_V_um_inv": vw("governing_mean_A_over_V_um_inv"),
        "annulus_mean_um": vw("annulus_mean_um"), "annulus_std_um": vw("annulus_std_um"),
        "annulus_min_um": float(min(float(x["annulus_min_um"]) for x in per_ensemble)),
        "annulus_max_um": float(max(float(x["annulus_max_um"]) for x in per_ensemble)),
        "annulus_q05_um": vw("annulus_q05_um"), "annulus_q50_um": vw("annulus_q50_um"),
        "annulus_q95_um": vw("annulus_q95_um"), "n_seeds_sim": int(tc),
        "sim_side_um": [float(x["sim_side_um"]) for x in per_ensemble],
        "source_side_um": [float(x["source_side_um"]) for x in per_ensemble],
    }


def ensemble_mean_subset_column_indices(columns: "sig.ColumnGrid") -> np.ndarray:
    """Full-vector indices of the declared v5 subset, schema order (madi/library.py:312-348)."""
    expected = np.asarray(ENSEMBLE_MEAN_SUBSET_B_VALUES_S_MM2, dtype=float)
    actual = np.asarray(columns.b_values, dtype=float)
    if actual.shape != expected.shape or not np.array_equal(actual, expected):
        raise ValueError("madi-library-v5 requires the declared 25-point b grid for the per-ensemble covariance "
                         "diagnostic subset")
    where = {(float(d), float(D)): i for i, (d, D) in enumerate(columns.delta_pairs)}
    if len(where) != len(columns.delta_pairs):
        raise ValueError("main storage grid contains duplicate (delta, Delta) pairs")
    missing = [p for p in ENSEMBLE_MEAN_SUBSET_DELTA_PAIRS_MS if p not in where]
    if missing:
        raise ValueError(f"madi-library-v5 main storage grid omits declared ensemble-mean diagnostic pairs: {missing!r}")
    return np.asarray([where[p] * columns.n_b + b for p in ENSEMBLE_MEAN_SUBSET_DELTA_PAIRS_MS
                       for b in range(columns.n_b)], dtype=np.int64)


def _v5_diagnostics_from_result(result: dict, columns: "sig.ColumnGrid", cfg: SimConfig):
    """signal_imag, signal_variance, ensemble subset, imaginary check (madi/library.py:351-438)."""
    er, ei = result.get("ensemble_S"), result.get("ensemble_S_imag")
    if er is None or ei is None:
        raise RuntimeError("v5 library build requires per-ensemble reduced signals")
    er = np.asarray(er, dtype=np.float64)
    ei = np.asarray(ei, dtype=np.float64)
    shape = (int(cfg.n_ensembles), columns.n_pairs, columns.n_b)
    if er.shape != shape or ei.shape != shape:
        raise RuntimeError("per-ensemble reduced signal shape does not match the configured ensemble/grid "
                           f"contract: got real={er.shape}, imag={ei.shape}, expected={shape}")
    if not (np.all(np.isfinite(er)) and np.all(np.isfinite(ei))):
        raise RuntimeError("per-ensemble reduced signals contain non-finite values")
    if int(cfg.n_ensembles) < 2:
        raise ValueError("madi-library-v5 requires at least two ensembles for sample variance")
    imag = np.asarray(result["S_imag"], dtype=np.float64).reshape(columns.n_pairs, columns.n_b)
    var = np.var(er, axis=0, ddof=1)
    subset = er.reshape(int(cfg.n_ensembles), -1)[:, ensemble_mean_subset_column_indices(columns)]
    mean_i = np.mean(ei, axis=0)
    se_i = np.std(ei, axis=0, ddof=1) / np.sqrt(cfg.n_ensembles)
    z = np.zeros_like(mean_i)
    pos = se_i > 0.0
    z[pos] = np.abs(mean_i[pos] / se_i[pos])
    zero_nz = (~pos) & (np.abs(mean_i) > 0.0)
    z[zero_nz] = np.inf
    k = int(np.argmax(z))
    pi, bi = np.unravel_index(k, z.shape)
    check = {
        "method": "absolute Student statistic of the mean of independent per-ensemble imaginary signals (ddof=1)",
        "n_ensembles": int(cfg.n_ensembles), "degrees_of_freedom": int(cfg.n_ensembles - 1),
        "max_abs_standardized_deviation": float(z.flat[k]),
        "max_column": {"delta_ms": float(columns.delta_pairs[pi][0]), "Delta_ms": float(columns.delta_pairs[pi][1]),
                       "b_s_mm2": float(columns.b_values[bi])},
        "mean_imaginary_signal": float(mean_i.flat[k]), "standard_error": float(se_i.flat[k]),
        "zero_standard_error_nonzero_count": int(np.count_nonzero(zero_nz)),
    }
    return (imag.astype(np.float32).ravel(), var.astype(np.float32).ravel(), subset.astype(np.float32), check)


def _entry_from_result(kio, rho, V, result, columns, cfg, entry_weights):
    vec = sig.signals_to_flat(result)
    imag, var, subset, check = _v5_diagnostics_from_result(result, columns, cfg)
    summary = _summarise_realised_geometry(result["geometry_stats"], cfg)
    free = rho == 0.0 and V == 0.0
    key = _entry_key(kio, rho, V)
    weight = None if entry_weights is None else entry_weights.get(key)
    meta = {
        "schema": MADI_LIBRARY_ENTRY_SCHEMA_V5,
        "nominal": {"kio": float(kio), "rho": float(rho), "V": float(V)},
        "realised_geometry": summary,
        "per_ensemble_geometry": result["geometry_stats"],
        "exchange": {
            "kio_analytic_eq5_s_inv": float(result["kio_analytic_eq5"]), "p_p_mean": float(result["pp"]),
            "definition": "MADI I Eq. 5 using SI §S.IV untrimmed governing-process arithmetic <A/V>; "
                          "no residence-time proxy is used.",
        },
        # a periodic box (this build's synthetic torus; the reference has no
        # periodic mode) has no exit, so nothing can escape and be fatal
        "boundary": {"mode": "periodic_torus" if cfg.periodic else cfg.boundary_mode,
                     "n_escaped": int(result["n_escaped"]),
                     "n_walkers_total": int(result["n_eff"] // 3), "fatal_on_escape": not cfg.periodic,
                     "occupancy_fraction": np.asarray(result["occupancy_fraction"]).tolist()},
        "weight": None if weight is None else float(weight),
        "imaginary_signal_check": check,
    }
    return LibraryEntry(
        kio=float("nan") if free else float(kio), rho=0.0 if free else summary["rho_per_uL"],
        V=0.0 if free else summary["mean_volume_pL"], vector=vec, signal_imag=imag, signal_variance=var,
        ensemble_means_subset=subset, vi=0.0 if free else summary["vi"],
        weight=None if weight is None else float(weight), kio_nominal=float(kio), rho_nominal=float(rho),
        V_nominal=float(V), pp=float(result["pp"]), kio_analytic_eq5=float(result["kio_analytic_eq5"]),
        is_free_water=free, metadata=meta,
    )


def _gpu_present() -> bool:
    try:
        import torch
        return bool(torch.cuda.is_available())
    except ImportError:
        return False


def build_library_from_triplets(triplets: List[Tuple[float, float, float]], cfg: SimConfig | None = None,
                                save_path: Optional[str] = None, existing_library: Optional[list] = None,
                                seed: int = 0, vi_min: float = 0.0, vi_max: float = VI_HARD_MAX,
                                entry_weights: Optional[Mapping[tuple, float]] = None,
                                grid_metadata: Optional[dict] = None, verbose: bool = True,
                                batched: Optional[bool] = None, pool=None) -> list:
    """Build or extend a library from (k_io, rho, V) triplets (madi/library.py:445-642).

    ``batched`` (default: whenever the configuration allows it) runs every
    walk of every group on the batched device scheduler (library_batch),
    with the same entries, order and checkpoints as the group-by-group loop;
    ``pool`` (a worker pool from serve.make_pool) builds the ensembles'
    CPU geometry in worker processes; their v_i measurement (the HIP
    classifier) runs in this process."""
    if cfg is None:
        cfg = SimConfig()
    if int(cfg.n_ensembles) < 2:
        raise ValueError("madi-library-v5 requires at least two independently constructed ensembles so that "
                         "per-column sample variance is defined")
    library = list(existing_library) if existing_library is not None else []
    done = _existing_keys(library)
    if library and not all(e.signal_imag is not None and e.signal_variance is not None
                           and e.ensemble_means_subset is not None for e in library):
        raise ValueError("cannot append v5 entries to a pre-v5 library: the existing artifact lacks required "
                         "Monte-Carlo diagnostic arrays")
    valid = _filter_valid(triplets, vi_min=vi_min, vi_max=vi_max)
    todo = [t for t in valid if _entry_key(*t) not in done]
    if verbose:
        print(f"  Requested: {len(triplets)} triplets; new to compute: {len(todo)}")
    if not todo:
        if save_path:
            _save_library(library, save_path, cfg=cfg, grid_metadata=grid_metadata, build_seed=seed)
        return library
    columns = sig.build_columns(cfg)
    groups = defaultdict(list)
    for k, r, v in todo:
        groups[(r, v)].append(k)
    t0 = time.time()
    ordered = sorted(groups.items(), key=lambda kv: kv[0][0] * kv[0][1])
    from . import library_batch
    if batched is None:
        batched = library_batch.batchable(cfg, columns) and _gpu_present()
    if batched:
        cp = library_batch.LibraryCheckpoint(save_path, cfg, columns, library, grid_metadata=grid_metadata,
                                             build_seed=seed)
        try:
            library_batch.build_groups_batched([(rv, sorted(ks)) for rv, ks in ordered], cfg, columns, seed, cp,
                                               entry_weights=entry_weights, pool=pool, verbose=verbose)
        finally:
            library[:] = cp.library
        if verbose:
            print(f"Library: {len(library)} entries total ({len(todo)} new in {time.time() - t0:.0f}s, batched)")
        return library
    for (rho, V), kios in ordered:
        kios = sorted(kios)
        results = sig.compute_signals_multi_kio(rho, V, kios, cfg, columns=columns, seed=seed, verbose=False)
        for kio in kios:
            library.append(_entry_from_result(kio, rho, V, results[kio], columns, cfg, entry_weights))
        if save_path:
            _save_library(library, save_path, cfg=cfg, columns=columns, grid_metadata=grid_metadata,
                          build_seed=seed)
    if verbose:
        print(f"Library: {len(library)} entries total ({len(todo)} new in {time.time() - t0:.0f}s)")
    return library


def build_library(kios=None, rhos=None, Vs=None, cfg=None, save_path=None, existing_library=None, seed=0,
                  vi_min=0.0, vi_max=VI_HARD_MAX, entry_weights=None, grid_metadata=None, verbose=True) -> list:
    """Full cross-product grid (madi/library.py:649-670)."""
    kios = DEFAULT_KIOS if kios is None else kios
    rhos = DEFAULT_RHOS if rhos is None else rhos
    Vs = DEFAULT_VS if Vs is None else Vs
    triplets = [(k, r, v) for k in kios for r in rhos for v in Vs]
    return build_library_from_triplets(triplets, cfg=cfg, save_path=save_path, existing_library=existing_library,
                                       seed=seed, vi_min=vi_min, vi_max=vi_max, entry_weights=entry_weights,
                                       grid_metadata=grid_metadata, verbose=verbose)


# ----------------------------------------------------------------------- I/O

def _json_default(value):
    if isinstance(value, np.generic):
        return value.item()
    if isinstance(value, np.ndarray):
        return value.tolist()
    raise TypeError(f"cannot JSON-serialise {type(value)!r}")


def _cfg_metadata(cfg: SimConfig, *, schema: str = MADI_LIBRARY_SCHEMA_V5, build_seed: Optional[int] = None) -> dict:
    """Build metadata embedded in the artifact (keys of madi/library.py:685-801)."""
    meta = {
        "schema": schema, "D0_um2_ms": cfg.D0, "ts_ms": cfg.ts, "T_max_ms": cfg.T_max_ms,
        "domain_sizing": "MADI I SI Eq. S8; per-entry W unless diagnostic L override",
        "diagnostic_domain_L_override_um": cfg.L, "source_fraction": cfg.source_fraction,
        "boundary_mode": cfg.boundary_mode, "kappa": cfg.kappa,
        "population_certification": {
            "equations": ["S10", "S11", "S12"],
            "initial_margin_cell_spacings": cfg.population_initial_margin_cell_spacings,
            "certification_spacing_um": cfg.population_certification_spacing_um,
            "max_expansions": cfg.population_max_expansions,
        },
        "geometry_reference": {
            "path": cfg.geometry_reference_path, "required_single_cells": cfg.geometry_reference_required_cells,
            "required_alpha_values": cfg.geometry_reference_required_alpha_values,
            "allow_uncertified": cfg.allow_uncertified_geometry_reference,
            "mean_estimator": "untrimmed_arithmetic_mean_A_over_V",
        },
        "geometry_validation_points": cfg.geometry_validation_points,
        "geometry_vi_tolerance": cfg.geometry_vi_tolerance, "walkers_per_ensemble": cfg.n_walkers,
        "ensembles_per_entry": cfg.n_ensembles, "axis_walks_per_entry": 3 * int(cfg.n_walkers) * int(cfg.n_ensembles),
        "build_seed": None if build_seed is None else int(build_seed), "phase_model": cfg.phase_model,
        "checkpoint_h_ms": cfg.h_ms, "common_random_numbers": "base_seed + ensemble_index; shared across rho,V,kio",
        "classifier": "exact full-facet shifted-Voronoi contraction at every endpoint; SI Eq. S2 "
                      + ("(FP64 KD classifier)" if cfg.precision == "fp64"
                         else "(FP32 Delaunay-graph classifier certified in FP64)"),
        "classifier_cache": {"mode": cfg.classifier_mode, "delta_max_um": float(cfg.classifier_cache_delta_max_um),
                             "min_safe_radius_um": float(cfg.classifier_cache_min_safe_radius_um),
                             "candidate_capacity": int(cfg.classifier_cache_candidate_capacity)},
        "walk_precision": cfg.precision, "periodic_geometry": bool(cfg.periodic),
        "kio_label": "MADI I Eq. 5 with governing-process untrimmed arithmetic <A/V>",
        "residence_time_measurement": "intentionally absent: SI §S.IV states k_io != 1/<tau_i>",
    }
    if schema == MADI_LIBRARY_SCHEMA_V5:
        meta["uncertainty"] = {
            "signal_imag": {"array": "signal_imag", "dtype": "float32", "shape": ["entry", "column"],
                            "definition": "collapsed mean(sin(phase)) over all axis-walks"},
            "between_ensemble_variance": {
                "array": "signal_variance", "dtype": "float32", "shape": ["entry", "column"],
                "definition": "sample variance (ddof=1) of per-ensemble mean real signals; consumer SE is "
                              "sqrt(signal_variance / n_ensembles)", "n_ensembles": int(cfg.n_ensembles)},
            "ensemble_means_subset": {
                "array": "ensemble_means_subset", "dtype": "float32",
                "shape": ["entry", "ensemble_index", "declared_column"],
                "declared_delta_Delta_pairs_ms": [[float(d), float(D)] for d, D in ENSEMBLE_MEAN_SUBSET_DELTA_PAIRS_MS],
                "declared_b_values_s_mm2": np.asarray(ENSEMBLE_MEAN_SUBSET_B_VALUES_S_MM2, dtype=float).tolist(),
                "column_order": "pair-major, then b-major within each pair"},
            "ensemble_index_ordering_contract": {
                "axis_position_in_ensemble_means_subset": 1, "index_values": "0..n_ensembles-1",
                "same_order_across_entries": True, "independent_of": ["rho", "V", "k_io"],
                "walk_seed_formula": "(build_seed + 104729 * ensemble_index) mod 2^31",
                "geometry_seed_formula": "(build_seed + 97003 * ensemble_index) mod 2^31",
                "purpose": "ensemble index e is the common-random-number partner across entries"},
        }
    return meta


def _v5_entry_arrays(lib: list, n_columns: int):
    """Stack v5 diagnostics when every entry has them (madi/library.py:804-869)."""
    have = [e.signal_imag is not None and e.signal_variance is not None and e.ensemble_means_subset is not None
            for e in lib]
    if not any(have):
        return None
    if not all(have):
        raise ValueError("library mixes entries with and without v5 Monte-Carlo diagnostics")
    imag, var, sub = [], [], []
    n_ens = None
    for i, e in enumerate(lib):
        a, b, c = np.asarray(e.signal_imag), np.asarray(e.signal_variance), np.asarray(e.ensemble_means_subset)
        if a.shape != (n_columns,):
            raise ValueError(f"entry {i} signal_imag shape {a.shape} != ({n_columns},)")
        if b.shape != (n_columns,):
            raise ValueError(f"entry {i} signal_variance shape {b.shape} != ({n_columns},)")
        if c.ndim != 2 or c.shape[1] != ENSEMBLE_MEAN_SUBSET_N_COLUMNS:
            raise ValueError(f"entry {i} ensemble_means_subset shape {c.shape} != "
                             f"(n_ensembles, {ENSEMBLE_MEAN_SUBSET_N_COLUMNS})")
        n_ens = c.shape[0] if n_ens is None else n_ens
        if c.shape[0] != n_ens:
            raise ValueError("v5 entries disagree on the number of ensembles")
        if n_ens < 2:
            raise ValueError("v5 diagnostic arrays require at least two ensembles")
        if not (np.all(np.isfinite(a)) and np.all(np.isfinite(b)) and np.all(np.isfinite(c))):
            raise ValueError(f"entry {i} has non-finite v5 diagnostic values")
        if np.any(b < 0.0):
            raise ValueError(f"entry {i} has negative between-ensemble variance")
        imag.append(a.astype(np.float32))
        var.append(b.astype(np.float32))
        sub.append(c.astype(np.float32))
    return np.stack(imag), np.stack(var), np.stack(sub)


def _save_library(lib: list, path: str, cfg: SimConfig | None = None, columns=None, grid_metadata=None,
                  build_metadata=None, build_seed=None):
    """Write the library .npz (keys of madi/library.py:872-972)."""
    if cfg is None:
        cfg = SimConfig()
    if columns is None:
        columns = sig.build_columns(cfg)
    n_cols = int(columns.n_pairs * columns.n_b)
    v5 = _v5_entry_arrays(lib, n_cols)
    schema = MADI_LIBRARY_SCHEMA_V5 if v5 is not None else MADI_LIBRARY_SCHEMA_V4

    def col(fn, dtype=float):
        return np.array([fn(e) for e in lib], dtype=dtype)

    meta = dict(build_metadata) if build_metadata is not None else {
        **_cfg_metadata(cfg, schema=schema, build_seed=build_seed), "grid": grid_metadata or {}}
    if meta.get("schema") != schema:
        raise ValueError(f"library-schema mismatch between entries and build metadata: entries imply {schema!r}, "
                         f"metadata has {meta.get('schema')!r}")
    if grid_metadata is not None:
        meta["grid"] = grid_metadata
    weights = col(lambda e: np.nan if e.weight is None else float(e.weight))
    payload = dict(
        library_schema=np.array(schema), kios=col(lambda e: e.kio), rhos=col(lambda e: e.rho),
        Vs=col(lambda e: e.V), vis=col(lambda e: e.realised_vi), vectors=np.array([e.vector for e in lib]),
        nominal_kios=col(lambda e: e.kio if e.kio_nominal is None else e.kio_nominal),
        nominal_rhos=col(lambda e: e.rho if e.rho_nominal is None else e.rho_nominal),
        nominal_Vs=col(lambda e: e.V if e.V_nominal is None else e.V_nominal),
        weights=weights, has_weights=np.array(bool(np.all(np.isfinite(weights)))),
        pps=col(lambda e: np.nan if e.pp is None else e.pp),
        kio_analytic_eq5=col(lambda e: np.nan if e.kio_analytic_eq5 is None else e.kio_analytic_eq5),
        is_free_water=col(lambda e: e.is_free_water, np.bool_),
        entry_metadata_json=np.array([json.dumps(e.metadata, sort_keys=True, default=_json_default) for e in lib]),
        build_metadata_json=np.array(json.dumps(meta, sort_keys=True, default=_json_default)),
        pair_deltas=np.array([d for d, _ in columns.delta_pairs], dtype=float),
        pair_Deltas=np.array([D for _, D in columns.delta_pairs], dtype=float),
        b_values=np.asarray(columns.b_values, dtype=float), n_b=np.array(columns.n_b), h_ms=np.array(float(cfg.h_ms)),
    )
    if v5 is not None:
        payload.update(
            signal_imag=v5[0], signal_variance=v5[1], ensemble_means_subset=v5[2],
            ensemble_subset_pair_deltas=np.asarray([d for d, _ in ENSEMBLE_MEAN_SUBSET_DELTA_PAIRS_MS], dtype=float),
            ensemble_subset_pair_Deltas=np.asarray([D for _, D in ENSEMBLE_MEAN_SUBSET_DELTA_PAIRS_MS], dtype=float),
            ensemble_subset_b_values=np.asarray(ENSEMBLE_MEAN_SUBSET_B_VALUES_S_MM2, dtype=float),
            ensemble_subset_n_b=np.array(len(ENSEMBLE_MEAN_SUBSET_B_VALUES_S_MM2)),
        )
    np.savez(path, **payload)


def load_library(path: str) -> list:
    """Read a library .npz (madi/library.py:975-1040)."""
    data = np.load(path, allow_pickle=False)
    kios, rhos, Vs, vectors = data["kios"], data["rhos"], data["Vs"], data["vectors"]
    n = len(kios)

    def opt(name, default):
        return data[name] if name in data.files else default

    vis = opt("vis", None)
    weights = opt("weights", np.full(n, np.nan))
    nk, nr, nv = opt("nominal_kios", kios), opt("nominal_rhos", rhos), opt("nominal_Vs", Vs)
    pps = opt("pps", np.full(n, np.nan))
    eq5 = opt("kio_analytic_eq5", np.full(n, np.nan))
    free = opt("is_free_water", np.zeros(n, dtype=bool))
    ej = opt("entry_metadata_json", np.asarray(["{}"] * n))
    si, sv, sm = opt("signal_imag", None), opt("signal_variance", None), opt("ensemble_means_subset", None)
    diag = (si, sv, sm)
    if any(x is not None for x in diag) and not all(x is not None for x in diag):
        raise ValueError(f"{path}: incomplete v5 Monte-Carlo diagnostic arrays")
    if si is not None and (si.shape[0] != n or sv.shape[0] != n or sm.shape[0] != n):
        raise ValueError(f"{path}: v5 Monte-Carlo diagnostic entry axis has wrong length")
    out = []
    for i in range(n):
        try:
            meta = json.loads(str(ej[i]))
        except (TypeError, ValueError):
            meta = {}
        vi = None if vis is None or not np.isfinite(vis[i]) else float(vis[i])
        out.append(LibraryEntry(
            kio=float(kios[i]), rho=float(rhos[i]), V=float(Vs[i]), vector=vectors[i],
            signal_imag=None if si is None else si[i], signal_variance=None if sv is None else sv[i],
            ensemble_means_subset=None if sm is None else sm[i], vi=vi,
            weight=None if not np.isfinite(weights[i]) else float(weights[i]), kio_nominal=float(nk[i]),
            rho_nominal=float(nr[i]), V_nominal=float(nv[i]), pp=None if not np.isfinite(pps[i]) else float(pps[i]),
            kio_analytic_eq5=None if not np.isfinite(eq5[i]) else float(eq5[i]), is_free_water=bool(free[i]),
            metadata=meta,
        ))
    return out


def load_library_meta(path: str) -> dict:
    """Grid metadata of a library file, new or legacy format (madi/library.py:1043-1120)."""
    data = np.load(path, allow_pickle=False)
    meta = {}
    if "library_schema" in data.files:
        meta["library_schema"] = str(data["library_schema"])
    meta["has_weights"] = bool(data["has_weights"]) if "has_weights" in data.files else False
    if "build_metadata_json" in data.files:
        try:
            meta["build_metadata"] = json.loads(str(data["build_metadata_json"]))
        except (TypeError, ValueError):
            meta["build_metadata"] = None
    if "pair_deltas" in data.files and "pair_Deltas" in data.files:
        meta["format"] = "v2"
        meta["delta_pairs"] = list(zip(np.asarray(data["pair_deltas"], dtype=float).tolist(),
                                       np.asarray(data["pair_Deltas"], dtype=float).tolist()))
        meta["n_b"] = int(data["n_b"]) if "n_b" in data.files else None
        meta["b_values"] = list(np.asarray(data["b_values"], dtype=float)) if "b_values" in data.files else None
        meta["h_ms"] = float(data["h_ms"]) if "h_ms" in data.files else None
        meta["has_free_water"] = bool(np.any(data["is_free_water"])) if "is_free_water" in data.files else False
        names = {"signal_imag", "signal_variance", "ensemble_means_subset", "ensemble_subset_pair_deltas",
                 "ensemble_subset_pair_Deltas", "ensemble_subset_b_values", "ensemble_subset_n_b"}
        meta["has_v5_diagnostics"] = names.issubset(set(data.files))
        if meta["has_v5_diagnostics"]:
            meta["signal_imag_shape"] = tuple(data["signal_imag"].shape)
            meta["signal_variance_shape"] = tuple(data["signal_variance"].shape)
            meta["ensemble_means_subset_shape"] = tuple(data["ensemble_means_subset"].shape)
            meta["ensemble_subset_delta_pairs"] = list(zip(
                np.asarray(data["ensemble_subset_pair_deltas"], dtype=float).tolist(),
                np.asarray(data["ensemble_subset_pair_Deltas"], dtype=float).tolist()))
            meta["ensemble_subset_b_values"] = list(np.asarray(data["ensemble_subset_b_values"], dtype=float))
            meta["ensemble_subset_n_b"] = int(data["ensemble_subset_n_b"])
        return meta
    meta["format"] = "legacy"
    deltas = list(data["deltas"]) if "deltas" in data.files else list(DELTAS_BIG)
    n_b = int(data["n_b"]) if "n_b" in data.files else len(BVALS_UNIQUE)
    small = float(data["small_delta"]) if "small_delta" in data.files else DELTA_SMALL
    bv = (list(np.asarray(data["b_values"], dtype=float)) if "b_values" in data.files
          else (list(BVALS_UNIQUE.astype(float)) if n_b == len(BVALS_UNIQUE) else None))
    meta.update(delta_pairs=[(small, D) for D in deltas], n_b=n_b, b_values=bv, small_delta=small, h_ms=None,
                has_v5_diagnostics=False)
    return meta


def merge_library_shards(paths: List[str], out_path: str, require_shards: Optional[int] = None,
                         verbose: bool = True) -> list:
    """Merge per-shard libraries into one file (scripts/merge_shards.py:54-132).

    Every shard must carry the first shard's (delta, Delta) pairs, b values,
    schema, build seed, grid metadata and v5 uncertainty contract.  Entries
    keep shard order; a (k_io, rho, V) key already taken (rounded to 4, 1 and
    6 decimals; the free-water atom is one key) is skipped, first occurrence
    winning.  With require_shards, the file names must carry shard ids
    0..require_shards-1 (``.shard<k>``) exactly once.  The merged file keeps
    the first shard's build metadata and grid."""
    import re
    if not paths:
        raise ValueError("no shard files given")
    if require_shards is not None:
        found = []
        for p in paths:
            m = re.search(r"\.shard(\d+)(?:\.|$)", os.path.basename(p))
            if m is None:
                raise ValueError(f"cannot infer shard id from {p}")
            found.append(int(m.group(1)))
        if set(found) != set(range(require_shards)) or len(found) != len(set(found)):
            raise ValueError(f"shard coverage mismatch; found={sorted(found)}, expected={list(range(require_shards))}")
    first = load_library_meta(paths[0])
    fb = first.get("build_metadata") or {}

    def contract(meta):
        b = meta.get("build_metadata") or {}
        return (meta["delta_pairs"], meta["b_values"], meta.get("library_schema"), b.get("build_seed"),
                json.dumps(b.get("grid", {}), sort_keys=True), json.dumps(b.get("uncertainty", {}), sort_keys=True))

    want = contract(first)
    merged, seen = [], set()
    for p in paths:
        if contract(load_library_meta(p)) != want:
            raise ValueError(f"shard metadata/schema/uncertainty mismatch: {p}")
        lib = load_library(p)
        before = len(merged)
        for e in lib:
            key = ("free_water",) if e.is_free_water else (round(e.kio, 4), round(e.rho, 1), round(e.V, 6))
            if key in seen:
                continue
            seen.add(key)
            merged.append(e)
        if verbose:
            print(f"  {p}: +{len(merged) - before} new  ({len(lib)} in file)")
    pairs = [tuple(x) for x in first["delta_pairs"]]
    empty = np.array([], dtype=np.int32)
    columns = sig.ColumnGrid(delta_pairs=pairs, b_values=np.asarray(first["b_values"], dtype=float), j_delta=empty,
                             j_Delta=empty, j_sum=empty, phase_coef=np.array([], dtype=np.float64),
                             n_pairs=len(pairs), n_b=int(first["n_b"]))
    cfg = SimConfig(h_ms=first.get("h_ms") or 1.0)
    os.makedirs(os.path.dirname(os.path.abspath(out_path)) or ".", exist_ok=True)
    _save_library(merged, out_path, cfg=cfg, columns=columns, build_metadata=fb or None)
    return merged


def library_summary(lib: list, meta: dict | None = None):
    """Print a short summary (madi/library.py:1123-1161)."""
    if not lib:
        print("  (empty library)")
        return
    cell = [e for e in lib if not e.is_free_water]
    print(f"  Entries: {len(lib)}")
    print(f"  kio  ({len({e.kio for e in cell})}), rho ({len({e.rho for e in cell})}), "
          f"V ({len({e.V for e in cell})})")
    print(f"  free-water atom: {any(e.is_free_water for e in lib)}")
    print(f"  Vector length: {lib[0].vector.size}")
    if meta is not None:
        print(f"  Format: {meta.get('format', '?')}, {len(meta['delta_pairs'])} pairs x {meta['n_b']} b-values")
