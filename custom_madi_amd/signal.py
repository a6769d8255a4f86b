"""(delta, Delta, b) columns and signal assembly (madi/signal.py:38-264).

The finite-lobe moment of a walker is dM = Y(delta) + Y(Delta) - Y(Delta +
delta) [µm ms]; the phase of column (delta, Delta, b) is
gamma * G(b, delta, Delta) * dM * 1e-9 and the stored signal is the mean of
its cosine over walkers and the three Cartesian axes.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from .config import GAMMA_RAD, SimConfig, grid_time_index
from .walker_gpu import ReducedResult, run_simulation_multi_kio_reduced, run_simulation_reduced


def G_from_b(b_s_mm2: float, delta_ms: float, Delta_ms: float) -> float:
    """Gradient [T/m] for b [s/mm²], delta, Delta [ms] (madi/signal.py:38-45)."""
    b_si = b_s_mm2 * 1e6
    d_si = delta_ms * 1e-3
    tD_si = (Delta_ms - delta_ms / 3.0) * 1e-3
    if b_si <= 0 or tD_si <= 0:
        return 0.0
    return float(np.sqrt(b_si / ((GAMMA_RAD * d_si) ** 2 * tD_si)))


@dataclass
class ColumnGrid:
    """Flattened column arrays, pair-major then b (madi/signal.py:52-72)."""

    delta_pairs: List[Tuple[float, float]]
    b_values: np.ndarray
    j_delta: np.ndarray
    j_Delta: np.ndarray
    j_sum: np.ndarray
    phase_coef: np.ndarray
    n_pairs: int
    n_b: int
    phase_model: str = "finite_lobe"
    narrow_snapshot_steps: Optional[np.ndarray] = None
    narrow_snapshot_index: Optional[np.ndarray] = None
    narrow_phase_coef: Optional[np.ndarray] = None
    narrow_tD_requested_ms: Optional[np.ndarray] = None
    narrow_tD_sampled_ms: Optional[np.ndarray] = None


def build_columns(cfg: Optional[SimConfig] = None, delta_pairs=None, b_values=None) -> ColumnGrid:
    """Column grid of cfg (or explicit pairs / b values) (madi/signal.py:75-150)."""
    if cfg is None:
        cfg = SimConfig()
    pairs = list(delta_pairs) if delta_pairs is not None else cfg.delta_pairs()
    bvals = np.asarray(b_values if b_values is not None else cfg.b_values, dtype=float)
    n_pairs, n_b = len(pairs), len(bvals)
    j_delta = np.empty(n_pairs * n_b, dtype=np.int32)
    j_Delta = np.empty_like(j_delta)
    j_sum = np.empty_like(j_delta)
    pc = np.empty(n_pairs * n_b, dtype=np.float64)
    tD_steps = np.empty(n_pairs * n_b, dtype=np.int32)
    narrow_pc = np.empty(n_pairs * n_b, dtype=np.float64)
    tD_req = np.empty(n_pairs * n_b, dtype=np.float64)
    tD_smp = np.empty(n_pairs * n_b, dtype=np.float64)
    for pi, (delta, Delta) in enumerate(pairs):
        jd = grid_time_index(delta, cfg.h_ms)
        jD = grid_time_index(Delta, cfg.h_ms)
        js = grid_time_index(Delta + delta, cfg.h_ms)
        tD = float(Delta - delta / 3.0)
        step = int(round(tD / cfg.ts))
        if step < 0 or step > cfg.n_steps:
            raise ValueError(f"narrow-pulse t_D={tD:g} ms lies outside the {cfg.T_max_ms:g} ms walk")
        for bi, b in enumerate(bvals):
            col = pi * n_b + bi
            j_delta[col], j_Delta[col], j_sum[col] = jd, jD, js
            pc[col] = GAMMA_RAD * G_from_b(float(b), delta, Delta) * 1e-9
            tD_steps[col] = step
            narrow_pc[col] = pc[col] * float(delta)
            tD_req[col] = tD
            tD_smp[col] = step * cfg.ts
    extra = dict(narrow_snapshot_steps=None, narrow_snapshot_index=None, narrow_phase_coef=None,
                 narrow_tD_requested_ms=None, narrow_tD_sampled_ms=None)
    if cfg.phase_model == "narrow_pulse":
        steps, index = np.unique(tD_steps, return_inverse=True)
        extra = dict(narrow_snapshot_steps=steps, narrow_snapshot_index=index, narrow_phase_coef=narrow_pc,
                     narrow_tD_requested_ms=tD_req, narrow_tD_sampled_ms=tD_smp)
    elif cfg.phase_model != "finite_lobe":
        raise ValueError(f"unknown phase_model {cfg.phase_model!r}")
    return ColumnGrid(delta_pairs=pairs, b_values=bvals, j_delta=j_delta, j_Delta=j_Delta, j_sum=j_sum,
                      phase_coef=pc, n_pairs=n_pairs, n_b=n_b, phase_model=cfg.phase_model, **extra)


def _assemble(res: ReducedResult, columns: ColumnGrid) -> dict:
    """Signal block and diagnostics of one entry (madi/signal.py:157-195)."""
    n_eff = res.n_eff
    S = (res.cos_sum / n_eff).reshape(columns.n_pairs, columns.n_b)
    S_imag = (res.sin_sum / n_eff).reshape(columns.n_pairs, columns.n_b)
    if len(res.ensemble_cos_means) != len(res.ensemble_sin_means):
        raise RuntimeError("reduction returned mismatched real/imaginary ensemble means")
    if res.ensemble_cos_means:
        e = len(res.ensemble_cos_means)
        ens_S = np.asarray(res.ensemble_cos_means, dtype=np.float64).reshape(e, columns.n_pairs, columns.n_b)
        ens_I = np.asarray(res.ensemble_sin_means, dtype=np.float64).reshape(e, columns.n_pairs, columns.n_b)
    else:
        ens_S = ens_I = None
    return {
        "delta_pairs": columns.delta_pairs, "b_values": columns.b_values, "S": S, "S_imag": S_imag,
        "ensemble_S": ens_S, "ensemble_S_imag": ens_I, "phase_model": columns.phase_model, "n_eff": n_eff,
        "n_escaped": res.n_escaped, "kio_analytic_eq5": res.analytic_kio_eq5, "pp": res.mean_pp,
        "occupancy_fraction": res.occupancy_fraction, "geometry_stats": res.geometry_stats,
    }


def compute_signals(rho, V, kio, cfg: Optional[SimConfig] = None, columns: Optional[ColumnGrid] = None,
                    seed: int = 0, verbose: bool = True) -> dict:
    if cfg is None:
        cfg = SimConfig()
    if columns is None:
        columns = build_columns(cfg)
    res = run_simulation_reduced(
        rho, V, kio, columns.j_delta, columns.j_Delta, columns.j_sum, columns.phase_coef, cfg, seed=seed,
        verbose=verbose, narrow_snapshot_steps=columns.narrow_snapshot_steps,
        narrow_snapshot_index=columns.narrow_snapshot_index, narrow_phase_coef=columns.narrow_phase_coef)
    return _assemble(res, columns)


def compute_signals_multi_kio(rho, V, kios, cfg: Optional[SimConfig] = None, columns: Optional[ColumnGrid] = None,
                              seed: int = 0, verbose: bool = True) -> dict:
    if cfg is None:
        cfg = SimConfig()
    if columns is None:
        columns = build_columns(cfg)
    results = run_simulation_multi_kio_reduced(
        rho, V, kios, columns.j_delta, columns.j_Delta, columns.j_sum, columns.phase_coef, cfg, seed=seed,
        verbose=verbose, narrow_snapshot_steps=columns.narrow_snapshot_steps,
        narrow_snapshot_index=columns.narrow_snapshot_index, narrow_phase_coef=columns.narrow_phase_coef)
    return {k: _assemble(r, columns) for k, r in results.items()}


def compute_adc(b_values, signal_1d, b_adc: float = 1000.0) -> float:
    idx = int(np.argmin(np.abs(np.asarray(b_values) - b_adc)))
    S = signal_1d[idx]
    b_int = b_values[idx] / 1e6
    if S <= 0 or b_int <= 0:
        return 0.0
    return float(-np.log(S) / b_int)


def signals_to_flat(result: dict) -> np.ndarray:
    return result["S"].ravel()
