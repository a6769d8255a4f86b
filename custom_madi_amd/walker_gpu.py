"""Drop-in walk API of madi/walker_gpu.py, running on MI355X HIP kernels.

Public names and signatures follow the reference (madi/walker_gpu.py:50-1233).
The walk itself runs in libmadi_hip.so:

* ``cfg.precision == "fp64"`` (default): the reference-semantics walk -- the
  exact FP64 KD classifier at every endpoint, the reference's step order and
  trapezoid Y -- with the phase reduction either fused into the walk kernel
  (column sets of at most 8 (delta, Delta) pairs x 16 b values) or run as a
  separate FP64 column reduction over the device-resident Y.
* ``cfg.precision == "fp32"``: the production walk (FP32 seed-relative state,
  Delaunay-graph classifier certified in FP64, compensated Y, FP64 sums).

There is no CPU walk: ``use_gpu=False`` raises, as does a missing GPU or
library.  ``classifier_mode="exact_cached"`` (the reference's
equivalence-preserving shortcut) runs the FP64 walk with graph labels
(``MADI_CLASSIFY_FP64_CACHED``): endpoints are labelled from the Delaunay
graph in FP64 from the walker's own cell, a label also certifies a ball that
later endpoints take it from, and decisions within rounding of a threshold
go to the exact KD classifier, so the results are the exact walk's bit for
bit.
"""

from __future__ import annotations

import math
import sys
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field, replace
from typing import Optional, Sequence, Tuple

import numpy as np

from . import _hip, engine
from .config import PRODUCTION_AXIS_WALKS_PER_ENTRY, SimConfig
from .ensemble import (
    Ensemble, build_ensemble_host, create_dummy_ensemble, create_ensemble, create_periodic_ensemble, finish_ensemble,
    geometry_arrays,
)

PAPER_WALKS_PER_ENTRY = PRODUCTION_AXIS_WALKS_PER_ENTRY


def _gpu_available() -> bool:
    try:
        import torch
        if not torch.cuda.is_available():
            return False
        _hip.load_library()
        return True
    except Exception:
        return False


def __getattr__(name):
    # HAS_CUDA / HAS_GPU: the reference's "GPU path is available" flag
    # (madi/walker_gpu.py:39-41), evaluated on first use so that importing
    # this module (for example in host-only worker processes) does not
    # initialise the GPU runtime.
    if name in ("HAS_CUDA", "HAS_GPU"):
        value = _gpu_available()
        globals()["HAS_CUDA"] = globals()["HAS_GPU"] = value
        return value
    raise AttributeError(name)


def pp_to_kio_eq5(pp: float, mean_AV: float, cfg: SimConfig) -> float:
    """MADI I Eq. 5 [s^-1] (madi/walker_gpu.py:50-52)."""
    return float(pp) * math.sqrt(cfg.D0 / (6.0 * cfg.ts)) * float(mean_AV) * 1000.0


def kio_to_pp(kio: float, mean_AV: float, cfg: SimConfig) -> float:
    """Inverse of Eq. 5; raises outside [0, 1] (madi/walker_gpu.py:55-62)."""
    if kio < 0.0:
        raise ValueError(f"k_io must be non-negative, got {kio}")
    if mean_AV <= 0.0:
        raise ValueError("governing-process mean_A_over_V must be positive")
    return _checked_pp((float(kio) / 1000.0) / (math.sqrt(cfg.D0 / (6.0 * cfg.ts)) * mean_AV))


def pp_to_kio(pp: float, mean_AV: float, cfg: SimConfig) -> float:
    return pp_to_kio_eq5(pp, mean_AV, cfg)


def _checked_pp(pp: float) -> float:
    pp = float(pp)
    if not np.isfinite(pp) or pp < 0.0 or pp > 1.0:
        raise ValueError(f"p_p must be finite and in [0, 1], got {pp!r}")
    return pp


def _ensemble_seed(base_seed: int, ensemble_index: int) -> int:
    """Geometry CRN seed (madi/walker_gpu.py:77-78)."""
    return (int(base_seed) + ensemble_index * 97_003) & 0x7FFFFFFF


def _walk_seed(base_seed: int, ensemble_index: int) -> int:
    """Walk CRN seed (madi/walker_gpu.py:81-83)."""
    return (int(base_seed) + ensemble_index * 104_729) & 0x7FFFFFFF


@dataclass(frozen=True)
class WalkRandomStream:
    """Injected walk inputs (madi/walker_gpu.py:86-112)."""

    initial_positions: np.ndarray
    increments: np.ndarray
    acceptance_uniforms: np.ndarray

    def validate(self, cfg: SimConfig, ensemble: Optional[Ensemble] = None) -> None:
        n = int(cfg.n_walkers)
        if np.asarray(self.initial_positions).shape != (n, 3):
            raise ValueError("golden initial_positions must have shape (n_walkers, 3)")
        if np.asarray(self.increments).shape != (int(cfg.n_steps), n, 3):
            raise ValueError("golden increments must have shape (n_steps, n_walkers, 3)")
        if np.asarray(self.acceptance_uniforms).shape != (int(cfg.n_steps), n):
            raise ValueError("golden acceptance_uniforms must have shape (n_steps, n_walkers)")
        arrays = (self.initial_positions, self.increments, self.acceptance_uniforms)
        if not all(np.all(np.isfinite(a)) for a in arrays):
            raise ValueError("golden stream contains non-finite values")
        u = np.asarray(self.acceptance_uniforms)
        if np.any(u < 0.0) or np.any(u >= 1.0):
            raise ValueError("golden acceptance uniforms lie outside [0,1)")
        if ensemble is not None:
            x0 = np.asarray(self.initial_positions)
            if np.any(x0 < ensemble.source_lo) or np.any(x0 >= ensemble.source_hi):
                raise ValueError("golden initial positions must lie in Ω_src")


def make_walk_random_stream(cfg: SimConfig, seed: int, ensemble: Optional[Ensemble] = None) -> WalkRandomStream:
    """Deterministic injected stream from numpy PCG64 (madi/walker_gpu.py:115-127)."""
    if ensemble is None:
        ensemble = create_dummy_ensemble(cfg)
    rng = np.random.default_rng(int(seed))
    n = int(cfg.n_walkers)
    stream = WalkRandomStream(
        initial_positions=rng.uniform(ensemble.source_lo, ensemble.source_hi, size=(n, 3)).astype(np.float64),
        increments=rng.normal(0.0, cfg.sigma, size=(int(cfg.n_steps), n, 3)).astype(np.float64),
        acceptance_uniforms=rng.uniform(0.0, 1.0, size=(int(cfg.n_steps), n)).astype(np.float64),
    )
    stream.validate(cfg, ensemble)
    return stream


@dataclass
class ReducedResult:
    """Per-entry reduced walk (madi/walker_gpu.py:130-165)."""

    cos_sum: np.ndarray
    sin_sum: np.ndarray
    n_walkers: int
    n_escaped: int
    occupancy_counts: np.ndarray
    pp_values: list = field(default_factory=list)
    analytic_kio_eq5_values: list = field(default_factory=list)
    geometry_stats: list = field(default_factory=list)
    ensemble_cos_means: list = field(default_factory=list)
    ensemble_sin_means: list = field(default_factory=list)
    classifier_cache_stats: list = field(default_factory=list)

    @property
    def n_eff(self) -> int:
        return 3 * self.n_walkers

    @property
    def analytic_kio_eq5(self) -> float:
        return float(np.mean(self.analytic_kio_eq5_values)) if self.analytic_kio_eq5_values else 0.0

    @property
    def mean_pp(self) -> float:
        return float(np.mean(self.pp_values)) if self.pp_values else 0.0

    @property
    def occupancy_fraction(self) -> np.ndarray:
        if self.n_walkers <= 0:
            return np.zeros_like(self.occupancy_counts, dtype=np.float64)
        return self.occupancy_counts.astype(np.float64) / self.n_walkers


@dataclass
class ClassifierCacheStats:
    """Classifier accounting of one walk (counter names of madi/walker_gpu.py:168-203).

    This build's counters map as: ``full`` = endpoint classifications that
    evaluated facets (KD or Delaunay graph), ``tier1`` = endpoints answered by
    the certified safe ball, ``tier2`` = 0 (no candidate-list tier).
    """

    full: int = 0
    tier1: int = 0
    tier2: int = 0
    candidate_overflow: int = 0
    full_inside: int = 0
    full_outside: int = 0
    tier1_inside: int = 0
    tier2_inside: int = 0
    tier2_outside: int = 0
    device_counters: dict = field(default_factory=dict)

    @property
    def requests(self) -> int:
        return self.full + self.tier1 + self.tier2

    def as_dict(self) -> dict:
        out = {k: int(getattr(self, k)) for k in (
            "full", "tier1", "tier2", "candidate_overflow", "full_inside", "full_outside",
            "tier1_inside", "tier2_inside", "tier2_outside")}
        out["requests"] = int(self.requests)
        out.update({f"device_{k}": int(v) for k, v in self.device_counters.items()})
        return out


def _stats_from_counters(counters: dict, n_requests: Optional[int] = None) -> ClassifierCacheStats:
    full = int(counters.get("classify", 0))
    if full == 0 and n_requests is not None:
        full = int(n_requests)
    return ClassifierCacheStats(full=full, tier1=int(counters.get("safe_hit", 0)), device_counters=dict(counters))


def _membrane_transition_counts(old_cell, old_inside, new_cell, new_inside) -> np.ndarray:
    """Endpoint-only membrane count m (madi/walker_gpu.py:206-217)."""
    old_inside = np.asarray(old_inside, dtype=bool)
    new_inside = np.asarray(new_inside, dtype=bool)
    m = np.where(old_inside != new_inside, 1, 0)
    m = np.where(old_inside & new_inside & (np.asarray(old_cell) != np.asarray(new_cell)), 2, m)
    return m.astype(np.int8)


def _require_device(use_gpu: Optional[bool]) -> None:
    if use_gpu is False:
        raise RuntimeError(
            "this build has no CPU walk: the MADI walk runs only in its HIP kernels "
            "(the CPU restatement lives in the test oracle)"
        )
    _hip.require_gpu()


def _run_walk_one_ensemble(ens: Ensemble, pp: float, cfg: SimConfig, seed: int, verbose: bool,
                           random_stream: Optional[WalkRandomStream] = None, use_gpu: Optional[bool] = None,
                           snapshot_steps: Optional[np.ndarray] = None):
    """One FP64 walk writing Y (madi/walker_gpu.py:843-951).

    Returns (Y device tensor, escaped, inside_trace, None, True, stats).
    """
    _ = verbose
    pp = _checked_pp(pp)
    if random_stream is not None:
        random_stream.validate(cfg, ens)
    if snapshot_steps is not None and len(snapshot_steps) > 0:
        raise NotImplementedError("narrow-pulse endpoint snapshots are a CPU-only diagnostic of the reference")
    _require_device(use_gpu)
    geom = ens.device_geometry()
    params = engine.walk_params(cfg.n_walkers, cfg, pp, seed)
    out = engine.walk_y(geom, params, random_stream)
    if np.any(out.err):
        raise RuntimeError("GPU full-facet KD-tree classification failed")
    n = int(cfg.n_walkers)
    stats = _stats_from_counters(out.counters, n * (cfg.n_steps + 1))
    return out.Y, out.escaped, out.inside_trace, None, True, stats


def run_walk_Y(ens: Ensemble, kio: float, cfg: Optional[SimConfig] = None, seed: int = 0, verbose: bool = True, *,
               pp: Optional[float] = None, return_telemetry: bool = False, classifier: str = "exact",
               random_stream: Optional[WalkRandomStream] = None, use_gpu: Optional[bool] = None):
    """Walk and return host Y (madi/walker_gpu.py:1093-1133)."""
    if cfg is None:
        cfg = SimConfig()
    if classifier not in {"exact", "exact_cached"}:
        raise ValueError("classifier must be 'exact' or 'exact_cached'")
    if classifier != cfg.classifier_mode:
        cfg = replace(cfg, classifier_mode=classifier)
    actual_pp = _checked_pp(pp if pp is not None else kio_to_pp(kio, ens.mean_AV, cfg))
    Y, escaped, inside, _, _, stats = _run_walk_one_ensemble(
        ens, actual_pp, cfg, seed, verbose, random_stream=random_stream, use_gpu=use_gpu)
    Y_host = Y.cpu().numpy()
    n_escape = int(np.sum(escaped))
    if n_escape:
        raise RuntimeError(f"SI §S.III fatal boundary condition: {n_escape} walkers escaped Ω_sim")
    if return_telemetry:
        return Y_host, n_escape, {
            "occupancy_counts": np.sum(inside, axis=0),
            "occupancy_fraction": np.mean(inside, axis=0),
            "pp": actual_pp,
            "classifier_cache": stats.as_dict(),
        }
    return Y_host, n_escape


def _with_n_walkers(cfg: SimConfig, n_walkers: int) -> SimConfig:
    return replace(cfg, n_walkers=int(n_walkers))


def _walk_and_reduce_one_ensemble(ens: Ensemble, pp: float, cfg: SimConfig, j_delta: np.ndarray,
                                  j_Delta: np.ndarray, j_sum: np.ndarray, phase_coef: np.ndarray, walk_seed: int, *,
                                  narrow_snapshot_steps=None, narrow_snapshot_index=None,
                                  narrow_phase_coef=None) -> ReducedResult:
    """Walk + reduce one ensemble over walker chunks (madi/walker_gpu.py:989-1063)."""
    narrow = (narrow_snapshot_steps, narrow_snapshot_index, narrow_phase_coef)
    if any(x is not None for x in narrow):
        if not all(x is not None for x in narrow):
            raise ValueError("narrow-pulse reduction requires all endpoint arrays together")
        raise NotImplementedError("the narrow-pulse diagnostic is CPU-only in the reference and not built here")
    pp = _checked_pp(pp)
    _require_device(None)
    n_cols = len(phase_coef)
    cos_total = np.zeros(n_cols)
    sin_total = np.zeros(n_cols)
    occupancy = np.zeros(cfg.n_grid, dtype=np.int64)
    total = 0
    stats = ClassifierCacheStats()
    chunk = int(cfg.walker_chunk or cfg.n_walkers)
    fused = engine.spec_from_columns(j_delta, j_Delta, j_sum, phase_coef)
    precision = 32 if cfg.precision == "fp32" else 64
    geom = ens.device_geometry()
    runner = None
    for offset in range(0, int(cfg.n_walkers), chunk):
        n_this = min(chunk, int(cfg.n_walkers) - offset)
        params = engine.walk_params(n_this, cfg, pp, walk_seed + offset)
        if fused is not None:
            spec, col_pair, col_b = fused
            if runner is None:
                runner = engine.FusedRunner(chunk, spec, cfg.n_grid)
            runner.launch(geom, params, precision)
            res = runner.result(n_this)
            if res.n_error:
                raise RuntimeError("GPU full-facet classification failed")
            if res.n_escaped:
                raise RuntimeError(f"SI §S.III fatal boundary condition: {res.n_escaped}/{n_this} walkers exited Ω_sim.")
            c = res.cos.sum(axis=2)[col_pair, col_b]
            s = res.sin.sum(axis=2)[col_pair, col_b]
            occ = res.occupancy
            counters = res.counters
        elif precision == 32:
            # any other column set: the FP32 walk writes Y, then the FP64 column reduce
            out = engine.walk_y_f32(geom, params)
            if out.n_error:
                raise RuntimeError("GPU full-facet classification failed")
            if out.n_escaped:
                raise RuntimeError(f"SI §S.III fatal boundary condition: {out.n_escaped}/{n_this} walkers exited Ω_sim.")
            c, s = engine.reduce_columns(out.Y, j_delta, j_Delta, j_sum, phase_coef)
            occ = out.occupancy
            counters = out.counters
            del out
        else:
            out = engine.walk_y(geom, params)
            if np.any(out.err):
                raise RuntimeError("GPU full-facet KD-tree classification failed")
            n_esc = int(np.sum(out.escaped))
            if n_esc:
                raise RuntimeError(f"SI §S.III fatal boundary condition: {n_esc}/{n_this} walkers exited Ω_sim.")
            c, s = engine.reduce_columns(out.Y, j_delta, j_Delta, j_sum, phase_coef)
            occ = np.sum(out.inside_trace, axis=0, dtype=np.int64)
            counters = out.counters
            del out
        cos_total += c
        sin_total += s
        occupancy += occ
        total += n_this
        st = _stats_from_counters(counters, n_this * (cfg.n_steps + 1))
        stats.full += st.full
        stats.tier1 += st.tier1
        for k, v in st.device_counters.items():
            stats.device_counters[k] = stats.device_counters.get(k, 0) + v
    return ReducedResult(
        cos_sum=cos_total, sin_sum=sin_total, n_walkers=total, n_escaped=0, occupancy_counts=occupancy,
        pp_values=[float(pp)],
        analytic_kio_eq5_values=[pp_to_kio_eq5(pp, ens.mean_AV, cfg)] if not ens.is_free_water else [0.0],
        geometry_stats=[ens.geometry.to_dict()],
        ensemble_cos_means=[cos_total / float(3 * total)], ensemble_sin_means=[sin_total / float(3 * total)],
        classifier_cache_stats=[stats.as_dict()],
    )


# Y of the FP64 walks run together in one launch (the rest of the sweep
# follows in further launches)
_MULTI_WALK_Y_BYTES = 16 << 30


def _walk_and_reduce_kio_sweep(ens: Ensemble, pps: Sequence[float], cfg: SimConfig, j_delta, j_Delta, j_sum,
                               phase_coef, walk_seed: int) -> list:
    """[_walk_and_reduce_one_ensemble(ens, pp, ...) for pp in pps], bit for bit.

    The reference walks an ensemble once per k_io with one walk seed
    (madi/walker_gpu.py:1193-1226).  The FP64 walks whose columns are reduced
    from Y run up to 64 per launch (madi_walk_f64_multi): a walker's walks in
    adjacent lanes, sharing each step's draws and their FP64 arithmetic,
    every value computed as in the one-walk kernel, so the results are the
    same.  One walk of 50,000 walkers is 196 blocks, under one per CU; the
    sweep together fills the GPU."""
    pps = [_checked_pp(pp) for pp in pps]
    fused = engine.spec_from_columns(j_delta, j_Delta, j_sum, phase_coef)
    if cfg.precision == "fp32" or fused is not None or len(pps) < 2:
        return [_walk_and_reduce_one_ensemble(ens, pp, cfg, j_delta, j_Delta, j_sum, phase_coef, walk_seed)
                for pp in pps]
    _require_device(None)
    K, n_cols, G = len(pps), len(phase_coef), int(cfg.n_grid)
    cos_t = [np.zeros(n_cols) for _ in range(K)]
    sin_t = [np.zeros(n_cols) for _ in range(K)]
    occ_t = [np.zeros(G, dtype=np.int64) for _ in range(K)]
    stats = [ClassifierCacheStats() for _ in range(K)]
    total = 0
    chunk = int(cfg.walker_chunk or cfg.n_walkers)
    geom = ens.device_geometry()
    for offset in range(0, int(cfg.n_walkers), chunk):
        n_this = min(chunk, int(cfg.n_walkers) - offset)
        params = engine.walk_params(n_this, cfg, pps[0], walk_seed + offset)
        per_launch = int(max(1, min(_hip.MADI_MAX_MULTI_WALKS, _MULTI_WALK_Y_BYTES // max(1, n_this * G * 24))))
        for k0 in range(0, K, per_launch):
            outs = engine.walk_y_multi(geom, params, pps[k0:k0 + per_launch], traces=False)
            for k, out in zip(range(k0, k0 + per_launch), outs):
                if np.any(out.err):
                    raise RuntimeError("GPU full-facet KD-tree classification failed")
                n_esc = int(np.sum(out.escaped))
                if n_esc:
                    raise RuntimeError(f"SI §S.III fatal boundary condition: {n_esc}/{n_this} walkers exited Ω_sim.")
                c, sn = engine.reduce_columns(out.Y, j_delta, j_Delta, j_sum, phase_coef)
                cos_t[k] += c
                sin_t[k] += sn
                occ_t[k] += out.occupancy
                st = _stats_from_counters(out.counters, n_this * (cfg.n_steps + 1))
                stats[k].full += st.full
                stats[k].tier1 += st.tier1
                for key, v in st.device_counters.items():
                    stats[k].device_counters[key] = stats[k].device_counters.get(key, 0) + v
            del outs
        total += n_this
    return [ReducedResult(
        cos_sum=cos_t[k], sin_sum=sin_t[k], n_walkers=total, n_escaped=0, occupancy_counts=occ_t[k],
        pp_values=[float(pps[k])],
        analytic_kio_eq5_values=[pp_to_kio_eq5(pps[k], ens.mean_AV, cfg)] if not ens.is_free_water else [0.0],
        geometry_stats=[ens.geometry.to_dict()],
        ensemble_cos_means=[cos_t[k] / float(3 * total)], ensemble_sin_means=[sin_t[k] / float(3 * total)],
        classifier_cache_stats=[stats[k].as_dict()],
    ) for k in range(K)]


def _merge_results(results: Sequence[ReducedResult]) -> ReducedResult:
    """Merge per-ensemble results in the given (ensemble) order (madi/walker_gpu.py:1066-1090)."""
    if not results:
        raise ValueError("cannot merge no results")
    merged = ReducedResult(
        cos_sum=np.zeros_like(results[0].cos_sum), sin_sum=np.zeros_like(results[0].sin_sum), n_walkers=0,
        n_escaped=0, occupancy_counts=np.zeros_like(results[0].occupancy_counts),
    )
    for r in results:
        merged.cos_sum += r.cos_sum
        merged.sin_sum += r.sin_sum
        merged.n_walkers += r.n_walkers
        merged.n_escaped += r.n_escaped
        merged.occupancy_counts += r.occupancy_counts
        merged.pp_values.extend(r.pp_values)
        merged.analytic_kio_eq5_values.extend(r.analytic_kio_eq5_values)
        merged.geometry_stats.extend(r.geometry_stats)
        merged.ensemble_cos_means.extend(r.ensemble_cos_means)
        merged.ensemble_sin_means.extend(r.ensemble_sin_means)
        merged.classifier_cache_stats.extend(r.classifier_cache_stats)
    return merged


def _build_ensembles_for_entry(rho: float, V: float, cfg: SimConfig, seed: int, verbose: bool) -> list:
    """Ensembles of one (rho, V) entry (madi/walker_gpu.py:1136-1142).

    With ``cfg.periodic`` the ensembles are periodic boxes of side cfg.L
    holding round(rho * L^3) cells.
    """
    return [_build_ensemble_for_entry(rho, V, cfg, seed, i, verbose) for i in range(int(cfg.n_ensembles))]


def _build_ensemble_for_entry(rho: float, V: float, cfg: SimConfig, seed: int, index: int,
                              verbose: bool = False) -> Ensemble:
    """Ensemble `index` of one (rho, V) entry: each ensemble depends only on
    its own seed, so they may be built apart (in parallel)."""
    if rho <= 0.0 or V <= 0.0:
        return create_dummy_ensemble(cfg)
    if cfg.periodic:
        if cfg.L is None:
            raise ValueError("a periodic geometry needs cfg.L")
        n_cells = max(2, int(round(rho * 1e-9 * cfg.L ** 3)))
        rho_box = n_cells / cfg.L ** 3 * 1e9
        vi = rho * V * 1e-6
        return create_periodic_ensemble(n_cells, rho_box, vi, cfg, seed=_ensemble_seed(seed, index))
    return create_ensemble(rho, V, cfg, seed=_ensemble_seed(seed, index), verbose=verbose)


def _ensembles_in_turn(rho: float, V: float, cfg: SimConfig, seed: int, verbose: bool):
    """The entry's ensembles one at a time, in order (the reference builds
    them all before the first walk, madi/walker_gpu.py:1136-1142, 1193).
    The CPU part of ensemble i + 1 (population, KD tree, Voronoi graph) is
    built on a worker thread while ensemble i walks, and an ensemble's
    device geometry is freed once the caller drops it.  Each ensemble
    depends only on its own seed, so the ensembles, and every result, are
    those of building them all first."""
    n = int(cfg.n_ensembles)
    if n <= 1 or rho <= 0.0 or V <= 0.0 or cfg.periodic:
        for i in range(n):
            yield _build_ensemble_for_entry(rho, V, cfg, seed, i, verbose)
        return

    def host(i):
        draft = build_ensemble_host(rho, V, cfg, seed=_ensemble_seed(seed, i))
        return draft, geometry_arrays(draft.ens)

    with ThreadPoolExecutor(max_workers=1) as pool:
        pending = pool.submit(host, 0)
        for i in range(n):
            draft, arrays = pending.result()
            if i + 1 < n:
                pending = pool.submit(host, i + 1)
            # the v_i measurement runs on the GPU, here in the calling thread
            yield finish_ensemble(draft, cfg, verbose=verbose, arrays=arrays)


def run_simulation_reduced(rho: float, V: float, kio: float, j_delta, j_Delta, j_sum, phase_coef,
                           cfg: Optional[SimConfig] = None, seed: int = 0, verbose: bool = True, *,
                           narrow_snapshot_steps=None, narrow_snapshot_index=None,
                           narrow_phase_coef=None) -> ReducedResult:
    """One Eq.-5-labelled entry (madi/walker_gpu.py:1145-1167)."""
    out = run_simulation_multi_kio_reduced(
        rho, V, [kio], j_delta, j_Delta, j_sum, phase_coef, cfg=cfg, seed=seed, verbose=verbose,
        narrow_snapshot_steps=narrow_snapshot_steps, narrow_snapshot_index=narrow_snapshot_index,
        narrow_phase_coef=narrow_phase_coef)
    return out[float(kio)]


def run_simulation_multi_kio_reduced(rho: float, V: float, kios: Sequence[float], j_delta, j_Delta, j_sum,
                                     phase_coef, cfg: Optional[SimConfig] = None, seed: int = 0,
                                     verbose: bool = True, *, narrow_snapshot_steps=None,
                                     narrow_snapshot_index=None, narrow_phase_coef=None) -> dict:
    """Geometry reused across a k_io sweep (madi/walker_gpu.py:1170-1233)."""
    if cfg is None:
        cfg = SimConfig()
    cfg.assert_grid_alignment()
    narrow = (narrow_snapshot_steps, narrow_snapshot_index, narrow_phase_coef)
    if any(x is not None for x in narrow) and not all(x is not None for x in narrow):
        raise ValueError("narrow-pulse reduction requires all endpoint arrays together")
    is_narrow = all(x is not None for x in narrow)
    if cfg.phase_model == "narrow_pulse":
        if not is_narrow:
            raise ValueError("narrow-pulse cfg requires endpoint displacement arrays")
        raise NotImplementedError("the narrow-pulse diagnostic is CPU-only in the reference and not built here")
    if is_narrow:
        raise ValueError("endpoint displacement arrays require phase_model='narrow_pulse'")
    kios = [float(k) for k in kios]
    if any(k < 0.0 for k in kios):
        raise ValueError("k_io grid values must be non-negative")
    total_walks = 3 * int(cfg.n_ensembles) * int(cfg.n_walkers)
    if total_walks < PAPER_WALKS_PER_ENTRY // 10:
        sys.stderr.write(
            f"    WARNING: only {total_walks:,} axis-walks per entry; production reference is "
            f"~{PAPER_WALKS_PER_ENTRY:,}.\n")
    outputs = {k: [] for k in kios}
    n_ens = int(cfg.n_ensembles)
    for e, ens in enumerate(_ensembles_in_turn(rho, V, cfg, seed, verbose)):
        ws = _walk_seed(seed, e)
        pps = [0.0 if (ens.is_free_water or kio == 0.0) else _checked_pp(kio_to_pp(kio, ens.mean_AV, cfg))
               for kio in kios]
        for kio, pp, res in zip(kios, pps, _walk_and_reduce_kio_sweep(ens, pps, cfg, j_delta, j_Delta, j_sum,
                                                                       phase_coef, ws)):
            outputs[kio].append(res)
            if verbose:
                print(f"    ensemble {e + 1}/{n_ens}, k_io={kio:g}: p_p={pp:.7g}, "
                      f"Eq.5={res.analytic_kio_eq5:.5g} s^-1")
    return {k: _merge_results(v) for k, v in outputs.items()}


def _walk_cpu(*args, **kwargs):
    raise RuntimeError("this build has no CPU walk; the CPU restatement of the reference walk is the test oracle")


def _reduce_cpu(*args, **kwargs):
    raise RuntimeError("this build has no CPU reduction; use engine.reduce_columns on a device Y")
