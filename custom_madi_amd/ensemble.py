"""Contracted Poisson-Voronoi ensembles and their device geometry.

Host-side construction follows the reference's SI geometry
(madi/ensemble.py:405-792): Eq. S8 domain side, Eq. S7a alpha*, a certified
Omega_pop population (Eqs. S10-S11), annulus cap min(alpha*, kappa d_nn / 2),
a balanced KD tree, and v_i acceptance by direct volume sampling.  Every
classification -- including the v_i measurement -- runs in the HIP kernels
(``madi_classify``); there is no CPU classifier here.

The device geometry (``DeviceGeometry``) adds what the MI355X kernels need:
the KD tree for the FP64 exact classifier, the Delaunay neighbour graph of
the seeds (edge vectors in FP64, unit vectors and half lengths in FP32) for
the FP32 production walk, and a hint grid that gives every walker a nearby
starting seed.  A periodic ensemble (a torus of side L) is represented by
periodic images: the KD tree indexes the seeds plus their images within a
certified margin, and every Delaunay edge carries the image shift, so the
walk kernels never wrap coordinates.

Every geometric query of the construction is native host code of the HIP
library: the KD build (``madi_kd_build``), nearest-point queries on a bin
grid (``madi_nearest_points``: nearest-neighbour distances, the Eq. S10
certification grid, the hint grid) and the Voronoi neighbour graph
(``madi_voronoi_graph``: each cell cut out by its neighbours' bisector
planes).
"""

from __future__ import annotations

import ctypes as C
import json
import math
from dataclasses import asdict, dataclass
from pathlib import Path
from typing import Any, Optional, Tuple

import numpy as np

from . import _hip
from .config import SimConfig

# SI Eq. S7a coefficients, x = rho^(1/3) alpha* (madi/ensemble.py:35-40).
_S7A = (-0.570444, 10.6047, -5.81457, 1.0)

DATA_DIR = Path(__file__).resolve().parent / "data"


def _rho_um3(rho_per_uL: float) -> float:
    return float(rho_per_uL) * 1.0e-9


def _V_um3(V_pL: float) -> float:
    return float(V_pL) * 1.0e3


def _rho_per_uL(rho_um3: float) -> float:
    return float(rho_um3) * 1.0e9


def _distance_to_cube(points: np.ndarray, lo: float, hi: float) -> np.ndarray:
    """Euclidean distance of points to the closed cube [lo, hi]^3."""
    p = np.asarray(points, dtype=np.float64)
    out = np.maximum(lo - p, 0.0) + np.maximum(p - hi, 0.0)
    return np.sqrt(np.sum(out * out, axis=-1))


@dataclass(frozen=True)
class GeometryReference:
    vi: np.ndarray
    alpha_x: np.ndarray
    mean_A_over_V_norm: np.ndarray
    metadata: dict
    path: str


@dataclass(frozen=True)
class PopulationCertificate:
    """Omega_pop certificate for SI Eqs. S10-S11 (madi/ensemble.py:75-97)."""

    pop_margin_um: float
    s10_distance_upper_um: float
    s10_pass: bool
    n_relevant_seeds: int
    s11_min_clearance_margin_um: float
    s11_pass: bool
    certification_spacing_um: float
    expansion_index: int

    def to_dict(self) -> dict:
        return asdict(self)


@dataclass(frozen=True)
class GeometryStats:
    """Realised-geometry provenance (madi/ensemble.py:100-131)."""

    target_vi: float
    realised_vi: float
    realised_vi_se: float
    realised_rho_per_uL: float
    realised_mean_volume_pL: float
    alpha_star_um: float
    governing_mean_A_over_V_um_inv: float
    annulus_mean_um: float
    annulus_std_um: float
    annulus_min_um: float
    annulus_max_um: float
    annulus_q05_um: float
    annulus_q50_um: float
    annulus_q95_um: float
    n_seeds_pop: int
    n_seeds_sim: int
    n_validation_points: int
    sim_side_um: float
    source_side_um: float
    population: PopulationCertificate
    reference_path: str
    reference_metadata: dict

    def to_dict(self) -> dict:
        out = asdict(self)
        out["population"] = self.population.to_dict()
        return out


# ------------------------------------------------------------------ SI maths

def si_domain_side_um(rho_per_uL: float, cfg: SimConfig) -> float:
    """Omega_sim side W of SI Eq. S8 (madi/ensemble.py:405-423)."""
    if cfg.L is not None:
        if cfg.L <= 0.0:
            raise ValueError("diagnostic L override must be positive")
        return float(cfg.L)
    L_rw = math.sqrt(2.0 * cfg.D0 * cfg.tRW_max)
    rho = _rho_um3(rho_per_uL)
    if rho <= 0.0:
        return float(20.0 * L_rw)
    return float(max(10.0 * L_rw, min(20.0 * L_rw, (8.0e5 / rho) ** (1.0 / 3.0))))


def _s7a_poly(x):
    a, b, c, d = _S7A
    return ((a * x + b) * x + c) * x + d


def alpha_star_from_vi(vi: float, rho_per_uL: float) -> float:
    """alpha* [µm] from the physical root of SI Eq. S7a (madi/ensemble.py:426-457).

    v_i(x) is decreasing on [0, x_turn], where x_turn is the first positive
    stationary point of the cubic; the physical root is the unique root of
    v_i(x) - vi on that branch.
    """
    vi = float(vi)
    rho = _rho_um3(rho_per_uL)
    if not (0.0 < vi < 1.0):
        raise ValueError(f"target v_i must lie in (0, 1), got {vi}")
    if rho <= 0.0:
        raise ValueError("rho must be positive for a cellular geometry")
    a, b, c, d = _S7A
    stationary = np.roots([3.0 * a, 2.0 * b, c])
    x_turn = min(float(r.real) for r in stationary if abs(r.imag) < 1e-10 and r.real > 0.0)
    candidates = [float(r.real) for r in np.roots([a, b, c, d - vi])
                  if abs(r.imag) < 1e-10 and -1e-12 <= r.real <= x_turn + 1e-12]
    if len(candidates) != 1:
        raise RuntimeError(f"SI Eq. S7a has no unique physical root for v_i={vi}")
    return candidates[0] / rho ** (1.0 / 3.0)


def _reference_default_path() -> Path:
    return DATA_DIR / "geometry_reference_si_kappa_0p9.npz"


def _load_reference(cfg: SimConfig) -> GeometryReference:
    """Load and validate the SI single-cell <A/V> table (madi/ensemble.py:464-509)."""
    path = Path(cfg.geometry_reference_path) if cfg.geometry_reference_path else _reference_default_path()
    if not path.is_file():
        raise RuntimeError(
            f"No SI-certified geometry reference table was found at {path}; "
            "refusing to substitute a finite Omega_sim A/V."
        )
    with np.load(path, allow_pickle=False) as data:
        missing = {"vi", "alpha_x", "mean_A_over_V_norm", "metadata_json"} - set(data.files)
        if missing:
            raise RuntimeError(f"geometry reference {path} is missing {sorted(missing)}")
        try:
            meta = json.loads(str(data["metadata_json"]))
        except (TypeError, ValueError) as exc:
            raise RuntimeError(f"geometry reference {path} has invalid metadata_json") from exc
        vi = np.asarray(data["vi"], dtype=np.float64)
        ax = np.asarray(data["alpha_x"], dtype=np.float64)
        av = np.asarray(data["mean_A_over_V_norm"], dtype=np.float64)
    if vi.ndim != 1 or len(vi) < 2 or ax.shape != vi.shape or av.shape != vi.shape:
        raise RuntimeError("geometry reference axes must be matching 1-D arrays")
    if not (np.all(np.diff(vi) > 0.0) and np.all(np.diff(ax) < 0.0)):
        raise RuntimeError("geometry reference v_i must increase while alpha_x decreases")
    if not np.all(np.isfinite(av)) or np.any(av <= 0.0):
        raise RuntimeError("geometry reference contains invalid <A/V> values")
    if abs(float(meta.get("kappa", np.nan)) - 0.90) > 1e-12:
        raise RuntimeError("geometry reference must use SI κ=0.9")
    if meta.get("mean_estimator") != "untrimmed_arithmetic_mean_A_over_V":
        raise RuntimeError("geometry reference must store the untrimmed arithmetic <A/V> mean")
    if meta.get("contraction_rule") != "all_shifted_voronoi_facets_S2":
        raise RuntimeError("geometry reference must use the full SI Eq. S2 shifted-facet contraction")
    if not cfg.allow_uncertified_geometry_reference:
        if int(meta.get("n_single_cells", 0)) < cfg.geometry_reference_required_cells:
            raise RuntimeError(
                "geometry reference is not SI-certified: it must contain at least "
                f"{cfg.geometry_reference_required_cells:,} independent single cells"
            )
        if int(meta.get("n_alpha", 0)) != cfg.geometry_reference_required_alpha_values:
            raise RuntimeError(
                "geometry reference is not SI-certified: it must contain exactly "
                f"{cfg.geometry_reference_required_alpha_values} alpha values"
            )
    return GeometryReference(vi, ax, av, meta, str(path.resolve()))


def governing_mean_A_over_V(vi: float, rho_per_uL: float, cfg: SimConfig) -> Tuple[float, GeometryReference]:
    """Table <A/V>_norm at v_i, scaled by rho^(1/3) [µm^-1] (madi/ensemble.py:512-522)."""
    ref = _load_reference(cfg)
    if vi < ref.vi[0] or vi > ref.vi[-1]:
        raise RuntimeError(
            f"requested v_i={vi:.6f} is outside certified reference range "
            f"[{ref.vi[0]:.6f}, {ref.vi[-1]:.6f}]; no clamping is allowed"
        )
    norm = float(np.interp(float(vi), ref.vi, ref.mean_A_over_V_norm))
    return norm * _rho_um3(rho_per_uL) ** (1.0 / 3.0), ref


# ------------------------------------------------------------------ KD tree

def build_kdtree(points: np.ndarray) -> Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """Balanced KD tree in pre-order (madi/ensemble.py:525-563), built by the
    native madi_kd_build (host code of the HIP library, no device call).

    Each node splits its subset on the axis of largest extent at the element
    of rank len//2 along that axis; the node ids are assigned in pre-order
    (root 0, left subtree before right).  The arrays are a function of the
    point set only, so they equal the reference's for the same seeds (whose
    split coordinates are distinct; equal coordinates are ordered by index).
    """
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    n = len(pts)
    seed = np.empty(n, dtype=np.int32)
    axis = np.empty(n, dtype=np.int8)
    left = np.empty(n, dtype=np.int32)
    right = np.empty(n, dtype=np.int32)
    parent = np.empty(n, dtype=np.int32)
    lib = _hip.load_library()
    rc = lib.madi_kd_build(pts.ctypes.data, n, seed.ctypes.data, axis.ctypes.data, left.ctypes.data,
                           right.ctypes.data, parent.ctypes.data)
    _hip.check(rc, "madi_kd_build")
    if n and parent[0] != -1:
        raise RuntimeError("KD tree root must occupy node zero")
    return seed, axis, left, right, parent


# ------------------------------------------------------- native point queries

def nearest_points(points: np.ndarray, queries: Optional[np.ndarray] = None, period: float = 0.0,
                   n_threads: int = 0) -> Tuple[np.ndarray, np.ndarray]:
    """(index, distance) of the nearest point to each query (madi_nearest_points).

    queries=None queries the points themselves and returns each point's
    nearest OTHER point (the k=2 column of a KD-tree query).  period > 0:
    periodic cube [0, period)^3, minimum-image distances."""
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    self_query = queries is None
    q = pts if self_query else np.ascontiguousarray(queries, dtype=np.float64).reshape(-1, 3)
    idx = np.empty(len(q), dtype=np.int32)
    dist = np.empty(len(q), dtype=np.float64)
    lib = _hip.load_library()
    rc = lib.madi_nearest_points(pts.ctypes.data, len(pts), q.ctypes.data, len(q), float(period), int(self_query),
                                 int(n_threads), idx.ctypes.data, dist.ctypes.data)
    _hip.check(rc, "madi_nearest_points")
    return idx, dist


def voronoi_graph(points: np.ndarray, owned: int, pad: float,
                  n_threads: int = 0) -> Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """CSR Voronoi neighbours of the first ``owned`` points (madi_voronoi_graph):
    (offsets (owned + 1,) int64, ids int32 into ``points``, each cell's largest
    vertex distance, 1 where a cell reaches the points' bounding box widened
    by ``pad``).

    Coincident seeds have no facets between them, so no label is defined
    near them: that is the reference's classifier failure (the exact
    classifier divides by the seeds' separation, madi/walker_gpu.py:272-335,
    938-940), raised here as RuntimeError before any walk starts."""
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    owned = int(owned)
    offsets = np.zeros(owned + 1, dtype=np.int64)
    radius = np.zeros(owned, dtype=np.float64)
    boxed = np.zeros(owned, dtype=np.int8)
    n_edges = C.c_int64(0)
    lib = _hip.load_library()
    capacity = 24 * owned + 64
    for _ in range(2):
        ids = np.empty(capacity, dtype=np.int32)
        rc = lib.madi_voronoi_graph(pts.ctypes.data, len(pts), owned, float(pad), int(n_threads),
                                    offsets.ctypes.data, ids.ctypes.data, capacity, C.byref(n_edges),
                                    radius.ctypes.data, boxed.ctypes.data)
        if rc == _hip.MADI_ERR_ARG and n_edges.value > capacity:
            capacity = int(n_edges.value)
            continue
        if rc == _hip.MADI_ERR_CLASSIFIER:
            raise RuntimeError(_hip.last_error())
        _hip.check(rc, "madi_voronoi_graph")
        return offsets, ids[: n_edges.value].copy(), radius, boxed
    raise RuntimeError("madi_voronoi_graph: edge count changed between calls")


# ---------------------------------------------------------------- population

def _nearest_distance_upper_bound(seeds: np.ndarray, side_um: float, spacing_um: float) -> Tuple[float, float]:
    """Grid + Lipschitz upper bound of the nearest-seed distance over [0, side]^3
    (madi/ensemble.py:566-582): the largest nearest-seed distance over the
    centres of an m^3 grid of spacing h <= spacing_um, plus half the grid
    cell's diagonal."""
    m = max(1, int(np.ceil(side_um / spacing_um)))
    h = side_um / m
    axis = (np.arange(m, dtype=np.float64) + 0.5) * h
    X, Y, Z = np.meshgrid(axis, axis, axis, indexing="ij")
    _, d = nearest_points(seeds, np.column_stack((X.ravel(), Y.ravel(), Z.ravel())))
    return float(np.max(d)) + 0.5 * math.sqrt(3.0) * h, h


def _make_population(rho_um3: float, side_um: float, cfg: SimConfig, rng: np.random.Generator):
    """Sample and certify Omega_pop (madi/ensemble.py:585-638).

    Draw order (poisson count, then uniform seeds) matches the reference, so
    a given generator state yields the same seeds.
    """
    spacing = rho_um3 ** (-1.0 / 3.0)
    margin = max(cfg.population_initial_margin_cell_spacings * spacing, 2.0 * cfg.population_certification_spacing_um)
    for expansion in range(int(cfg.population_max_expansions)):
        lo, hi = -margin, side_um + margin
        count = int(rng.poisson(rho_um3 * (hi - lo) ** 3))
        if count < 2:
            margin *= 1.75
            continue
        seeds = rng.uniform(lo, hi, size=(count, 3)).astype(np.float64)
        _, nn = nearest_points(seeds)
        s10_upper, h = _nearest_distance_upper_bound(seeds, side_um, cfg.population_certification_spacing_um)
        s10_pass = bool(s10_upper < margin)
        relevant = _distance_to_cube(seeds, 0.0, side_um) <= s10_upper
        clearance = np.min(np.minimum(seeds - lo, hi - seeds), axis=1)
        s11 = clearance[relevant] - nn[relevant]
        min_s11 = float(np.min(s11)) if s11.size else float("inf")
        s11_pass = bool(np.all(s11 > 0.0))
        cert = PopulationCertificate(
            pop_margin_um=float(margin), s10_distance_upper_um=float(s10_upper), s10_pass=s10_pass,
            n_relevant_seeds=int(np.count_nonzero(relevant)), s11_min_clearance_margin_um=min_s11,
            s11_pass=s11_pass, certification_spacing_um=float(h), expansion_index=expansion,
        )
        if s10_pass and s11_pass:
            return seeds, nn, cert
        margin *= 1.75
    raise RuntimeError(
        f"Could not certify Ω_pop against SI Eqs. S10--S12 after {cfg.population_max_expansions} expansions."
    )


# -------------------------------------------------------------- the ensemble

@dataclass
class Ensemble:
    """One contracted Poisson-Voronoi realisation (fields of madi/ensemble.py:183-206).

    ``periodic`` marks a torus of side L (this build's synthetic benchmark
    boxes): ``seeds`` are then the cells of one period.
    """

    seeds: np.ndarray
    annulus: np.ndarray
    rho: float
    V: float
    vi: float
    alpha_star: float
    L: float
    source_lo: float
    source_hi: float
    mean_AV: float
    geometry: GeometryStats
    rho_requested: float
    V_requested: float
    kd_node_seed: np.ndarray
    kd_node_axis: np.ndarray
    kd_node_left: np.ndarray
    kd_node_right: np.ndarray
    kd_node_parent: np.ndarray
    tree: Optional[Any] = None
    is_free_water: bool = False
    periodic: bool = False
    _device: Optional["DeviceGeometry"] = None

    def _tree(self):
        """A scipy cKDTree of the seeds (the reference's private accessor, for
        callers that query it; nothing in this package does)."""
        if self.tree is None:
            from scipy.spatial import cKDTree
            self.tree = cKDTree(self.seeds, boxsize=self.L if self.periodic else None)
        return self.tree

    def device_geometry(self) -> "DeviceGeometry":
        """Device-resident geometry (built once, reused by every walk)."""
        if self._device is None:
            self._device = DeviceGeometry(self)
        return self._device

    def kernel_arrays(self) -> "GeometryArrays":
        """The arrays the kernels take: those of the device geometry when it
        has been built (the v_i measurement builds it), else derived now."""
        return self._device.arrays if self._device is not None else geometry_arrays(self)

    def classify(self, positions: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
        """Exact SI Eq. S2 labels (cell ids, inside flags) from the HIP classifier."""
        pts = np.asarray(positions, dtype=np.float64)
        if pts.ndim != 2 or pts.shape[1] != 3:
            raise ValueError("positions must have shape (N, 3)")
        if self.is_free_water:
            return np.zeros(len(pts), dtype=np.int32), np.zeros(len(pts), dtype=bool)
        if len(self.seeds) < 2:
            raise RuntimeError("full-facet classifier requires at least two population seeds")
        cell, inside, ok = self.device_geometry().classify(pts)
        if not np.all(ok):
            raise RuntimeError("Poisson seed realization contains coincident seeds")
        return cell, inside

    # Name of madi/ensemble.py:213 kept for callers; runs the HIP classifier.
    classify_cpu = classify
    classify_exact_cpu = classify


def _geometry_stats(target_vi, realised_vi, vi_se, rho_pul, V_pL, alpha_star, mean_av, ann, n_pop, n_sim,
                    n_points, side, source_side, population, reference) -> GeometryStats:
    return GeometryStats(
        target_vi=float(target_vi), realised_vi=float(realised_vi), realised_vi_se=float(vi_se),
        realised_rho_per_uL=float(rho_pul), realised_mean_volume_pL=float(V_pL), alpha_star_um=float(alpha_star),
        governing_mean_A_over_V_um_inv=float(mean_av),
        annulus_mean_um=float(np.mean(ann)),
        annulus_std_um=float(np.std(ann, ddof=1)) if len(ann) > 1 else 0.0,
        annulus_min_um=float(np.min(ann)), annulus_max_um=float(np.max(ann)),
        annulus_q05_um=float(np.quantile(ann, 0.05)), annulus_q50_um=float(np.quantile(ann, 0.50)),
        annulus_q95_um=float(np.quantile(ann, 0.95)),
        n_seeds_pop=int(n_pop), n_seeds_sim=int(n_sim), n_validation_points=int(n_points),
        sim_side_um=float(side), source_side_um=float(source_side), population=population,
        reference_path=reference.path if reference else "", reference_metadata=reference.metadata if reference else {},
    )


def _measure_vi(ens: Ensemble, n: int, rng: np.random.Generator) -> Tuple[float, float]:
    """Direct uniform-volume v_i of Omega_sim (madi/ensemble.py:641-646), classified on the GPU."""
    points = rng.uniform(0.0, ens.L, size=(int(n), 3))
    _, inside = ens.classify(points)
    vi = float(np.mean(inside))
    se = float(np.sqrt(max(vi * (1.0 - vi), 0.0) / max(len(points), 1)))
    return vi, se


@dataclass
class EnsembleDraft:
    """An ensemble before its v_i measurement: everything create_ensemble
    computes on the CPU, and the generator in the state the measurement
    draws its points from."""

    ens: "Ensemble"
    rng: np.random.Generator
    target_vi: float
    rho_real: float
    alpha_star: float
    mean_av: float
    annulus_sim: np.ndarray
    n_seeds: int
    n_sim: int
    side: float
    source_side: float
    population: object
    reference: object


def build_ensemble_host(rho: float, V: float, cfg: SimConfig, seed: int | None = None) -> EnsembleDraft:
    """The CPU part of create_ensemble (population, KD tree, annulus); no GPU
    call, so it may run in a worker process."""
    cfg.assert_grid_alignment()
    rho_um3 = _rho_um3(rho)
    target_vi = rho_um3 * _V_um3(V)
    if not (0.0 < target_vi < 1.0):
        raise ValueError(f"rho={rho:g}, V={V:g} give invalid v_i={target_vi:g}")
    alpha_star = alpha_star_from_vi(target_vi, rho)
    mean_av, reference = governing_mean_A_over_V(target_vi, rho, cfg)
    side = si_domain_side_um(rho, cfg)
    rng = np.random.default_rng(seed)
    seeds, nn, population = _make_population(rho_um3, side, cfg, rng)
    annulus = np.minimum(alpha_star, cfg.kappa * nn / 2.0).astype(np.float64)
    source_side = cfg.source_fraction * side
    source_lo = 0.5 * (side - source_side)
    source_hi = source_lo + source_side
    nodes = build_kdtree(seeds)
    in_sim = np.all((seeds >= 0.0) & (seeds < side), axis=1)
    n_sim = int(np.count_nonzero(in_sim))
    if n_sim < 2:
        raise RuntimeError("Ω_sim contains fewer than two seed points; increase W or density")
    rho_real = _rho_per_uL(n_sim / side ** 3)
    stats = _geometry_stats(target_vi, float("nan"), float("nan"), rho_real, float("nan"), alpha_star, mean_av,
                            annulus[in_sim], len(seeds), n_sim, cfg.geometry_validation_points, side, source_side,
                            population, reference)
    ens = Ensemble(
        seeds=seeds, annulus=annulus, rho=rho_real, V=float("nan"), vi=float("nan"), alpha_star=alpha_star,
        L=side, source_lo=source_lo, source_hi=source_hi, mean_AV=mean_av, geometry=stats,
        rho_requested=float(rho), V_requested=float(V), kd_node_seed=nodes[0], kd_node_axis=nodes[1],
        kd_node_left=nodes[2], kd_node_right=nodes[3], kd_node_parent=nodes[4],
    )
    return EnsembleDraft(ens, rng, target_vi, rho_real, alpha_star, mean_av, annulus[in_sim], len(seeds), n_sim,
                         side, source_side, population, reference)


def finish_ensemble(draft: EnsembleDraft, cfg: SimConfig, verbose: bool = False, verify_vi: bool = True,
                    arrays: Optional["GeometryArrays"] = None) -> "Ensemble":
    """The GPU part of create_ensemble: v_i measured with the HIP classifier,
    the SI acceptance test, the final geometry statistics.  arrays: the
    ensemble's geometry_arrays computed elsewhere (uploaded instead of
    derived again)."""
    ens = draft.ens
    if arrays is not None:
        ens._device = DeviceGeometry(arrays)
    realised_vi, vi_se = _measure_vi(ens, cfg.geometry_validation_points, draft.rng)
    allowed = max(float(cfg.geometry_vi_tolerance), 4.0 * vi_se)
    if verify_vi and abs(realised_vi - draft.target_vi) > allowed:
        raise RuntimeError(
            "Finite Ω_sim packing failed SI geometry acceptance: "
            f"target={draft.target_vi:.6f}, measured={realised_vi:.6f}, allowed={allowed:.6f}."
        )
    side, n_sim = draft.side, draft.n_sim
    V_real = realised_vi * side ** 3 / n_sim / 1.0e3
    ens.geometry = _geometry_stats(draft.target_vi, realised_vi, vi_se, draft.rho_real, V_real, draft.alpha_star,
                                   draft.mean_av, draft.annulus_sim, draft.n_seeds, n_sim,
                                   cfg.geometry_validation_points, side, draft.source_side, draft.population,
                                   draft.reference)
    ens.V = V_real
    ens.vi = realised_vi
    if verbose:
        print(f"    [SI geometry] W={side:.3f} um, seeds(sim/pop)={n_sim}/{draft.n_seeds}, "
              f"target v_i={draft.target_vi:.5f}, realised={realised_vi:.5f}±{vi_se:.5f}, "
              f"alpha*={draft.alpha_star:.5f} um, <A/V>={draft.mean_av:.5f} um^-1")
    return ens


def create_ensemble(rho: float, V: float, cfg: SimConfig | None = None, seed: int | None = None,
                    verbose: bool = False, verify_vi: bool = True) -> Ensemble:
    """SI-conformant ensemble for requested (rho, V) (madi/ensemble.py:649-757)."""
    if cfg is None:
        cfg = SimConfig()
    return finish_ensemble(build_ensemble_host(rho, V, cfg, seed), cfg, verbose=verbose, verify_vi=verify_vi)


def create_dummy_ensemble(cfg: SimConfig) -> Ensemble:
    """Acellular free-water control in the Eq. S8 box (madi/ensemble.py:760-792)."""
    side = si_domain_side_um(0.0, cfg)
    source_side = cfg.source_fraction * side
    source_lo = 0.5 * (side - source_side)
    cert = PopulationCertificate(0.0, 0.0, True, 0, float("inf"), True, 0.0, 0)
    zeros = np.zeros(1)
    stats = _geometry_stats(0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, zeros, 0, 0, 0, side, source_side, cert, None)
    return Ensemble(
        seeds=np.zeros((2, 3)), annulus=np.zeros(2), rho=0.0, V=0.0, vi=0.0, alpha_star=0.0, L=side,
        source_lo=source_lo, source_hi=source_lo + source_side, mean_AV=0.0, geometry=stats,
        rho_requested=0.0, V_requested=0.0,
        kd_node_seed=np.array([0, 1], dtype=np.int32), kd_node_axis=np.array([0, 0], dtype=np.int8),
        kd_node_left=np.array([-1, -1], dtype=np.int32), kd_node_right=np.array([1, -1], dtype=np.int32),
        kd_node_parent=np.array([-1, 0], dtype=np.int32), is_free_water=True,
    )


def estimate_vi(ens: Ensemble, n: int = 200_000, seed: int = 42) -> float:
    if ens.is_free_water:
        return 0.0
    vi, _ = _measure_vi(ens, n, np.random.default_rng(seed))
    return vi


def governing_reference_summary(ens: Ensemble) -> dict:
    return {
        "mean_A_over_V_um_inv": ens.mean_AV,
        "source": "SI §S.IV 5e6 independent single-cell reference cloud",
        "reference_path": ens.geometry.reference_path,
        "reference_metadata": ens.geometry.reference_metadata,
    }


# ----------------------------------------------------------- periodic boxes

def periodic_nearest_distances(seeds: np.ndarray, L: float) -> np.ndarray:
    """Each seed's distance to its nearest other seed on the torus of side L."""
    return nearest_points(seeds, period=float(L))[1]


def create_periodic_ensemble(n_cells: int, rho: float, vi: float, cfg: SimConfig | None = None,
                             seed: int | None = 0, verify_vi: bool = False, n_vi_points: int = 200_000,
                             with_kd: bool = True) -> Ensemble:
    """Periodic cube of n_cells uniform seeds at density rho [cells/µL].

    Side L = (n_cells / rho)^(1/3); alpha* from Eq. S7a at the target v_i;
    annulus min(alpha*, kappa d_nn / 2) with periodic nearest distances;
    walkers start uniformly in the whole box.  <A/V> is the governing-process
    table value, as for the reference's ensembles.
    """
    if cfg is None:
        cfg = SimConfig()
    rho_um3 = _rho_um3(rho)
    L = (n_cells / rho_um3) ** (1.0 / 3.0)
    alpha_star = alpha_star_from_vi(vi, rho)
    mean_av, reference = governing_mean_A_over_V(vi, rho, cfg)
    rng = np.random.default_rng(seed)
    seeds = rng.uniform(0.0, L, size=(int(n_cells), 3)).astype(np.float64)
    nn = periodic_nearest_distances(seeds, L)
    if np.any(nn <= 0.0):
        raise RuntimeError("Poisson seed realization contains coincident seeds")
    annulus = np.minimum(alpha_star, cfg.kappa * nn / 2.0).astype(np.float64)
    empty = np.zeros(0, dtype=np.int32)
    nodes = build_kdtree(seeds) if with_kd else (empty, empty.astype(np.int8), empty, empty, empty)
    cert = PopulationCertificate(0.0, 0.0, True, int(n_cells), float("inf"), True, 0.0, 0)
    stats = _geometry_stats(vi, float("nan"), float("nan"), rho, float("nan"), alpha_star, mean_av, annulus,
                            n_cells, n_cells, n_vi_points, L, L, cert, reference)
    ens = Ensemble(
        seeds=seeds, annulus=annulus, rho=float(rho), V=float("nan"), vi=float("nan"), alpha_star=alpha_star,
        L=L, source_lo=0.0, source_hi=L, mean_AV=mean_av, geometry=stats, rho_requested=float(rho),
        V_requested=vi / rho_um3 / 1.0e3, kd_node_seed=nodes[0], kd_node_axis=nodes[1], kd_node_left=nodes[2],
        kd_node_right=nodes[3], kd_node_parent=nodes[4], periodic=True,
    )
    if verify_vi:
        realised, se = _measure_vi(ens, n_vi_points, rng)
        ens.vi = realised
        ens.V = realised * L ** 3 / n_cells / 1.0e3
        ens.geometry = _geometry_stats(vi, realised, se, rho, ens.V, alpha_star, mean_av, annulus, n_cells, n_cells,
                                       n_vi_points, L, L, cert, reference)
    return ens


def periodic_images(seeds: np.ndarray, L: float, margin: float) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Seeds and their periodic images inside [-margin, L + margin)^3.

    Returns (points, label, shift) with label the cell of each point and
    shift its integer image offset (points = seeds[label] + shift * L); the
    first len(seeds) points are the seeds themselves.
    """
    k = int(math.ceil(margin / L))
    rng = range(-k, k + 1)
    shifts = np.array([(a, b, c) for a in rng for b in rng for c in rng if (a, b, c) != (0, 0, 0)], dtype=np.int64)
    pts = [seeds]
    labels = [np.arange(len(seeds), dtype=np.int32)]
    shs = [np.zeros((len(seeds), 3), dtype=np.int64)]
    for s in shifts:
        img = seeds + s * L
        keep = np.all((img >= -margin) & (img < L + margin), axis=1)
        if np.any(keep):
            pts.append(img[keep])
            labels.append(np.nonzero(keep)[0].astype(np.int32))
            shs.append(np.repeat(s[None, :], int(np.count_nonzero(keep)), axis=0))
    return np.concatenate(pts), np.concatenate(labels), np.concatenate(shs)


# ------------------------------------------------------------- device data

def _hint_grid(seeds: np.ndarray, L: float, n_hint: int, periodic: bool) -> np.ndarray:
    """The nearest seed to the centre of each cell of an n_hint^3 grid over [0, L]^3."""
    h = L / n_hint
    axis = (np.arange(n_hint) + 0.5) * h
    X, Y, Z = np.meshgrid(axis, axis, axis, indexing="ij")
    idx, _ = nearest_points(seeds, np.column_stack((X.ravel(), Y.ravel(), Z.ravel())), period=L if periodic else 0.0)
    return idx


@dataclass
class GeometryArrays:
    """Host arrays of one device geometry (all FP64 / int, C order)."""

    kd_xyz: np.ndarray
    kd_annulus: np.ndarray
    kd_label: np.ndarray
    nodes: tuple
    cell_xyz: np.ndarray
    cell_annulus: np.ndarray
    nb_offsets: np.ndarray
    nb_ids: np.ndarray
    nb_vec: np.ndarray
    hint_n: int
    hint: np.ndarray
    L: float
    source_lo: float
    source_hi: float
    periodic: bool
    free_water: bool
    pad_margin: float = 0.0
    cell_boxed: Optional[np.ndarray] = None  # 1 where a cell's neighbour list may miss a facet


def geometry_arrays(ens: Ensemble, with_kd: bool = True) -> GeometryArrays:
    """Everything the kernels need, derived from an ensemble.

    with_kd=False skips the exact classifier's KD tree (and, for periodic
    boxes, the KD padding margin): the FP32 production walk needs only the
    Delaunay graph and the hint grid.
    """
    L = float(ens.L)
    if ens.is_free_water:
        z = np.zeros(0)
        return GeometryArrays(z.reshape(0, 3), z, z.astype(np.int32), (), z.reshape(0, 3), z,
                              np.zeros(1, np.int64), z.astype(np.int32), z.reshape(0, 3), 0, z.astype(np.int32),
                              L, float(ens.source_lo), float(ens.source_hi), False, True)
    seeds = np.ascontiguousarray(ens.seeds, dtype=np.float64)
    ann = np.ascontiguousarray(ens.annulus, dtype=np.float64)
    n = len(seeds)
    spacing = (L ** 3 / n) ** (1.0 / 3.0)
    n_hint = int(min(64, max(4, round(1.5 * L / spacing))))
    if not ens.periodic:
        # cells at the population's hull are cut by the box one spacing
        # beyond the seeds; walks stay in Omega_sim, deep inside it
        offs, ids, _, boxed = voronoi_graph(seeds, n, pad=spacing)
        nb_vec = seeds[ids] - np.repeat(seeds, np.diff(offs), axis=0)
        nodes = (ens.kd_node_seed, ens.kd_node_axis, ens.kd_node_left, ens.kd_node_right, ens.kd_node_parent)
        return GeometryArrays(
            kd_xyz=seeds if with_kd else seeds[:0], kd_annulus=ann, kd_label=np.arange(n, dtype=np.int32),
            nodes=nodes if with_kd else (),
            cell_xyz=seeds, cell_annulus=ann, nb_offsets=offs, nb_ids=ids,
            nb_vec=np.ascontiguousarray(nb_vec), hint_n=n_hint, hint=_hint_grid(seeds, L, n_hint, False), L=L,
            source_lo=float(ens.source_lo), source_hi=float(ens.source_hi), periodic=False, free_water=False,
            cell_boxed=boxed,
        )
    # Periodic: pad with images.  The margin must cover (a) the exact
    # classifier's ball d1 + 2 alpha1 around any point of the box and (b)
    # twice each cell's largest vertex distance: every image within that
    # distance of a seed of the box (all of them within `margin`) is then
    # among the points its cell was cut by, so the cell is exact.
    # (twice the largest vertex distance is 2.5-3.1 spacings in boxes of
    # 50-4,000 cells, so 3.5 spacings rarely needs a second pass)
    margin = min(L, 3.5 * spacing)
    if with_kd:
        d1_upper, _ = _nearest_distance_upper_bound_periodic(seeds, L, min(spacing / 4.0, 2.0))
        margin = max(margin, min(L, d1_upper + 2.0 * float(np.max(ann)) + 1e-6), 2.0 * spacing)
    for _ in range(8):
        pts, label, shift = periodic_images(seeds, L, margin)
        offs, ids, radius, boxed = voronoi_graph(pts, n, pad=spacing)
        need = 2.0 * float(np.max(radius))
        if need <= margin and not boxed.any():
            break
        margin = max(1.25 * margin, 1.05 * need)
    else:
        raise RuntimeError("periodic Delaunay padding could not be certified")
    owner = np.repeat(np.arange(n), np.diff(offs))
    nb_vec = pts[ids] - pts[owner]
    nodes = build_kdtree(pts) if with_kd else ()
    return GeometryArrays(
        kd_xyz=np.ascontiguousarray(pts) if with_kd else pts[:0], kd_annulus=np.ascontiguousarray(ann[label]),
        kd_label=label, nodes=nodes, cell_xyz=seeds, cell_annulus=ann, nb_offsets=offs,
        nb_ids=np.ascontiguousarray(label[ids]).astype(np.int32), nb_vec=np.ascontiguousarray(nb_vec),
        hint_n=n_hint, hint=_hint_grid(seeds, L, n_hint, True), L=L, source_lo=float(ens.source_lo),
        source_hi=float(ens.source_hi), periodic=True, free_water=False, pad_margin=float(margin),
    )


def _nearest_distance_upper_bound_periodic(seeds: np.ndarray, L: float, spacing: float) -> Tuple[float, float]:
    """The periodic twin of _nearest_distance_upper_bound over one period."""
    m = max(1, int(np.ceil(L / spacing)))
    h = L / m
    axis = (np.arange(m) + 0.5) * h
    X, Y, Z = np.meshgrid(axis, axis, axis, indexing="ij")
    _, d = nearest_points(seeds, np.column_stack((X.ravel(), Y.ravel(), Z.ravel())), period=L)
    return float(np.max(d)) + 0.5 * math.sqrt(3.0) * h, h


class DeviceGeometry:
    """Owner of one madi_geom handle (device copies of the arrays)."""

    def __init__(self, ens_or_arrays):
        arr = ens_or_arrays if isinstance(ens_or_arrays, GeometryArrays) else geometry_arrays(ens_or_arrays)
        self.arrays = arr
        lib = _hip.load_library()
        _hip.require_gpu()
        keep = []

        def f64(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(a)
            return a.ctypes.data_as(_hip._f64p)

        def i32(a):
            a = np.ascontiguousarray(a, dtype=np.int32)
            keep.append(a)
            return a.ctypes.data_as(_hip._i32p)

        def i64(a):
            a = np.ascontiguousarray(a, dtype=np.int64)
            keep.append(a)
            return a.ctypes.data_as(_hip._i64p)

        def i8(a):
            a = np.ascontiguousarray(a, dtype=np.int8)
            keep.append(a)
            return a.ctypes.data_as(_hip._i8p)

        d = _hip.GeomDesc()
        d.L = arr.L
        d.source_lo = arr.source_lo
        d.source_hi = arr.source_hi
        d.periodic = int(arr.periodic)
        d.free_water = int(arr.free_water)
        if not arr.free_water:
            if len(arr.kd_xyz):
                seed, axis, left, right, parent = arr.nodes
                d.n_kd = len(arr.kd_xyz)
                d.kd_xyz = f64(arr.kd_xyz)
                d.kd_annulus = f64(arr.kd_annulus)
                d.kd_label = i32(arr.kd_label)
                d.node_seed, d.node_axis = i32(seed), i8(axis)
                d.node_left, d.node_right, d.node_parent = i32(left), i32(right), i32(parent)
            d.n_cells = len(arr.cell_xyz)
            d.cell_xyz = f64(arr.cell_xyz)
            d.cell_annulus = f64(arr.cell_annulus)
            d.n_edges = len(arr.nb_ids)
            d.nb_offsets = i64(arr.nb_offsets)
            d.nb_ids = i32(arr.nb_ids)
            d.nb_vec = f64(arr.nb_vec)
            d.hint_n = arr.hint_n
            d.hint = i32(arr.hint)
            if arr.cell_boxed is not None:
                d.cell_incomplete = i8(arr.cell_boxed)
        handle = C.c_void_p()
        _hip.check(lib.madi_geom_create(C.byref(d), C.byref(handle)), "madi_geom_create")
        self._lib = lib
        self.handle = handle

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                self._lib.madi_geom_destroy(h)
            except Exception:
                pass
            self.handle = None

    def release(self, stream: int) -> None:
        """Free the device copy in the order of `stream` without waiting.  Every
        kernel that reads the geometry must have finished: the library refuses
        (RuntimeError, MADI_ERR_BUSY) while the last one it launched is still
        queued or running, and the handle is then kept."""
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            _hip.check(self._lib.madi_geom_destroy_async(h, stream), "madi_geom_destroy_async")
            self.handle = None

    @property
    def device_bytes(self) -> int:
        out = C.c_int64()
        _hip.check(self._lib.madi_geom_device_bytes(self.handle, C.byref(out)), "madi_geom_device_bytes")
        return int(out.value)

    def classify(self, pts: np.ndarray, graph: bool = False, counters: bool = False):
        """Labels of world points: exact FP64 KD (default) or certified FP32 graph."""
        torch = _hip.require_gpu()
        n = len(pts)
        d_pts = _hip.to_device(np.asarray(pts, dtype=np.float64).reshape(n, 3))
        d_cell = torch.empty(n, dtype=torch.int32, device="cuda")
        d_in = torch.empty(n, dtype=torch.int8, device="cuda")
        d_ok = torch.empty(n, dtype=torch.int8, device="cuda")
        d_cnt = torch.zeros(_hip.N_COUNTERS, dtype=torch.int64, device="cuda")
        s = _hip.stream_ptr()
        if graph:
            rc = self._lib.madi_classify_graph(self.handle, _hip.dptr(d_pts), n, _hip.dptr(d_cell), _hip.dptr(d_in),
                                               _hip.dptr(d_ok), _hip.dptr(d_cnt), s)
            _hip.check(rc, "madi_classify_graph")
        else:
            rc = self._lib.madi_classify(self.handle, _hip.dptr(d_pts), n, _hip.dptr(d_cell), _hip.dptr(d_in),
                                         _hip.dptr(d_ok), s)
            _hip.check(rc, "madi_classify")
        torch.cuda.synchronize()
        out = (d_cell.cpu().numpy(), d_in.cpu().numpy().astype(bool), d_ok.cpu().numpy().astype(bool))
        if counters:
            return out + (dict(zip(_hip.COUNTER_NAMES, d_cnt.cpu().numpy().tolist())),)
        return out
