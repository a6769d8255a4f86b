"""Walk configuration for the MI355X MADI random walk.

Units follow the reference: lengths in µm, times in ms, b in s/mm².

``SimConfig`` keeps the field names and derived quantities of the reference
configuration (``madi/config.py:178-359``) so a caller's existing keyword
arguments work unchanged.  Two fields are additions of this build and are
documented where they are declared: ``periodic`` (a torus geometry for
synthetic benchmark boxes) and ``precision`` (which walk kernel runs).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

# Physical constants (madi/config.py:36-38).
D0_UM2_MS = 3.0
D0_MOUSE = 2.0
GAMMA_RAD = 2.675222e8

# Library storage grid (madi/config.py:61-72).
H_MS = 1.0
SMALL_DELTAS_MS: List[float] = [float(v) for v in range(1, 31)]
BIG_DELTAS_MS: List[float] = [float(v) for v in range(1, 51)] + [float(v) for v in range(55, 81, 5)]
T_MAX_MS = 128.0

# Production Monte-Carlo allocation (madi/config.py:76-90).
PRODUCTION_WALKERS_PER_ENSEMBLE = 50_000
PRODUCTION_ENSEMBLES_PER_ENTRY = 40
PRODUCTION_AXIS_WALKS_PER_ENTRY = 3 * PRODUCTION_WALKERS_PER_ENSEMBLE * PRODUCTION_ENSEMBLES_PER_ENTRY
PRODUCTION_CLASSIFIER_MODE = "exact_cached"
PRODUCTION_CLASSIFIER_CACHE_DELTA_MAX_UM = 2.0
PRODUCTION_CLASSIFIER_CACHE_MIN_SAFE_RADIUS_UM = 0.0
PRODUCTION_CLASSIFIER_CACHE_CANDIDATE_CAPACITY = 256

B_MAX_S_MM2 = 12000.0
B_STEP_S_MM2 = 500.0

# v5 per-ensemble diagnostic subset (madi/config.py:107-113): the order of
# these pairs and b values is part of the on-disk library format.
ENSEMBLE_MEAN_SUBSET_DELTA_PAIRS_MS: Tuple[Tuple[float, float], ...] = (
    (5.0, 15.0), (7.0, 25.0), (10.0, 30.0), (12.0, 36.0),
    (15.0, 40.0), (20.0, 50.0), (25.0, 60.0), (30.0, 80.0),
)
ENSEMBLE_MEAN_SUBSET_B_VALUES_S_MM2 = np.arange(int(B_MAX_S_MM2 / B_STEP_S_MM2) + 1, dtype=float) * B_STEP_S_MM2


def evenly_spaced_bvalues(step: float = B_STEP_S_MM2, b_max: float = B_MAX_S_MM2) -> np.ndarray:
    """b grid 0, step, ..., b_max (madi/config.py:135-146)."""
    count = int(round(b_max / step)) + 1
    return step * np.arange(count, dtype=float)


def sqrt_spaced_bvalues(b_max: float = B_MAX_S_MM2, n: int = 28) -> np.ndarray:
    """b values uniform in sqrt(b) (madi/config.py:116-132)."""
    return np.square(np.linspace(0.0, np.sqrt(b_max), n))


def valid_delta_pairs(small_deltas=None, big_deltas=None) -> List[Tuple[float, float]]:
    """Triangular (δ, Δ) grid with Δ ≥ δ, δ-major (madi/config.py:149-153)."""
    smalls = SMALL_DELTAS_MS if small_deltas is None else small_deltas
    bigs = BIG_DELTAS_MS if big_deltas is None else big_deltas
    pairs = []
    for d in smalls:
        for D in bigs:
            if D >= d:
                pairs.append((float(d), float(D)))
    return pairs


def grid_time_index(t_ms: float, h_ms: float = H_MS, tol: float = 1e-6) -> int:
    """Index of ``t_ms`` on the Y(t) grid; raises off-grid (madi/config.py:156-172)."""
    ratio = t_ms / h_ms
    nearest = round(ratio)
    if abs(ratio - nearest) > tol:
        raise ValueError(
            f"t={t_ms} ms is not an exact integer multiple of h={h_ms} ms "
            "(grid misalignment: Y would have to be interpolated, which is not allowed)."
        )
    return int(nearest)


@dataclass
class SimConfig:
    """Walk, geometry and acquisition knobs (field names of madi/config.py:178-258)."""

    D0: float = D0_UM2_MS
    ts: float = 1e-3
    T_max_ms: float = T_MAX_MS
    n_walkers: int = PRODUCTION_WALKERS_PER_ENSEMBLE
    n_ensembles: int = 4
    walker_chunk: Optional[int] = None
    classifier_mode: str = "exact"
    classifier_cache_delta_max_um: float = 1.0
    classifier_cache_min_safe_radius_um: float = 0.0
    classifier_cache_candidate_capacity: int = 256
    L: Optional[float] = None
    source_fraction: float = 0.40
    kappa: float = 0.90
    population_initial_margin_cell_spacings: float = 3.0
    population_certification_spacing_um: float = 8.0
    population_max_expansions: int = 8
    geometry_reference_path: Optional[str] = None
    allow_uncertified_geometry_reference: bool = False
    geometry_reference_required_cells: int = 5_000_000
    geometry_reference_required_alpha_values: int = 26
    geometry_validation_points: int = 200_000
    geometry_vi_tolerance: float = 0.005
    boundary_mode: str = "si_fatal_escape"
    phase_model: str = "finite_lobe"
    h_ms: float = H_MS
    small_deltas: list = field(default_factory=lambda: list(SMALL_DELTAS_MS))
    big_deltas: list = field(default_factory=lambda: list(BIG_DELTAS_MS))
    b_values: list = field(default_factory=lambda: list(evenly_spaced_bvalues()))
    # Additions of this build.
    # ``periodic``: geometry is a torus of side L (synthetic benchmark boxes);
    # the walk then never escapes.  The reference has no periodic mode, so
    # parity runs leave this False.
    periodic: bool = False
    # ``precision``: "fp64" runs the reference-semantics walk (exact KD
    # classifier at every endpoint, FP64 state); "fp32" runs the production
    # walk (FP32 seed-relative state, Delaunay-graph classifier certified in
    # FP64, FP64 final sums).
    precision: str = "fp64"

    @property
    def sigma(self) -> float:
        """Per-axis Gaussian step std sqrt(2 D0 ts) [µm]."""
        return float(np.sqrt(2.0 * self.D0 * self.ts))

    @property
    def ls_rms(self) -> float:
        """RMS 3-D step length sqrt(6 D0 ts) [µm]."""
        return float(np.sqrt(6.0 * self.D0 * self.ts))

    @property
    def n_steps(self) -> int:
        return int(round(self.T_max_ms / self.ts))

    @property
    def tRW_max(self) -> float:
        return self.n_steps * self.ts

    @property
    def steps_per_h(self) -> int:
        ratio = self.h_ms / self.ts
        k = round(ratio)
        if abs(ratio - k) > 1e-6:
            raise ValueError(
                f"h_ms={self.h_ms} is not an exact integer multiple of ts={self.ts}: "
                "Y(t) samples would drift off the step grid."
            )
        return int(k)

    @property
    def n_grid(self) -> int:
        return int(round(self.T_max_ms / self.h_ms)) + 1

    @staticmethod
    def tD(delta: float, Delta: float) -> float:
        return Delta - delta / 3.0

    def delta_pairs(self) -> List[Tuple[float, float]]:
        return valid_delta_pairs(self.small_deltas, self.big_deltas)

    def assert_grid_alignment(self) -> None:
        """Raise on any inconsistent knob (checks of madi/config.py:307-359)."""
        for d in self.small_deltas:
            grid_time_index(d, self.h_ms)
        for D in self.big_deltas:
            grid_time_index(D, self.h_ms)
        longest = 0.0
        for d, D in self.delta_pairs():
            grid_time_index(D + d, self.h_ms)
            longest = max(longest, D + d)
        if longest > self.T_max_ms:
            raise ValueError(f"(δ,Δ) grid needs Δ+δ up to {longest} ms but T_max_ms={self.T_max_ms} ms.")
        self.steps_per_h  # raises when h is not a multiple of ts
        checks = [
            (self.phase_model in ("finite_lobe", "narrow_pulse"),
             f"phase_model must be 'finite_lobe' or 'narrow_pulse', got {self.phase_model!r}."),
            (self.boundary_mode == "si_fatal_escape",
             "boundary_mode must be 'si_fatal_escape'; periodic wrapping and survivor pruning "
             f"contradict SI §S.III, got {self.boundary_mode!r}."),
            (0.0 < self.source_fraction <= 1.0, "source_fraction must lie in (0, 1]."),
            (abs(self.kappa - 0.90) <= 1e-12, "kappa must be the SI-specified 0.90."),
            (self.population_initial_margin_cell_spacings > 0.0,
             "population_initial_margin_cell_spacings must be positive."),
            (self.population_certification_spacing_um > 0.0, "population_certification_spacing_um must be positive."),
            (self.population_max_expansions >= 1, "population_max_expansions must be positive."),
            (self.geometry_validation_points > 0, "geometry_validation_points must be positive."),
            (self.geometry_reference_required_cells >= 1, "geometry_reference_required_cells must be positive."),
            (self.geometry_reference_required_alpha_values >= 1,
             "geometry_reference_required_alpha_values must be positive."),
            (self.classifier_mode in ("exact", "exact_cached"), "classifier_mode must be 'exact' or 'exact_cached'."),
            (self.classifier_cache_delta_max_um > 0.0, "classifier_cache_delta_max_um must be positive."),
            (self.classifier_cache_min_safe_radius_um >= 0.0,
             "classifier_cache_min_safe_radius_um must be non-negative."),
            (self.classifier_cache_candidate_capacity >= 2,
             "classifier_cache_candidate_capacity must be at least two."),
            (self.precision in ("fp64", "fp32"), "precision must be 'fp64' or 'fp32'."),
        ]
        for ok, message in checks:
            if not ok:
                raise ValueError(message)

    @property
    def grid_size(self) -> int:
        """The exact classifier has no voxel grid (madi/config.py:361-369)."""
        raise RuntimeError("the SI-exact classifier has no voxel cache grid")
