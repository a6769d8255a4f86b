"""The benchmark workloads as library objects (used by bench.py and the tests).

* ``c0_entry``: the CPU-runnable configuration -- a periodic cube of 64
  contracted cells (seeds uniform from seed 0), rho = 6e5 cells/µL, v_i = 0.6,
  k_io = 20 /s, D = 2.0 µm²/ms, dt = 0.05 ms x 920 steps (46 ms), 20,000
  walkers, FP64.
* ``c1_entry``: the production single entry -- 1,000 cells, rho = 5e5
  cells/µL, V = 1.2 pL, k_io = 22 /s, dt = 0.02 ms x 2,300 steps, 1e6
  walkers.
* ``c2_requests``: 256 uneven library entries (k_io log-U 1-130 /s, V log-U
  0.1-20 pL, v_i U 0.5-0.95, 50-4,000 cells, 2e5-1e6 walkers).
* ``c3_requests``: Poisson arrivals at 20 /s, entries as C2 with 1e5-1e6
  walkers.
* ``c4_requests``: the full-range sweep (V 0.01-206 pL, v_i up to 0.994,
  k_io 0-130 in 65 steps, 32-8,000 cells, 5e5 walkers, dt 0.01 ms below
  V = 0.5 pL).
Every workload is measured with the same PGSE acquisition: delta = 6 ms,
Delta = 15/25/30/40 ms, b = 0/1000/2500/4000/6000 s/mm², 24 directions
uniform on the sphere from seed 1 (480 signals).
"""

from __future__ import annotations

from typing import List, Sequence, Tuple

import numpy as np

from . import batch, engine, serve
from .config import SimConfig

DELTA_MS = 6.0
DELTAS_MS = (15.0, 25.0, 30.0, 40.0)
B_VALUES = (0.0, 1000.0, 2500.0, 4000.0, 6000.0)
N_DIRS = 24


def acquisition(b_values: Sequence[float] = B_VALUES, n_dirs: int = N_DIRS) -> engine.PhaseSpec:
    """The 4 Delta x 5 b x 24 direction acquisition (480 signals)."""
    return batch.acquisition_spec(n_dirs, delta=DELTA_MS, Deltas=DELTAS_MS, b_values=tuple(b_values))


def acquisition_with_variance(b_values: Sequence[float] = B_VALUES) -> Tuple[engine.PhaseSpec, np.ndarray,
                                                                              np.ndarray]:
    """The acquisition plus the columns that give each signal's Monte Carlo variance.

    The phase of column (pair, b, dir) is pc(b) * (g . dM) with pc
    proportional to sqrt(b), so the signal at 4b is the mean of cos(2 phi) and
    Var(cos phi) = (1 + S(4b)) / 2 - S(b)^2.  Returns (spec, idx, idx4):
    spec's b axis holds b_values followed by the missing 4b values; idx[k]
    and idx4[k] are the b-axis positions of b_values[k] and 4 b_values[k]."""
    b = [float(x) for x in b_values]
    extra = [4.0 * x for x in b if 4.0 * x not in b]
    allb = b + sorted(set(extra))
    idx = np.arange(len(b))
    idx4 = np.array([allb.index(4.0 * x) for x in b])
    return acquisition(allb), idx, idx4


def signal_se(cos_sum: np.ndarray, n: int, idx: np.ndarray, idx4: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """(S, its standard error) for the b positions idx from column sums shaped
    (pair, b, dir) over n walkers, with the 4b columns at idx4."""
    S_all = cos_sum / float(n)
    S, S4 = S_all[:, idx, :], S_all[:, idx4, :]
    var = np.maximum(0.5 * (1.0 + S4) - S * S, 0.0)
    return S, np.sqrt(var / float(n))


def c1_entry(n_walkers: int = 1_000_000):
    """(cfg, ensemble, p_p, spec) of the production single entry."""
    from .ensemble import create_periodic_ensemble
    from .walker_gpu import kio_to_pp

    rho, V, kio = 5.0e5, 1.2, 22.0
    vi = rho * V * 1e-6
    cfg = SimConfig(D0=2.0, ts=0.02, T_max_ms=46.0, h_ms=1.0, n_walkers=n_walkers, small_deltas=[DELTA_MS],
                    big_deltas=list(DELTAS_MS), b_values=list(B_VALUES), periodic=True, precision="fp32")
    ens = create_periodic_ensemble(1000, rho, vi, cfg, seed=0)
    return cfg, ens, kio_to_pp(kio, ens.mean_AV, cfg), acquisition()


def c2_requests(n: int = 256) -> List[batch.EntryRequest]:
    return batch.library_requests(n, seed=2)


def c3_requests(rate_hz: float = 20.0, duration_s: float = 600.0):
    """(requests, arrival offsets in seconds) of the served workload."""
    arrivals = serve.poisson_arrivals(rate_hz, duration_s, seed=3)
    return batch.library_requests(len(arrivals), seed=3, walkers=(100_000, 1_000_000)), arrivals


def c4_requests() -> List[batch.EntryRequest]:
    return batch.sweep_requests()


def universal_reduce(n_walkers: int = 50_000, seed: int = 7):
    """The universal library's column reduce at production size: the default
    storage grid (1245 (delta, Delta) pairs x 25 b = 31,125 columns, h = 1 ms
    over T_max = 128 ms, so 129 Y samples) over n_walkers synthetic walkers.
    Y is the integrated displacement of free diffusion at the default D0,
    sampled every h, so phases have their production magnitudes.  Returns (Y, columns)."""
    from . import signal
    cfg = SimConfig()
    cols = signal.build_columns(cfg)
    G = cfg.n_grid
    rng = np.random.default_rng(seed)
    inc = rng.normal(0.0, np.sqrt(2.0 * cfg.D0 * cfg.h_ms), size=(n_walkers, G, 3))
    inc[:, 0, :] = 0.0
    Y = np.cumsum(np.cumsum(inc, axis=1), axis=1) * cfg.h_ms
    return Y, cols
