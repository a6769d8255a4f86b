"""Autotuner for the batched FP32 walk.

The knobs it searches change only how the GPU does the work, never the
result (the batch tests check bitwise equality across them):

* ``classify_mode``: 1 per-lane state machine per tile, 3 per-lane state
  machine with lane refill, 4 walker waves with classifier waves;
* ``scan_rounds``: classification rounds per state-machine iteration (1, 2),
  or for mode 4 the workgroup shape (1: 3 walker + 1 classifier waves, 2: 2 + 2);
* ``grid_per_cu``: persistent workgroups per CU of the batch kernel.

The plane-tier cap is left at its default: it changes which proposals take a
full classification and therefore where the anchor moves, i.e. the low-order
bits of the result.

A trial runs a fixed sample of uneven entries (seeded, prepared once) through
``DeviceScheduler`` with each candidate and keeps the fastest.  The choice is
cached in a JSON file keyed by device, architecture, CU count and plan shape
(``MADI_TUNE_CACHE`` or ``~/.cache/custom_madi_amd/tune.json``).
"""

from __future__ import annotations

import json
import os
import time
from dataclasses import asdict, dataclass
from pathlib import Path
from typing import Iterable, List, Optional

from . import _hip, engine

DEFAULT_CANDIDATES = [
    (4, 1, 3), (4, 1, 4),
    (3, 1, 3), (3, 1, 4),
    (1, 2, 4),
]  # (classify_mode, scan_rounds, grid_per_cu)


@dataclass
class TuneChoice:
    classify_mode: int
    scan_rounds: int
    grid_per_cu: int
    trial_s: float = 0.0

    def kwargs(self) -> dict:
        return {"classify_mode": self.classify_mode, "scan_rounds": self.scan_rounds,
                "grid_per_cu": self.grid_per_cu}


def cache_path() -> Path:
    p = os.environ.get("MADI_TUNE_CACHE")
    return Path(p) if p else Path.home() / ".cache" / "custom_madi_amd" / "tune.json"


def _key(spec: engine.PhaseSpec, tag: str = "") -> str:
    info = _hip.device_info()
    key = f"{info['name']}|{info['arch']}|{info['cu_count']}|{tuple(spec.shape)}"
    return f"{key}|{tag}" if tag else key


def _load() -> dict:
    try:
        return json.loads(cache_path().read_text())
    except (OSError, ValueError):
        return {}


def cached(spec: engine.PhaseSpec, tag: str = "") -> Optional[TuneChoice]:
    rec = _load().get(_key(spec, tag))
    return TuneChoice(**rec) if rec else None


def _store(spec: engine.PhaseSpec, choice: TuneChoice, tag: str = "") -> None:
    path = cache_path()
    data = _load()
    data[_key(spec, tag)] = asdict(choice)
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        path.write_text(json.dumps(data, indent=1))
    except OSError:
        pass  # a read-only home only loses the cache


def sample_entries(n: int = 6, walkers=(250_000, 600_000), seed: int = 71):
    """Prepared uneven entries for trials (host geometry built here)."""
    from . import batch
    reqs = batch.library_requests(n, seed=seed, walkers=walkers, cells=(100, 2000))
    return [batch.prepare_entry(r) for r in reqs]


def autotune(spec: engine.PhaseSpec, preps: Optional[list] = None,
             candidates: Iterable[tuple] = DEFAULT_CANDIDATES, use_cache: bool = True,
             log=None, tag: str = "") -> TuneChoice:
    """Time each candidate on the sample and return (and cache) the fastest.

    ``preps`` is the sample (default: ``sample_entries()``); a workload with
    its own sample passes a ``tag`` so its choice is cached apart (the best
    variant depends on how often a workload's walkers need a classification).
    """
    from . import batch
    if use_cache:
        c = cached(spec, tag)
        if c is not None:
            return c
    torch = _hip.require_gpu()
    preps = preps if preps is not None else sample_entries()
    best: Optional[TuneChoice] = None
    trials: List[TuneChoice] = []
    batch.run_batched(preps[:1], spec)  # first-use costs (module load, allocator) out of the trials
    for mode, rounds, per_cu in candidates:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        batch.run_batched(preps, spec, classify_mode=mode, scan_rounds=rounds, grid_per_cu=per_cu)
        torch.cuda.synchronize()
        ch = TuneChoice(mode, rounds, per_cu, time.perf_counter() - t0)
        trials.append(ch)
        if log:
            log(f"[tune] classify_mode={mode} scan_rounds={rounds} grid_per_cu={per_cu}: {ch.trial_s * 1e3:.1f} ms")
        if best is None or ch.trial_s < best.trial_s:
            best = ch
    _store(spec, best, tag)
    return best
