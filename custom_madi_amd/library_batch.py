"""Library builds and library serving on the batched device scheduler.

``build_library_from_triplets`` (madi/library.py:445-642) walks one (rho, V)
group at a time: the group's ``n_ensembles`` ensembles (Eq. S8 geometry, or
periodic boxes) each walked once per k_io, one launch per walk.  Here every
(group, ensemble, k_io, walker chunk) walk is an entry of ``DeviceScheduler``:
walks of uneven size from many groups share persistent launches, and the
host geometry of the next groups is built while the device walks the
previous ones.

Each walk keeps the sequential path's inputs (ensemble e of the group,
geometry seed and walk seed of ensemble e, p_p of k_io on that ensemble, the
walker chunks of ``cfg.walker_chunk`` with seeds walk_seed + offset), and a
scheduled walk's sums equal the one-entry launch bit for bit, so the merged
results (chunks in order, then ensembles in order, as
walker_gpu._walk_and_reduce_one_ensemble and _merge_results do) and the
library entries built from them are the sequential build's, bit for bit.

Two orders:

* ``ordered=True`` (library builds): a group's entries are released only
  after every earlier group's, so the library grows in the sequential
  order and each checkpoint holds what the sequential build's checkpoint
  after the same group holds.  A failed walk (escape: the reference's fatal
  boundary condition, or a classifier failure) raises the sequential path's
  RuntimeError when its group's turn comes, after the earlier groups have
  been released and checkpointed.
* ``ordered=False`` (serving): a group's entries are released as soon as its
  walks finish; a failed group is reported and left out, the others go on.

The FP32 walk is batched (FP64 builds take the sequential path), in the
sequential FP32 path's two forms: a column set that fuses (at most 8
(delta, Delta) pairs x 16 b values) runs the fused walk; any other (the
universal grid, every v5 build) runs the scheduler's Y mode, walks that
store Y followed by the FP64 column reduce, as walk_y_f32 + reduce_columns.
"""

from __future__ import annotations

import os
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _hip, batch, engine
from . import signal as sig
from .config import SimConfig
from .ensemble import GeometryArrays


@dataclass
class GroupEnsemble:
    """One ensemble of a (rho, V) group, as the walks need it."""

    arrays: GeometryArrays
    mean_AV: float
    is_free_water: bool
    geometry: dict


def prepare_ensemble(rho: float, V: float, cfg: SimConfig, seed: int, index: int) -> GroupEnsemble:
    """Ensemble `index` of one (rho, V) group, as the sequential build
    constructs it (walker_gpu._build_ensemble_for_entry, the same seed),
    reduced to the arrays its walks upload (those DeviceGeometry(ens) holds on
    the sequential path) and what the entry metadata needs.  Building a
    cellular ensemble measures v_i with the HIP classifier, so a worker
    process running this holds a HIP context of its own."""
    from .walker_gpu import _build_ensemble_for_entry
    ens = _build_ensemble_for_entry(rho, V, cfg, seed, index)
    return GroupEnsemble(ens.kernel_arrays(), float(ens.mean_AV), bool(ens.is_free_water), ens.geometry.to_dict())


def prepare_ensemble_host(rho: float, V: float, cfg: SimConfig, seed: int, index: int) -> tuple:
    """The CPU part of prepare_ensemble, for a worker process with no GPU:
    ("done", GroupEnsemble) for the free-water and periodic ensembles, whose
    construction is CPU only; ("draft", EnsembleDraft, arrays) for an Eq. S8
    ensemble, whose v_i measurement finish_prepared runs on the GPU."""
    from .ensemble import build_ensemble_host, geometry_arrays
    from .walker_gpu import _build_ensemble_for_entry, _ensemble_seed
    if rho > 0.0 and V > 0.0 and not cfg.periodic:
        draft = build_ensemble_host(rho, V, cfg, _ensemble_seed(seed, index))
        return ("draft", draft, geometry_arrays(draft.ens))
    ens = _build_ensemble_for_entry(rho, V, cfg, seed, index)
    return ("done", GroupEnsemble(ens.kernel_arrays(), float(ens.mean_AV), bool(ens.is_free_water),
                                  ens.geometry.to_dict()))


def finish_prepared(item: tuple, cfg: SimConfig) -> GroupEnsemble:
    """The rest of prepare_ensemble for a prepare_ensemble_host result (in the
    process that holds the GPU): the v_i measurement and acceptance test of
    create_ensemble, on the arrays the worker derived."""
    if item[0] == "done":
        return item[1]
    from .ensemble import finish_ensemble
    _, draft, arrays = item
    ens = finish_ensemble(draft, cfg, arrays=arrays)
    return GroupEnsemble(arrays, float(ens.mean_AV), bool(ens.is_free_water), ens.geometry.to_dict())


def prepare_group(rho: float, V: float, cfg: SimConfig, seed: int) -> List[GroupEnsemble]:
    """Every ensemble of one (rho, V) group (prepare_ensemble in index order)."""
    return [prepare_ensemble(rho, V, cfg, seed, i) for i in range(int(cfg.n_ensembles))]


def batchable(cfg: SimConfig, columns) -> bool:
    """The scheduler runs FP32 walks (any column set)."""
    return cfg.precision == "fp32" and len(columns.phase_coef) <= _hip.MADI_MAX_REDUCE_COLS


@dataclass
class _Group:
    index: int
    rho: float
    V: float
    kios: List[float]
    submitted: float
    ensembles: Optional[List[GroupEnsemble]] = None
    walks: Dict[tuple, object] = field(default_factory=dict)  # (e, kio index, offset) -> EntryResult
    n_walks: int = 0
    error: Optional[str] = None


@dataclass
class GroupResult:
    """A finished (rho, V) group: its library entries in k_io order, or its error."""

    rho: float
    V: float
    kios: List[float]
    entries: list
    error: Optional[str]
    submitted: float
    finished: float

    @property
    def latency_s(self) -> float:
        return self.finished - self.submitted


class LibraryBuilder:
    """Submit (rho, V, k_io list) groups; ``step()`` prepares geometry, admits
    walks, launches generations and releases finished groups as
    :class:`GroupResult` (in submission order when ``ordered``)."""

    def __init__(self, cfg: SimConfig, columns=None, seed: int = 0, entry_weights=None, ordered: bool = True,
                 pool=None, prepare: Callable = prepare_group, **scheduler_kw):
        cfg.assert_grid_alignment()
        if cfg.phase_model == "narrow_pulse":
            raise NotImplementedError("the narrow-pulse diagnostic is CPU-only in the reference and not built here")
        self.cfg = cfg
        self.columns = columns if columns is not None else sig.build_columns(cfg)
        if cfg.precision != "fp32":
            raise ValueError("the batched library build runs the FP32 walk (precision 'fp32')")
        cols = self.columns
        fused = engine.spec_from_columns(cols.j_delta, cols.j_Delta, cols.j_sum, cols.phase_coef)
        self.y_mode = fused is None
        if self.y_mode:
            scheduler_kw["y_columns"] = (cols.j_delta, cols.j_Delta, cols.j_sum, cols.phase_coef)
            self.spec = None
        else:
            self.spec, self.col_pair, self.col_b = fused
        self.seed = int(seed)
        self.entry_weights = entry_weights
        self.ordered = bool(ordered)
        self.pool = pool
        self.prepare = prepare
        self.sch = batch.DeviceScheduler(self.spec, **scheduler_kw)
        self.groups: Dict[int, _Group] = {}
        self.preparing: deque = deque()  # (group, AsyncResult or None)
        self.waiting: deque = deque()    # PreparedEntry not yet admitted
        self.next_index = 0
        self.next_release = 0
        self.released: List[GroupResult] = []
        chunk = int(cfg.walker_chunk or cfg.n_walkers)
        self.chunks = [(off, min(chunk, int(cfg.n_walkers) - off)) for off in range(0, int(cfg.n_walkers), chunk)]

    # ------------------------------------------------------------ submission
    def submit(self, rho: float, V: float, kios: Sequence[float]) -> int:
        """Queue one (rho, V) group walked at every k_io in kios; returns its index."""
        kios = sorted(float(k) for k in kios)
        if any(k < 0.0 for k in kios):
            raise ValueError("k_io grid values must be non-negative")
        g = _Group(self.next_index, float(rho), float(V), kios, time.perf_counter())
        self.groups[g.index] = g
        self.next_index += 1
        jobs = None
        if self.pool is not None:
            # one job per ensemble: a group's ensembles are built in parallel
            jobs = [self.pool.apply_async(prepare_ensemble_host, (g.rho, g.V, self.cfg, self.seed, i))
                    for i in range(int(self.cfg.n_ensembles))]
        self.preparing.append((g, jobs))
        return g.index

    def _walks_of(self, g: _Group) -> List[batch.PreparedEntry]:
        from .walker_gpu import _checked_pp, _walk_seed, kio_to_pp
        cfg = self.cfg
        out = []
        for e, ge in enumerate(g.ensembles):
            ws = _walk_seed(self.seed, e)
            for ki, kio in enumerate(g.kios):
                pp = 0.0 if (ge.is_free_water or kio == 0.0) else _checked_pp(kio_to_pp(kio, ge.mean_AV, cfg))
                for off, n_this in self.chunks:
                    req = batch.EntryRequest(key=(g.index, e, ki, off), kio=kio, rho=g.rho, V=g.V, n_cells=0,
                                             n_walkers=n_this, walk_seed=ws + off, D0=cfg.D0, ts=cfg.ts,
                                             T_max_ms=cfg.T_max_ms, h_ms=cfg.h_ms, submitted=g.submitted)
                    out.append(batch.PreparedEntry(req, ge.arrays, pp, ge.mean_AV, 0.0, 0.0, 0.0, cfg=cfg))
        g.n_walks = len(out)
        return out

    def _collect_prepared(self, block: bool) -> None:
        while self.preparing:
            g, jobs = self.preparing[0]
            if jobs is None:
                # built here: one group per step, so the device walks the
                # groups already admitted meanwhile
                g.ensembles = self.prepare(g.rho, g.V, self.cfg, self.seed)
                self.preparing.popleft()
                self.waiting.extend(self._walks_of(g))
                return
            if not block and not all(j.ready() for j in jobs):
                return
            g.ensembles = [finish_prepared(j.get(), self.cfg) for j in jobs]
            self.preparing.popleft()
            self.waiting.extend(self._walks_of(g))
            block = False  # at most one blocking wait per step

    # ---------------------------------------------------------------- device
    def _admit(self) -> None:
        while self.waiting and self.sch.admit(self.waiting[0]):
            self.waiting.popleft()

    def _absorb(self, results) -> None:
        for r in results:
            gi, e, ki, off = r.key
            g = self.groups[gi]
            g.walks[(e, ki, off)] = r

    def _group_error(self, g: _Group) -> Optional[str]:
        # the first failure in the sequential order (ensemble, k_io, chunk),
        # with the sequential path's message
        for e in range(len(g.ensembles)):
            for ki in range(len(g.kios)):
                for off, n_this in self.chunks:
                    r = g.walks[(e, ki, off)]
                    if r.n_error:
                        return "GPU full-facet classification failed"
                    if r.n_escaped:
                        return f"SI §S.III fatal boundary condition: {r.n_escaped}/{n_this} walkers exited Ω_sim."
        return None

    def _finish(self, g: _Group) -> GroupResult:
        from . import library as L
        from .walker_gpu import ReducedResult, _merge_results, pp_to_kio_eq5
        cfg, cols = self.cfg, self.columns
        err = self._group_error(g)
        entries = []
        if err is None:
            n_cols = len(cols.phase_coef)
            for ki, kio in enumerate(g.kios):
                per_ens = []
                for e, ge in enumerate(g.ensembles):
                    cos_total = np.zeros(n_cols)
                    sin_total = np.zeros(n_cols)
                    occupancy = np.zeros(cfg.n_grid, dtype=np.int64)
                    total = 0
                    for off, n_this in self.chunks:
                        r = g.walks[(e, ki, off)]
                        if self.y_mode:
                            cos_total += r.cos
                            sin_total += r.sin
                        else:
                            cos_total += r.cos.sum(axis=2)[self.col_pair, self.col_b]
                            sin_total += r.sin.sum(axis=2)[self.col_pair, self.col_b]
                        occupancy += r.occupancy
                        total += n_this
                    pp = g.walks[(e, ki, 0)].pp
                    per_ens.append(ReducedResult(
                        cos_sum=cos_total, sin_sum=sin_total, n_walkers=total, n_escaped=0,
                        occupancy_counts=occupancy, pp_values=[float(pp)],
                        analytic_kio_eq5_values=[pp_to_kio_eq5(pp, ge.mean_AV, cfg)] if not ge.is_free_water
                        else [0.0],
                        geometry_stats=[ge.geometry], ensemble_cos_means=[cos_total / float(3 * total)],
                        ensemble_sin_means=[sin_total / float(3 * total)], classifier_cache_stats=[{}],
                    ))
                result = sig._assemble(_merge_results(per_ens), cols)
                entries.append(L._entry_from_result(kio, g.rho, g.V, result, cols, cfg, self.entry_weights))
        del self.groups[g.index]
        return GroupResult(g.rho, g.V, g.kios, entries, err, g.submitted, time.perf_counter())

    def _release(self) -> List[GroupResult]:
        done = [g for g in self.groups.values() if g.ensembles is not None and g.n_walks and len(g.walks) == g.n_walks]
        out = []
        if self.ordered:
            ready = {g.index for g in done}
            while self.next_release in ready:
                out.append(self._finish(self.groups[self.next_release]))
                self.next_release += 1
        else:
            for g in sorted(done, key=lambda x: x.index):
                out.append(self._finish(g))
        self.released += out
        return out

    def step(self, block: bool = False) -> List[GroupResult]:
        """Advance every stage once; returns the groups released by this step."""
        self._collect_prepared(block and not self.sch.active and not self.waiting)
        self._admit()
        launched = self.sch.launch_generation()
        if not launched and not self.sch.inflight and self.waiting:
            raise MemoryError("a walk does not fit in free HBM even with the device idle")
        self._absorb(self.sch.poll(block=block and not launched and bool(self.sch.inflight)))
        self._admit()
        return self._release()

    def idle(self) -> bool:
        return not (self.groups or self.preparing or self.waiting or self.sch.active or self.sch.inflight)

    def drain(self) -> List[GroupResult]:
        out = []
        while not self.idle():
            self._collect_prepared(True)
            self._admit()
            if not self.sch.launch_generation() and not self.sch.inflight and self.waiting:
                raise MemoryError("a walk does not fit in free HBM even with the device idle")
            self._absorb(self.sch.poll(block=True))
            out += self._release()
        return out

    def close(self) -> None:
        # generations still in flight read the admitted walks' geometry and
        # buffers: let them finish before any of it is released
        self.sch.torch.cuda.synchronize()
        self.sch.close()


class LibraryCheckpoint:
    """A v5 library file rewritten from the entries released so far (the
    format of madi/library.py:872-972; each write goes to a temporary file
    that replaces the previous one)."""

    def __init__(self, path: Optional[str], cfg: SimConfig, columns, library: Optional[list] = None,
                 grid_metadata=None, build_seed: Optional[int] = None):
        self.path = path
        self.cfg, self.columns = cfg, columns
        self.grid_metadata, self.build_seed = grid_metadata, build_seed
        self.library = list(library) if library is not None else []
        self.writes = 0

    def extend(self, entries) -> None:
        self.library.extend(entries)

    def write(self) -> None:
        if not self.path:
            return
        from . import library as L
        # numpy.savez's rule: a path without the .npz suffix gets it
        target = self.path if str(self.path).endswith(".npz") else f"{self.path}.npz"
        tmp = f"{target}.partial.npz"
        L._save_library(self.library, tmp, cfg=self.cfg, columns=self.columns, grid_metadata=self.grid_metadata,
                        build_seed=self.build_seed)
        os.replace(tmp, target)
        self.writes += 1


def build_groups_batched(groups: Sequence[Tuple[Tuple[float, float], List[float]]], cfg: SimConfig, columns,
                         seed: int, checkpoint: LibraryCheckpoint, entry_weights=None, pool=None,
                         max_ahead: int = 4, verbose: bool = False, prepare: Callable = prepare_group,
                         **scheduler_kw) -> None:
    """The library build of madi/library.py:546-629 on the scheduler: groups in
    the given order, entries appended and checkpointed group by group in that
    order; raises the sequential path's RuntimeError for the first group with
    a failed walk (after checkpointing the groups before it)."""
    b = LibraryBuilder(cfg, columns, seed=seed, entry_weights=entry_weights, ordered=True, pool=pool,
                       prepare=prepare, **scheduler_kw)
    try:
        todo = deque(groups)

        def take(results):
            for gr in results:
                if gr.error is not None:
                    raise RuntimeError(gr.error)
                checkpoint.extend(gr.entries)
                checkpoint.write()
                if verbose:
                    print(f"    rho={gr.rho:g} V={gr.V:g}: {len(gr.entries)} entries in {gr.latency_s:.1f} s")

        while todo or not b.idle():
            # keep a few groups' geometry in preparation ahead of the device
            while todo and len(b.groups) < max_ahead:
                (rho, V), kios = todo.popleft()
                b.submit(rho, V, kios)
            take(b.step(block=not todo or len(b.groups) >= max_ahead))
        take(b.drain())
    finally:
        b.close()


def serve_library(builder: LibraryBuilder, arrivals: Sequence[Tuple[float, float, float, Sequence[float]]],
                  checkpoint: Optional[LibraryCheckpoint] = None) -> Tuple[List[GroupResult], List[GroupResult]]:
    """Serve (arrival time in s from now, rho, V, k_io list) requests with
    continuous batching: each is submitted at its time, and each finished
    group's entries are appended to the v5 checkpoint at once (completion
    order; builder.ordered should be False).  Returns (finished, failed):
    a failed group is left out of the library and the others go on."""
    todo = deque(sorted(arrivals, key=lambda a: a[0]))
    t0 = time.perf_counter()
    done, failed = [], []

    def take(results):
        for gr in results:
            if gr.error is not None:
                failed.append(gr)
                continue
            done.append(gr)
            if checkpoint is not None:
                checkpoint.extend(gr.entries)
                checkpoint.write()

    while todo or not builder.idle():
        now = time.perf_counter() - t0
        while todo and todo[0][0] <= now:
            _, rho, V, kios = todo.popleft()
            builder.submit(rho, V, kios)
        if builder.idle():
            time.sleep(min(0.01, max(0.0, todo[0][0] - now)))
            continue
        take(builder.step(block=not todo))
    take(builder.drain())
    return done, failed
