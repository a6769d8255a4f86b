"""Batched and served library-entry simulation on one GPU.

An entry is one (k_io, rho, V) simulation request with its own cell count and
walker count.  Its walkers are cut into tiles of 256 (one workgroup each).
``DeviceScheduler`` keeps admitted entries' geometry and per-tile result slots
on the device and launches *generations*: one persistent work-stealing kernel
(madi_batch_launch) over the pending tiles of every active entry.  New entries
join the next generation as soon as they are admitted, so entries of uneven
size share launches and the GPU stays full; an entry's tile slots are summed in
tile order, and copied to pinned host memory, right behind the generation that
walks its last tiles, on the same stream.
Because every tile writes its own slot, an entry's result does not depend on
which other entries shared its launches.

Host geometry preparation (seeds, annulus, Delaunay graph, hint grid) is pure
CPU work and runs in worker processes (``prepare_entry``), overlapped with the
device work.  The plain form (``run_plain``) is the one-entry-per-launch path:
prepare, upload, one launch, synchronise, next entry.
"""

from __future__ import annotations

import ctypes as C
import sys
import time
from collections import deque
from dataclasses import dataclass
from typing import Callable, Iterable, List, Optional

import numpy as np

from . import _hip, engine
from .config import SimConfig
from .ensemble import DeviceGeometry, GeometryArrays, create_periodic_ensemble, geometry_arrays


@dataclass
class EntryRequest:
    """One library entry to simulate (periodic box of n_cells contracted cells)."""

    key: tuple
    kio: float
    rho: float        # cells / µL
    V: float          # pL
    n_cells: int
    n_walkers: int
    seed: int = 0        # geometry seed
    walk_seed: int = 0
    D0: float = 2.0
    ts: float = 0.02
    T_max_ms: float = 46.0
    h_ms: float = 1.0
    submitted: float = 0.0

    def config(self) -> SimConfig:
        return SimConfig(D0=self.D0, ts=self.ts, T_max_ms=self.T_max_ms, h_ms=self.h_ms, n_walkers=self.n_walkers,
                         small_deltas=[1.0], big_deltas=[1.0], b_values=[0.0], periodic=True, precision="fp32")


@dataclass
class PreparedEntry:
    req: EntryRequest
    arrays: GeometryArrays
    pp: float
    mean_AV: float
    alpha_star: float
    L: float
    prep_s: float
    cfg: Optional[SimConfig] = None  # the walk's configuration (default: req.config())

    def walk_config(self) -> SimConfig:
        return self.cfg if self.cfg is not None else self.req.config()


def prepare_entry(req: EntryRequest) -> PreparedEntry:
    """Host geometry for one entry (CPU only; safe to run in a worker process)."""
    from .walker_gpu import kio_to_pp

    t0 = time.perf_counter()
    cfg = req.config()
    vi = req.rho * req.V * 1e-6
    ens = create_periodic_ensemble(int(req.n_cells), req.rho, vi, cfg, seed=req.seed, with_kd=False)
    arrays = geometry_arrays(ens, with_kd=False)
    pp = 0.0 if req.kio == 0.0 else kio_to_pp(req.kio, ens.mean_AV, cfg)
    return PreparedEntry(req, arrays, pp, ens.mean_AV, ens.alpha_star, ens.L, time.perf_counter() - t0)


@dataclass
class EntryResult:
    key: tuple
    cos: np.ndarray        # (n_pairs, n_b, n_dirs) sums over walkers
    sin: np.ndarray
    occupancy: np.ndarray  # (n_grid,)
    n_walkers: int
    n_escaped: int
    n_error: int
    pp: float
    submitted: float
    admitted: float
    finished: float
    prep_s: float
    error: Optional[str] = None  # set when the entry failed (escape or classifier failure)

    @property
    def failed(self) -> bool:
        return self.error is not None

    @property
    def signal(self) -> np.ndarray:
        return self.cos / self.n_walkers

    @property
    def latency_s(self) -> float:
        return self.finished - self.submitted


@dataclass
class _Active:
    prep: PreparedEntry
    geom: DeviceGeometry
    params: _hip.WalkParams
    n_tiles: int
    partial: object
    occ: object
    flags: object
    moments: object
    next_tile: int = 0
    admitted: float = 0.0
    device_bytes: int = 0
    Y: object = None  # Y mode: the entry's (n_walkers, n_grid, 3) FP64 samples


class DeviceScheduler:
    """Admission, generations and finalisation of entries on the current GPU."""

    def __init__(self, spec: Optional[engine.PhaseSpec], grid_per_cu: int = 0, tier_cap_um: float = 0.0,
                 scan_rounds: int = 0, classify_mode: int = 0,
                 max_tiles_per_generation: int = 65536, hbm_reserve_bytes: int = 8 << 30,
                 max_inflight: int = 2, use_graphs: bool = True, graph_capacity: int = 1024,
                 y_columns: Optional[tuple] = None):
        """spec: the fused phase plan.  Y mode (y_columns = (j_delta, j_Delta,
        j_sum, phase_coef), any number of columns; spec is then unused): each
        entry's walkers store Y, and behind the entry's last generation the
        FP64 column reduce (madi_reduce_f64) sums those columns; results hold
        the column sums (n_columns,) summed over the three axes."""
        torch = _hip.require_gpu()
        self.torch = torch
        self.lib = _hip.load_library()
        self.y_n = None
        if y_columns is not None:
            if int(classify_mode) not in (0, 3, 4):
                raise ValueError("Y mode needs a streamed walk (classify_mode 0, 3 or 4)")
            jd, jD, js, pc = y_columns
            if not (len(jd) == len(jD) == len(js) == len(pc)):
                raise ValueError("y_columns: j_delta, j_Delta, j_sum and phase_coef need one entry per column")
            self.y_n = len(pc)
            if self.y_n > _hip.MADI_MAX_REDUCE_COLS:
                raise ValueError(f"{self.y_n} columns exceed one reduce call's {_hip.MADI_MAX_REDUCE_COLS}")
            self.y_dev = tuple(torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).cuda()
                               for a, dt in ((jd, np.int32), (jD, np.int32), (js, np.int32), (pc, np.float64)))
            self.y_max_grid = int(max(np.max(jd), np.max(jD), np.max(js))) + 1 if self.y_n else 0
            spec = engine.single_column_spec()
        self.spec = spec
        self.plan = spec.to_ctypes()
        self.n_cols = int(np.prod(spec.shape))
        self.tune = _hip.Tune()
        self.tune.grid_per_cu = int(grid_per_cu)
        self.tune.tier_cap_um = float(tier_cap_um)
        self.tune.scan_rounds = int(scan_rounds)
        self.tune.classify_mode = int(classify_mode)
        self.max_tiles = int(max_tiles_per_generation)
        self.reserve = int(hbm_reserve_bytes)
        self.max_inflight = int(max_inflight)
        self.active: List[_Active] = []
        self.inflight: deque = deque()   # (gen id, event, workspace, [(entry, pinned host results)])
        self.gen = 0
        self.counters = torch.zeros(_hip.N_COUNTERS, dtype=torch.int64, device="cuda")
        self.hbm_samples: list = []
        # one captured batch launch per in-flight generation slot
        self.graphs: List[C.c_void_p] = []
        self.graph_capacity = int(graph_capacity)
        if use_graphs:
            for _ in range(self.max_inflight):
                h = C.c_void_p()
                _hip.check(self.lib.madi_batch_graph_create(C.byref(self.plan), self.graph_capacity,
                                                            C.byref(self.tune), _hip.dptr(self.counters),
                                                            C.byref(h)), "madi_batch_graph_create")
                self.graphs.append(h)
        self.graph_launches = 0

    def close(self) -> None:
        """Release the captured graphs (waits for their last replays)."""
        for h in self.graphs:
            self.lib.madi_batch_graph_destroy(h)
        self.graphs = []

    def __del__(self):
        if sys.is_finalizing():
            return  # the HIP runtime may already be gone
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ admission
    def entry_bytes(self, prep: PreparedEntry) -> int:
        tiles = -(-prep.req.n_walkers // _hip.TILE_WALKERS)
        a = prep.arrays
        geom = (a.nb_vec.nbytes + a.nb_ids.nbytes * 2 + a.nb_offsets.nbytes + a.cell_xyz.nbytes * 2
                + a.hint.nbytes + a.cell_annulus.nbytes * 3)
        y = 0
        if self.y_n is not None:
            # Y, and the reduce's grid-major copy of it
            y = 2 * prep.req.n_walkers * prep.walk_config().n_grid * 24 + self.y_n * 16
        return tiles * self.n_cols * 16 + self.moments_bytes(prep.req.n_walkers) + geom + y + 4096

    def moments_bytes(self, n_walkers: int) -> int:
        b = C.c_int64()
        _hip.check(self.lib.madi_moments_bytes(int(n_walkers), C.byref(self.plan), C.byref(b)), "madi_moments_bytes")
        return int(b.value)

    def can_admit(self, prep: PreparedEntry) -> bool:
        free, _ = self.torch.cuda.mem_get_info()
        return free - self.entry_bytes(prep) > self.reserve

    def admit(self, prep: PreparedEntry) -> bool:
        """Upload an entry if free HBM allows; returns False otherwise."""
        if not self.can_admit(prep):
            return False
        torch = self.torch
        req = prep.req
        cfg = prep.walk_config()
        geom = DeviceGeometry(prep.arrays)
        params = engine.walk_params(req.n_walkers, cfg, prep.pp, req.walk_seed)
        tiles = -(-req.n_walkers // _hip.TILE_WALKERS)
        partial = torch.empty(tiles * self.n_cols * 2, dtype=torch.float64, device="cuda")
        # per-walker moments of the streamed variant (unused by the tile variants)
        moments = torch.empty(self.moments_bytes(req.n_walkers), dtype=torch_uint8(torch), device="cuda")
        occ = torch.empty(cfg.n_grid, dtype=torch.int64, device="cuda")
        flags = torch.empty(_hip.N_FLAGS, dtype=torch.int64, device="cuda")
        s = _hip.stream_ptr()
        _hip.check(self.lib.madi_fill_i64(_hip.dptr(occ), cfg.n_grid, 0, s), "madi_fill_i64")
        _hip.check(self.lib.madi_fill_i64(_hip.dptr(flags), _hip.N_FLAGS, 0, s), "madi_fill_i64")
        Y = None
        if self.y_n is not None:
            if cfg.n_grid < self.y_max_grid:
                raise ValueError(f"entry {req.key}: {cfg.n_grid} Y samples, the columns need {self.y_max_grid}")
            Y = torch.zeros((req.n_walkers, cfg.n_grid, 3), dtype=torch.float64, device="cuda")
        self.active.append(_Active(prep, geom, params, tiles, partial, occ, flags, moments, admitted=time.perf_counter(),
                                   device_bytes=self.entry_bytes(prep), Y=Y))
        return True

    # ---------------------------------------------------------- generations
    def pending_tiles(self) -> int:
        return sum(a.n_tiles - a.next_tile for a in self.active)

    def launch_generation(self) -> bool:
        """Launch the pending tiles of active entries (FIFO, up to the budget)."""
        if len(self.inflight) >= self.max_inflight:
            return False
        items, owners = [], []
        budget = self.max_tiles
        for a in self.active:
            if budget <= 0:
                break
            left = a.n_tiles - a.next_tile
            if left <= 0:
                continue
            take = min(left, budget)
            w = _hip.Work()
            w.geom = a.geom.handle.value
            w.walk = a.params
            w.tile_lo = a.next_tile
            w.n_tiles = take
            w.d_partial = _hip.dptr(a.partial)
            w.d_occupancy = _hip.dptr(a.occ)
            w.d_flags = _hip.dptr(a.flags)
            w.d_moments = _hip.dptr(a.moments)
            w.d_Y = _hip.dptr(a.Y) if a.Y is not None else None
            items.append(w)
            owners.append(a)
            a.next_tile += take
            budget -= take
        if not items:
            return False
        arr = (_hip.Work * len(items))(*items)
        s = _hip.stream_ptr()
        if self.graphs and len(items) <= self.graph_capacity:
            # generation g replays graph g mod max_inflight; its previous
            # replay (generation g - max_inflight) has completed, since at
            # most max_inflight generations are in flight
            ws = None
            g = self.graphs[self.gen % len(self.graphs)]
            _hip.check(self.lib.madi_batch_graph_launch(g, arr, len(items), s), "madi_batch_graph_launch")
            self.graph_launches += 1
        else:
            need = C.c_int64()
            _hip.check(self.lib.madi_batch_workspace(len(items), C.byref(need)), "madi_batch_workspace")
            ws = self.torch.empty(int(need.value), dtype=torch_uint8(self.torch), device="cuda")
            rc = self.lib.madi_batch_launch(arr, len(items), C.byref(self.plan), C.byref(self.tune), _hip.dptr(ws),
                                            int(need.value), _hip.dptr(self.counters), s)
            _hip.check(rc, "madi_batch_launch")
        # entries whose last tiles are in this generation: their reduction and
        # result copies follow it on the same stream
        staged = [self._stage(a, s) for a in owners if a.next_tile >= a.n_tiles]
        ev = self.torch.cuda.Event()
        ev.record()
        self.inflight.append((self.gen, ev, ws, staged))
        self.gen += 1
        return True

    def _stage(self, a: _Active, s: int) -> tuple:
        """Queue, behind the generation that walks the entry's last tiles and on
        its stream, the sum of the tile slots and the copies of the sums,
        occupancy and flags into pinned host memory.  poll() only reads host
        memory once the generation's event has completed: no result waits on a
        second stream, which the runtime may map onto the hardware queue of the
        launch stream (and so behind the next generation).

        The sum's scratch and output come from torch's allocator on the launch
        stream, whose blocks are reused only by later work on that stream; the
        geometry is released by poll() after the generation has completed (the
        stream-ordered allocator may hand a freed block to a geometry uploaded
        on another stream before the launch stream reaches the free)."""
        torch = self.torch
        if a.Y is not None:
            return self._stage_y(a, s)
        n, g = self.n_cols, a.occ.numel()
        need = C.c_int64()
        _hip.check(self.lib.madi_tiles_finalize_workspace(C.byref(self.plan), a.n_tiles, C.byref(need)),
                   "madi_tiles_finalize_workspace")
        sums = torch.empty(2 * n + -(-int(need.value) // 8), dtype=torch.float64, device="cuda")
        base = _hip.dptr(sums)
        rc = self.lib.madi_tiles_finalize_ws(C.byref(self.plan), _hip.dptr(a.partial), a.n_tiles, base, base + 8 * n,
                                             base + 16 * n, int(need.value), s)
        _hip.check(rc, "madi_tiles_finalize_ws")
        host = torch.empty(2 * n + g + _hip.N_FLAGS, dtype=torch.int64, pin_memory=True)
        host[:2 * n].view(torch.float64).copy_(sums[:2 * n], non_blocking=True)
        host[2 * n:2 * n + g].copy_(a.occ, non_blocking=True)
        host[2 * n + g:].copy_(a.flags, non_blocking=True)
        return a, host

    def _stage_y(self, a: _Active, s: int) -> tuple:
        """Y mode's _stage: the column reduce of the entry's Y, then the same
        host copies, all on the launch stream."""
        torch = self.torch
        n, g = self.y_n, a.occ.numel()
        nw, G = a.Y.shape[0], a.Y.shape[1]
        sums = torch.empty(2 * n, dtype=torch.float64, device="cuda")
        base = _hip.dptr(sums)
        jd, jD, js, pc = (_hip.dptr(t) for t in self.y_dev)
        rc = self.lib.madi_reduce_f64(_hip.dptr(a.Y), nw, G, jd, jD, js, pc, n, base, base + 8 * n, s)
        _hip.check(rc, "madi_reduce_f64")
        host = torch.empty(2 * n + g + _hip.N_FLAGS, dtype=torch.int64, pin_memory=True)
        host[:2 * n].view(torch.float64).copy_(sums, non_blocking=True)
        host[2 * n:2 * n + g].copy_(a.occ, non_blocking=True)
        host[2 * n + g:].copy_(a.flags, non_blocking=True)
        return a, host

    def _result(self, a: _Active, host) -> EntryResult:
        """The entry's results from its staged host copy (its generation has completed)."""
        n, g = (self.y_n if a.Y is not None else self.n_cols), a.occ.numel()
        h = host.numpy()
        shape = (n,) if a.Y is not None else self.spec.shape
        sums = h[:2 * n].view(np.float64)
        cos, sin = sums[:n].reshape(shape).copy(), sums[n:].reshape(shape).copy()
        occ, flags = h[2 * n:2 * n + g].copy(), h[2 * n + g:]
        req = a.prep.req
        n_esc, n_err = int(flags[0]), int(flags[1])
        return EntryResult(
            key=req.key, cos=cos, sin=sin, occupancy=occ, n_walkers=req.n_walkers, n_escaped=n_esc,
            n_error=n_err, pp=a.prep.pp, submitted=req.submitted, admitted=a.admitted,
            finished=time.perf_counter(), prep_s=a.prep.prep_s, error=entry_error(n_esc, n_err, req.n_walkers),
        )

    def poll(self, block: bool = False) -> List[EntryResult]:
        """Results of entries whose last generation has completed."""
        out, finished = [], set()
        while self.inflight:
            gen, ev, ws, staged = self.inflight[0]
            if block:
                ev.synchronize()
            elif not ev.query():
                break
            self.inflight.popleft()
            block = False
            for a, host in staged:
                # every kernel reading the geometry has completed; the free is
                # ordered behind whatever the launch stream runs now
                a.geom.release(_hip.stream_ptr())
                out.append(self._result(a, host))
                finished.add(id(a))
        if not out:
            return []
        self.active = [a for a in self.active if id(a) not in finished]
        if out:
            free, total = self.torch.cuda.mem_get_info()
            self.hbm_samples.append((time.perf_counter(), total - free))
        return out

    def drain(self) -> List[EntryResult]:
        out = []
        while self.active or self.inflight:
            self.launch_generation()
            out += self.poll(block=True)
        return out

    def counter_dict(self) -> dict:
        return dict(zip(_hip.COUNTER_NAMES, self.counters.cpu().numpy().tolist()))


def torch_uint8(torch):
    return torch.uint8


def entry_error(n_escaped: int, n_error: int, n_walkers: int) -> Optional[str]:
    """The reference's fatal outcome of a walk, or None: a classifier failure
    (madi/walker_gpu.py:938-940) or walkers leaving the simulation domain
    (madi/walker_gpu.py:1023-1028).  A failed entry has no valid signal."""
    if n_error:
        return f"GPU full-facet classification failed for {n_error}/{n_walkers} walkers"
    if n_escaped:
        return f"SI §S.III fatal boundary condition: {n_escaped}/{n_walkers} walkers exited Ω_sim."
    return None


def failed_entry(req: EntryRequest, error: str, shape, admitted: float = 0.0) -> EntryResult:
    """Result of an entry that failed before or during its walk (no signal)."""
    now = time.perf_counter()
    return EntryResult(key=req.key, cos=np.zeros(shape), sin=np.zeros(shape), occupancy=np.zeros(0, np.int64),
                       n_walkers=req.n_walkers, n_escaped=0, n_error=0, pp=float("nan"), submitted=req.submitted,
                       admitted=admitted or now, finished=now, prep_s=0.0, error=error)


def split_failed(results: List[EntryResult], failed: Optional[list]) -> List[EntryResult]:
    """Good results; failed ones go to `failed`, or raise RuntimeError (after
    every entry has finished) when the caller gave no list for them."""
    good = [r for r in results if not r.failed]
    bad = [r for r in results if r.failed]
    if failed is not None:
        failed.extend(bad)
    elif bad:
        raise RuntimeError("; ".join(f"entry {r.key}: {r.error}" for r in bad))
    return good


def run_batched(preps: Iterable[PreparedEntry], spec: engine.PhaseSpec, failed: Optional[list] = None,
                **kw) -> List[EntryResult]:
    """Admit every prepared entry as it arrives and run generations until all finish.

    An entry whose walk failed (escape or classifier failure) fails alone: it
    is left out of the returned results and appended to `failed` (without a
    list, RuntimeError once every entry has finished); the other entries'
    results do not depend on it."""
    sch = DeviceScheduler(spec, **kw)
    waiting: deque = deque()
    results: List[EntryResult] = []
    for prep in preps:
        waiting.append(prep)
        while waiting and sch.admit(waiting[0]):
            waiting.popleft()
        if sch.pending_tiles() >= sch.max_tiles or len(sch.inflight) == 0:
            sch.launch_generation()
        results += sch.poll()
    while waiting or sch.active or sch.inflight:
        while waiting and sch.admit(waiting[0]):
            waiting.popleft()
        if not sch.launch_generation() and not sch.inflight and waiting:
            sch.close()
            raise MemoryError("an entry does not fit in free HBM even with the device idle")
        results += sch.poll(block=True)
    sch.close()
    return split_failed(results, failed)


def run_plain(requests: Iterable[EntryRequest], spec: engine.PhaseSpec,
              prepare: Callable[[EntryRequest], PreparedEntry] = prepare_entry, classify_mode: int = 0,
              scan_rounds: int = 0, failed: Optional[list] = None) -> List[EntryResult]:
    """One entry per launch: prepare, upload, launch, synchronise, in sequence
    (failed entries as in run_batched)."""
    results = []
    for req in requests:
        t_admit = time.perf_counter()
        try:
            prep = prepare(req)
        except RuntimeError as e:  # a geometry with no defined labels fails its entry alone
            results.append(failed_entry(req, str(e), spec.shape, t_admit))
            continue
        geom = DeviceGeometry(prep.arrays)
        params = engine.walk_params(req.n_walkers, req.config(), prep.pp, req.walk_seed, scan_rounds=scan_rounds,
                                    classify_mode=classify_mode)
        res = engine.walk_fused(geom, params, spec, precision=32)
        results.append(EntryResult(
            key=req.key, cos=res.cos, sin=res.sin, occupancy=res.occupancy, n_walkers=req.n_walkers,
            n_escaped=res.n_escaped, n_error=res.n_error, pp=prep.pp, submitted=req.submitted, admitted=t_admit,
            finished=time.perf_counter(), prep_s=prep.prep_s,
            error=entry_error(res.n_escaped, res.n_error, req.n_walkers),
        ))
        del geom
    return split_failed(results, failed)


def library_requests(n: int, seed: int = 2, walkers=(200_000, 1_000_000), cells=(50, 4000), kio=(1.0, 130.0),
                     V=(0.1, 20.0), vi=(0.5, 0.95), ts: float = 0.02, D0: float = 2.0,
                     T_max_ms: float = 46.0) -> List[EntryRequest]:
    """Uneven library entries: k_io and V log-uniform, v_i uniform, cells and walkers uniform."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        k = float(np.exp(rng.uniform(np.log(kio[0]), np.log(kio[1]))))
        v = float(np.exp(rng.uniform(np.log(V[0]), np.log(V[1]))))
        f = float(rng.uniform(vi[0], vi[1]))
        rho = f / (v * 1e-6)
        c = int(rng.integers(cells[0], cells[1] + 1))
        w = int(rng.integers(walkers[0], walkers[1] + 1))
        out.append(EntryRequest(key=(i, round(k, 4), round(rho, 1), round(v, 6)), kio=k, rho=rho, V=v, n_cells=c,
                                n_walkers=w, seed=1000 + i, walk_seed=104_729 * (i + 1) & 0x7FFFFFFF, ts=ts, D0=D0,
                                T_max_ms=T_max_ms))
    return out


def sweep_requests(n_V: int = 77, vi_values=(0.5, 0.747, 0.994), rho_range=(4.4e3, 5.8e7), V_range=(0.01, 206.0),
                   n_kio: int = 65, kio_max: float = 130.0, walkers: int = 500_000, cells=(32, 8000),
                   seed: int = 4, D0: float = 2.0, T_max_ms: float = 46.0) -> List[EntryRequest]:
    """Full-range library sweep: V log-spaced over V_range, v_i from vi_values
    (rho = v_i / V, kept inside rho_range), k_io evenly spaced over [0, kio_max]
    (k_io = 0 gives p_p = 0), one geometry per (rho, V) group with a cell count
    log-uniform over `cells`, dt = 0.01 ms below V = 0.5 pL and 0.02 ms above."""
    rng = np.random.default_rng(seed)
    out = []
    kios = np.linspace(0.0, kio_max, n_kio)
    group = 0
    for v in np.exp(np.linspace(np.log(V_range[0]), np.log(V_range[1]), n_V)):
        for f in vi_values:
            rho = f / (v * 1e-6)
            if not (rho_range[0] <= rho <= rho_range[1]):
                continue
            c = int(round(np.exp(rng.uniform(np.log(cells[0]), np.log(cells[1])))))
            ts = 0.01 if v < 0.5 else 0.02
            for k in kios:
                i = len(out)
                out.append(EntryRequest(key=(i, round(float(k), 4), round(rho, 1), round(float(v), 6)), kio=float(k),
                                        rho=float(rho), V=float(v), n_cells=c, n_walkers=walkers,
                                        seed=50_000 + group, walk_seed=104_729 * (i + 1) & 0x7FFFFFFF, ts=ts,
                                        D0=D0, T_max_ms=T_max_ms))
            group += 1
    return out


def acquisition_spec(n_dirs: int = 24, delta: float = 6.0, Deltas=(15.0, 25.0, 30.0, 40.0),
                     b_values=(0.0, 1000.0, 2500.0, 4000.0, 6000.0), h_ms: float = 1.0,
                     dir_seed: int = 1) -> engine.PhaseSpec:
    """PGSE acquisition: one delta, several Delta, b values and n_dirs directions uniform on the sphere."""
    if n_dirs == 3:
        dirs = np.eye(3)
    else:
        dirs = np.random.default_rng(dir_seed).normal(size=(n_dirs, 3))
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    pc = np.array([[engine.phase_coef(b, delta, D) for b in b_values] for D in Deltas])
    j = lambda t: int(round(t / h_ms))  # noqa: E731
    return engine.PhaseSpec(j_delta=np.array([j(delta)] * len(Deltas)), j_Delta=np.array([j(D) for D in Deltas]),
                            j_sum=np.array([j(D + delta) for D in Deltas]), pc=pc, dirs=dirs)
