"""Continuous-batching server for library-entry simulation requests.

Requests (k_io, rho, V with their own cell and walker counts) arrive at any
time.  Host geometry is prepared in worker processes; a prepared request is
admitted to the GPU when free HBM allows (``DeviceScheduler.admit``), and its
tiles join the next generation of the persistent kernel alongside whatever
else is running.  Finished entries are appended to a library stream file
(one JSON line per entry) and can be written as a v5 .npz library.
Latency of an entry = finish time - submit time.
"""

from __future__ import annotations

import heapq
import json
import multiprocessing as mp
import time
from collections import deque
from pathlib import Path
from typing import Callable, List, Optional

import numpy as np

from .batch import DeviceScheduler, EntryRequest, EntryResult, PreparedEntry, failed_entry, prepare_entry


def _worker_init() -> None:
    # import the geometry code (numpy and the library's native host
    # geometry, which makes no device call) when the worker starts, not
    # inside the first request it prepares
    from . import _hip, batch, ensemble  # noqa: F401
    _hip.load_library()


def make_pool(workers: int):
    """Spawned worker pool for host geometry; create it before the GPU is touched."""
    ctx = mp.get_context("spawn")
    return ctx.Pool(processes=max(1, int(workers)), initializer=_worker_init)


class LibraryStream:
    """Append-only record of finished entries (one JSON object per line)."""

    def __init__(self, path: Optional[str]):
        self.path = Path(path) if path else None
        self.fh = open(self.path, "a") if self.path else None
        self.count = 0

    def append(self, res: EntryResult, spec_shape) -> None:
        self.count += 1
        if self.fh is None:
            return
        rec = {
            "key": list(res.key), "n_walkers": res.n_walkers, "pp": res.pp,
            "signal": (res.cos / res.n_walkers).ravel().tolist(),
            "signal_imag": (res.sin / res.n_walkers).ravel().tolist(), "shape": list(spec_shape),
            "occupancy_fraction": (res.occupancy / res.n_walkers).tolist(),
            "n_escaped": res.n_escaped, "n_error": res.n_error, "latency_s": res.latency_s,
        }
        self.fh.write(json.dumps(rec) + "\n")
        self.fh.flush()

    def close(self):
        if self.fh:
            self.fh.close()


class Server:
    """Submit requests; ``step()`` advances preparation, admission, launches and finalisation."""

    def __init__(self, scheduler: DeviceScheduler, pool=None, stream: Optional[LibraryStream] = None,
                 max_prep_inflight: int = 32, prepare: Callable[[EntryRequest], PreparedEntry] = prepare_entry):
        self.sch = scheduler
        self.pool = pool
        # without a pool, requests are prepared in the calling process by
        # `prepare` (for example a lookup of geometry built beforehand)
        self.prepare = prepare
        self.stream = stream or LibraryStream(None)
        self.queued: list = []            # heap of (cells, seq, request) not yet handed to the pool
        self.preparing: deque = deque()   # (request, AsyncResult) or PreparedEntry
        self.ready: deque = deque()       # prepared, waiting for HBM
        self.results: List[EntryResult] = []
        self.failed: List[EntryResult] = []  # entries whose walk failed: not in results, not in the stream
        self.max_prep = max(1, int(max_prep_inflight))
        self.rejected = 0
        self.seq = 0

    def submit(self, req: EntryRequest) -> None:
        req.submitted = time.perf_counter()
        if self.pool is None:
            try:
                self.preparing.append(self.prepare(req))
            except RuntimeError as e:
                self.failed.append(failed_entry(req, str(e), self.sch.spec.shape))
            return
        heapq.heappush(self.queued, (req.n_cells, self.seq, req))
        self.seq += 1
        self._dispatch()

    def _dispatch(self) -> None:
        """Hand queued requests to the pool, fewest cells first (the cheapest
        geometry reaches the GPU soonest), keeping at most max_prep in the
        pool so later, cheaper arrivals can still go ahead."""
        while self.queued and len(self.preparing) < self.max_prep:
            _, _, req = heapq.heappop(self.queued)
            self.preparing.append((req, self.pool.apply_async(prepare_entry, (req,))))

    def _collect_prepared(self) -> None:
        """Move every finished preparation to `ready` (in the order they finish:
        a slow geometry does not hold back the ones submitted after it)."""
        pending: deque = deque()
        while self.preparing:
            head = self.preparing.popleft()
            if isinstance(head, PreparedEntry):
                self.ready.append(head)
                continue
            req, job = head
            if not job.ready():
                pending.append(head)
                continue
            try:
                self.ready.append(job.get())
            except RuntimeError as e:  # the geometry has no defined labels: this entry fails alone
                self.failed.append(failed_entry(req, str(e), self.sch.spec.shape))
        self.preparing = pending
        if self.pool is not None:
            self._dispatch()

    def step(self, block: bool = False) -> List[EntryResult]:
        self._collect_prepared()
        while self.ready and self.sch.admit(self.ready[0]):
            self.ready.popleft()
        self.sch.launch_generation()
        done = self.sch.poll(block=block)
        good = []
        for r in done:
            if r.failed:
                # the reference aborts such an entry (fatal escape / classifier
                # failure); here it fails alone and is never written
                self.failed.append(r)
                continue
            self.stream.append(r, self.sch.spec.shape)
            good.append(r)
        self.results += good
        return done

    def idle(self) -> bool:
        return not (self.queued or self.preparing or self.ready or self.sch.active or self.sch.inflight)

    def drain(self, progress: Optional[Callable[[], None]] = None, every_s: float = 30.0) -> None:
        """Run until every submitted request has finished; `progress` is called
        every `every_s` seconds meanwhile."""
        last = time.perf_counter()
        while not self.idle():
            if progress is not None and time.perf_counter() - last >= every_s:
                progress()
                last = time.perf_counter()
            self._collect_prepared()
            if not self.preparing and self.ready and not self.sch.active and not self.sch.inflight:
                if not self.sch.admit(self.ready[0]):
                    raise MemoryError("a request does not fit in free HBM even with the device idle")
                self.ready.popleft()
            # while geometries are still being prepared, keep admitting and
            # launching as they arrive instead of blocking on a generation
            preparing = bool(self.preparing or self.queued)
            done = self.step(block=bool(self.sch.inflight) and not preparing)
            if preparing and not done:
                time.sleep(0.0005)


def poisson_arrivals(rate_hz: float, duration_s: float, seed: int = 3) -> np.ndarray:
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate_hz)
        if t >= duration_s:
            return np.asarray(out)
        out.append(t)


def serve_poisson(server: Server, requests: List[EntryRequest], arrivals: np.ndarray,
                  progress=None) -> dict:
    """Feed requests at the given arrival offsets (seconds) and run until all finish."""
    t0 = time.perf_counter()
    k = 0
    hbm = []
    last_report = t0
    while k < len(requests) or not server.idle():
        now = time.perf_counter() - t0
        while k < len(requests) and arrivals[k] <= now:
            server.submit(requests[k])
            k += 1
        server.step(block=False)
        if server.sch.inflight and k < len(requests):
            time.sleep(0.001)
        elif not server.sch.inflight and k < len(requests):
            time.sleep(min(0.005, max(0.0, arrivals[k] - (time.perf_counter() - t0))))
        elif server.sch.inflight:
            server.step(block=True)
        else:
            time.sleep(0.002)
        if progress and time.perf_counter() - last_report > 10.0:
            last_report = time.perf_counter()
            progress(f"t={now:.0f}s submitted={k} done={len(server.results)} active={len(server.sch.active)} "
                     f"ready={len(server.ready)} preparing={len(server.preparing)}")
            free, total = server.sch.torch.cuda.mem_get_info()
            hbm.append((now, total - free))
    wall = time.perf_counter() - t0
    lat = np.array([r.latency_s for r in server.results])
    return {
        "requests": len(requests), "finished": len(server.results), "failed": len(server.failed), "wall_s": wall,
        "entries_per_hour": len(server.results) / wall * 3600.0,
        "latency_s_p50": float(np.percentile(lat, 50)) if len(lat) else None,
        "latency_s_p99": float(np.percentile(lat, 99)) if len(lat) else None,
        "latency_s_max": float(lat.max()) if len(lat) else None,
        "hbm_used_bytes_samples": hbm,
    }
