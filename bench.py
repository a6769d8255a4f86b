"""Benchmark of the MADI walk on one MI355X.

Default run: one JSON line whose headline is the production single entry
(C1): a periodic cube of 1,000 contracted Voronoi cells, rho = 5e5 cells/µL,
V = 1.2 pL (v_i = 0.6), k_io = 22 s^-1 (p_p by Eq. 5), D = 2.0 µm²/ms,
dt = 0.02 ms for 2,300 steps (46 ms), 1e6 walkers, PGSE delta = 6 ms,
Delta = 15/25/30/40 ms, b = 0/1000/2500/4000/6000 s/mm², 24 directions:
480 signals.  The FP32 production walk with the fused phase reduction is the
timed path; a "step" is one full entry (1e6 walkers x 2,300 steps), inputs
resident on the device.  The same line carries:

* ``c2``: 256 uneven library entries, the batched/served form against the
  plain one-entry-per-launch form (all 256 entries each), entries/h, their
  ratio, the bitwise count and the batched entries' latency;
* ``c3``: 60 s of Poisson arrivals at 20 requests/s served with continuous
  batching, p50/p99 latency, entries/h and HBM in use;
* ``fp64_check``: the FP32 entry against the FP64 exact-classifier walk on the
  same seeds, per (Delta, b), with each signal's Monte Carlo standard error;
* ``cpu_baseline``: the reference's CPU walk algorithm (numpy/scipy
  restatement) on the host cores, single process and across the worker
  processes, and the FP64 C oracle with OpenMP.

* ``universal_reduce``: the universal library's FP64 column reduce at
  production size (50,000 walkers x 129 Y samples x 31,125 columns);
* ``production``: the reference's own published benchmark (walk plus
  31,125-column reduction of one Eq. S8 ensemble, 50,000 walkers x 128,000
  steps, k_io 20 /s, at the production grid's low / centre / high rho)
  through the drop-in API: the FP64 exact_cached walk (graph labels) and the
  FP32 walk at each coordinate, the FP64 exact walk at low rho (bitwise the
  exact_cached one), the FP64 walk's seed-to-seed |dS| beside FP32's, and
  the library's 51-k_io sweep of the ensemble in one launch, each beside the
  published A100 time.

Other workloads (--workload): c2, c3 on their own, c4 (a seeded sample of the
full-range sweep), map (MAP matching of 1e6 voxels against a 14,105-entry
library, FP64), reduce (the column reduce alone, with its CPU baseline),
library (the drop-in library build at the default configuration, group by
group against the batched scheduler), production (the production object
alone, with the exact walk at all three coordinates; --prod-coords and
--prod-modes narrow it), group (one production task of the reference's
library build: a (rho, V) group of 40 ensembles x 51 k_io through the drop-in
API, against the published A100 task time).  --gpus N > 1 exits: this
benchmark uses one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# Algorithmic FP32 operation counts (FMA = 2, other ops = 1) per unit of work:
# per walker-step (3 Box-Muller normals, step, escape test, safe-ball test,
# accept/reject, compensated Y), per facet (edge) evaluation, per walker
# (480-column phase stage).  The facet count comes from the kernel's counters.
F_STEP = 79
F_FACET = 13
F_WALKER = 2400
PEAK_FP32_TFLOPS = 157.3        # MI355X FP32 vector peak (packed FMA)
PEAK_FP32_NONPACKED_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def cpu_baseline_c_oracle(ens, cfg, pp, threads, target_s=10.0, max_walkers=1_000_000):
    """The FP64 C oracle (the reference walk's semantics, exact classifier) with
    OpenMP on host cores; a 2,048-walker calibration run sizes the timed
    sample to about target_s seconds.  Returns (walker-steps/s, s, walkers)."""
    from oracle import oracle as O
    from custom_madi_amd.ensemble import geometry_arrays
    O.set_threads(threads)
    og = O.Geometry.from_arrays(geometry_arrays(ens))

    def run(n):
        t0 = time.perf_counter()
        O.walk(og, n, cfg.n_steps, cfg.steps_per_h, cfg.n_grid, cfg.sigma, cfg.ts, pp, ens.source_lo,
               ens.source_hi, seed=0)
        return time.perf_counter() - t0

    dt = run(2048)
    n = int(min(max_walkers, max(2048, target_s / dt * 2048)))
    dt = run(n)
    return n * cfg.n_steps / dt, dt, n


def _reference_cpu_walk(job):
    """One process's share of the reference CPU walk algorithm: the numpy /
    scipy restatement of _walk_cpu + classify_cpu (madi/walker_gpu.py:698-840,
    madi/ensemble.py:213-256) over the image-padded periodic seed set."""
    from oracle import oracle as O
    kd_xyz, kd_ann, kd_label, L, n, steps, sph, n_grid, sigma, ts, pp, seed = job
    t0 = time.perf_counter()
    O.walk_cpu_reference(kd_xyz, kd_ann, L, 0.0, L, n, steps, sph, n_grid, sigma, ts, pp, seed=seed,
                         label=kd_label)
    return time.perf_counter() - t0


def cpu_baseline_reference(ens, cfg, pp, pool, processes, target_s, log):
    """The reference's CPU walk algorithm timed single-process and split by
    walker ranges over `processes` worker processes (each its own seed).
    A 64-walker x 230-step calibration sizes each process's walker count for
    about target_s seconds of the full 2,300-step walk."""
    from custom_madi_amd.ensemble import geometry_arrays
    arr = geometry_arrays(ens)
    base = (arr.kd_xyz, arr.kd_annulus, arr.kd_label, ens.L)
    tail = (cfg.steps_per_h, cfg.n_grid, cfg.sigma, cfg.ts, pp)
    cal = _reference_cpu_walk((*base, 64, 230, *tail, 1))
    n = int(max(16, min(20_000, target_s / (cal * 10.0) * 64)))
    log(f"reference CPU walk: {n} walkers x {cfg.n_steps} steps per process (calibration {cal:.2f} s)")
    dt1 = _reference_cpu_walk((*base, n, cfg.n_steps, *tail, 2))
    single = {"value": n * cfg.n_steps / dt1, "cores": 1, "walkers": n, "wall_s": dt1}
    t0 = time.perf_counter()
    pool.map(_reference_cpu_walk, [(*base, n, cfg.n_steps, *tail, 100 + k) for k in range(processes)], chunksize=1)
    dtn = time.perf_counter() - t0
    multi = {"value": processes * n * cfg.n_steps / dtn, "cores": processes, "walkers": processes * n,
             "wall_s": dtn}
    return single, multi


def visible_cpus() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _noop(x):
    return x


def _tuning(args, spec, log, sample=None, tag=""):
    """Tuner knobs for the batched runs: explicit flags, else the autotuner's
    choice, timed on `sample` (requests of the workload itself) when given."""
    if args.classify_mode or args.scan_rounds or args.grid_per_cu or args.no_tune:
        return {"classify_mode": args.classify_mode, "scan_rounds": args.scan_rounds,
                "grid_per_cu": args.grid_per_cu}, "flags"
    from custom_madi_amd import batch, tune
    t0 = time.perf_counter()
    preps = [batch.prepare_entry(r) for r in sample] if sample else None
    choice = tune.autotune(spec, preps=preps, log=log, tag=tag)
    log(f"tuned in {time.perf_counter() - t0:.1f} s: {choice}")
    src = f"autotuner ({tune.cache_path()}" + (f", sample of 6 {tag} entries)" if tag else ")")
    return choice.kwargs(), src


def measure_c2(args, pool, log):
    """256 uneven library entries: plain one-entry-per-launch form vs the batched form.

    Both forms run every entry.  The plain form restates the reference's
    library loop (madi/library.py:546-629): host geometry, upload, one
    launch, synchronise, per entry.  The batched form prepares geometry in
    the worker pool and streams entries through DeviceScheduler."""
    from custom_madi_amd import _hip, batch, serve, workloads
    reqs = workloads.c2_requests(args.entries)
    spec = workloads.acquisition()
    _hip.require_gpu()
    info = _hip.device_info()
    walker_steps = sum(r.n_walkers * int(round(r.T_max_ms / r.ts)) for r in reqs)
    log(f"{len(reqs)} entries, {walker_steps:.3e} walker-steps, device {info['name']} CUs={info['cu_count']}")

    # warm-up (kernel load, jump tables)
    warm = batch.library_requests(2, seed=99, walkers=(20_000, 20_000), cells=(60, 80))
    batch.run_plain(warm, spec)

    knobs, knob_src = _tuning(args, spec, log, sample=reqs[:: max(1, len(reqs) // 6)][:6], tag="c2")
    walk_knobs = {k: knobs[k] for k in ("classify_mode", "scan_rounds")}

    # batched / served form: all requests submitted at t0
    sch = batch.DeviceScheduler(spec, tier_cap_um=args.tier_cap, max_tiles_per_generation=args.gen_tiles, **knobs)
    srv = serve.Server(sch, pool=pool)
    t0 = time.perf_counter()
    for r in reqs:
        srv.submit(r)
    srv.drain()
    t_batch = time.perf_counter() - t0
    sch.close()
    res_b = {r.key: r for r in srv.results}
    log(f"batched: {t_batch:.1f} s, generations {sch.gen}")

    # plain form on the first n_plain entries (sequential host prep, one launch each)
    n_plain = min(args.plain_entries, len(reqs))
    t0 = time.perf_counter()
    res_p = batch.run_plain(reqs[:n_plain], spec, **walk_knobs)
    t_plain = time.perf_counter() - t0
    log(f"plain: {n_plain} entries in {t_plain:.1f} s")

    # plain form with host geometry prepared beforehand (device-only comparison)
    preps = pool.map(batch.prepare_entry, reqs[:n_plain])
    t0 = time.perf_counter()
    batch.run_plain([p.req for p in preps], spec, prepare=_prepared_lookup(preps), **walk_knobs)
    t_plain_dev = time.perf_counter() - t0

    same = sum(int(np.array_equal(res_b[r.key].cos, r.cos) and np.array_equal(res_b[r.key].sin, r.sin)
                   and np.array_equal(res_b[r.key].occupancy, r.occupancy)) for r in res_p)
    ws_plain = sum(r.n_walkers * int(round(r.T_max_ms / r.ts)) for r in reqs[:n_plain])
    eph_b = len(reqs) / t_batch * 3600.0
    eph_p = n_plain / t_plain * 3600.0
    eph_pd = n_plain / t_plain_dev * 3600.0
    lat = [r.latency_s for r in srv.results]
    esc = len(srv.failed)  # failed entries (escape or classifier failure)
    out = {
        "workload": "c2: 256 uneven entries (k_io log-U 1-130/s, V log-U 0.1-20 pL, v_i U 0.5-0.95, 50-4000 cells, "
                    "2e5-1e6 walkers, dt 0.02 ms x 2300 steps, 480 signals), FP32 walk, FP64 sums",
        "batched": {"entries": len(reqs), "wall_s": t_batch, "entries_per_hour": eph_b,
                    "walker_steps_per_s": walker_steps / t_batch, "generations": sch.gen,
                    "graph_launches": sch.graph_launches,
                    "latency_s_p50": float(np.percentile(lat, 50)), "latency_s_p99": float(np.percentile(lat, 99)),
                    "host_workers": args.workers, "escapes_or_errors": esc},
        "plain": {"entries": n_plain, "wall_s": t_plain, "entries_per_hour": eph_p,
                  "walker_steps_per_s": ws_plain / t_plain,
                  "host_prep_ms_mean": 1e3 * float(np.mean([r.prep_s for r in res_p])),
                  "host_prep_ms_max": 1e3 * float(np.max([r.prep_s for r in res_p]))},
        "plain_prebuilt_geometry": {"entries": n_plain, "wall_s": t_plain_dev, "entries_per_hour": eph_pd,
                                    "walker_steps_per_s": ws_plain / t_plain_dev},
        "batched_over_plain": eph_b / eph_p, "batched_over_plain_prebuilt": eph_b / eph_pd,
        "bitwise_equal_plain_vs_batched": f"{same}/{n_plain}",
        "tuning": {**knobs, "source": knob_src},
        "classifier": sch.counter_dict(),
    }
    return out


def run_c2(args):
    from custom_madi_amd import serve
    pool = serve.make_pool(args.workers)           # spawned before the GPU is initialised
    pool.map(_noop, range(args.workers))
    log = lambda m: print(f"[bench c2] {m}", file=sys.stderr, flush=True)  # noqa: E731
    out = measure_c2(args, pool, log)
    pool.close()
    pool.join()      # workers exit on their own (no SIGTERM at interpreter exit)
    print(json.dumps({"metric": "library entries/h", "value": out["batched"]["entries_per_hour"],
                      "unit": "entries/h", "n_gpus": 1, "higher_is_better": True, **out}), flush=True)
    return 0


PEAK_FP64_TFLOPS = 78.6


def map_workload(n_vox, n_lib, seed=4):
    """Synthetic matcher inputs: library rows of direction-averaged signals
    S = exp(-b D(Delta)) for 4 Delta x 5 b (20 columns, pair-major) with
    per-entry diffusivities, and voxels = a random entry's row + N(0, 0.02)."""
    rng = np.random.default_rng(seed)
    b = np.array([0.0, 1.0, 2.5, 4.0, 6.0])            # ms/µm² (s/mm² x 1e-3)
    Deltas = np.array([15.0, 25.0, 30.0, 40.0])
    D0 = rng.uniform(0.2, 3.0, (n_lib, 1, 1))
    r = rng.uniform(0.0, 0.6, (n_lib, 1, 1))
    Dd = D0 * (1.0 - r * Deltas[None, :, None] / 40.0)
    R = np.exp(-b[None, None, :] * Dd).reshape(n_lib, -1)
    M = R[rng.integers(0, n_lib, n_vox)] + rng.normal(0.0, 0.02, (n_vox, R.shape[1]))
    return np.ascontiguousarray(M), np.ascontiguousarray(R)


def run_map(args):
    """MAP matching of a voxel batch against a library, inputs resident in HBM."""
    import torch
    from custom_madi_amd import _hip
    from custom_madi_amd.match import DeviceLibrary
    _hip.require_gpu()
    n_vox, n_lib = args.voxels, args.library_entries
    M, R = map_workload(n_vox, n_lib)
    n_feat = R.shape[1]
    lib = DeviceLibrary(R)
    dM = torch.from_numpy(M).to("cuda")
    ws = lib.workspace(n_vox)
    outs = {fs: (torch.empty(n_vox, dtype=torch.int64, device="cuda"), torch.empty(n_vox, dtype=torch.float64,
                                                                                     device="cuda"),
                 torch.empty(n_vox, dtype=torch.float64, device="cuda") if fs else None) for fs in (False, True)}
    stream = torch.cuda.current_stream()

    def timed(fit_s0, steps, warmup):
        for _ in range(warmup):
            lib.match_device(dM, fit_s0, workspace=ws, out=outs[fit_s0])
        torch.cuda.synchronize()
        starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
        ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
        t0 = time.perf_counter()
        for k in range(steps):
            starts[k].record(stream)
            lib.match_device(dM, fit_s0, workspace=ws, out=outs[fit_s0])
            ends[k].record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return wall, [s.elapsed_time(e) for s, e in zip(starts, ends)]

    wall, kern_ms = timed(False, args.steps, args.warmup)
    wall_f, kern_f = timed(True, max(1, args.steps), 1)
    flops = 2.0 * n_vox * n_lib * n_feat
    kavg = float(np.mean(kern_ms)) * 1e-3
    achieved = flops / kavg / 1e12
    traffic = None
    pmc = ROOT / "profiles" / "map_match_traffic.json"
    if pmc.is_file():
        traffic = json.loads(pmc.read_text()).get("bytes_per_launch")
    out = {
        "metric": "voxels/s", "value": n_vox * args.steps / wall, "unit": "voxels/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"map: MAP matching (S0 fixed) of {n_vox} voxels against {n_lib} library entries "
                               f"x {n_feat} columns (4 Delta x 5 b), FP64", "voxels": n_vox,
                   "library_entries": n_lib, "columns": n_feat},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic, "kernel_ms": kavg * 1e3,
                     "algorithmic_tflop_per_launch": flops / 1e12},
        "free_s0": {"ms_per_step": 1e3 * wall_f / max(1, args.steps), "kernel_ms": float(np.mean(kern_f)),
                    "tflops": flops / (float(np.mean(kern_f)) * 1e-3) / 1e12},
    }
    if not args.no_cpu_baseline:
        from threadpoolctl import threadpool_limits
        from oracle import oracle as O
        try:
            cores = len(os.sched_getaffinity(0))
        except AttributeError:
            cores = os.cpu_count() or 1
        threads = max(1, min(16, cores))
        batch = 1000                       # the reference's own matching batch (1,000 voxels per call)
        with threadpool_limits(limits=threads):
            t0 = time.perf_counter()
            O.map_match_cpu_reference(M[:batch], R)
            one = time.perf_counter() - t0
            n_b = int(max(1, min(n_vox // batch, args.cpu_seconds / max(one, 1e-6))))
            t0 = time.perf_counter()
            for q in range(n_b):
                O.map_match_cpu_reference(M[q * batch:(q + 1) * batch], R)
            dt = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": n_b * batch / dt, "unit": "voxels/s", "cores": threads, "kind": "port",
            "sample": f"{n_b} batches of {batch} voxels against the same {n_lib}-entry library with the reference's "
                      f"CPU matcher arithmetic (numpy/BLAS, {threads} threads) in {dt:.1f} s on {cpu_model()}",
        }
    print(json.dumps(out), flush=True)
    return 0


# FP64 operations per (sample, column) of the column reduce: dM's two adds
# shared by the pair's b values are left out; phase 1 multiply; Cody-Waite
# reduction 1 multiply + 3 FMA; r^2 and r^3 factor 2 multiplies; the sin and
# cos polynomials 8 FMA each; the two sums 2 adds (FMA = 2).
F_REDUCE = 1 + (1 + 3 * 2) + 2 + 2 * 8 * 2 + 2


def measure_reduce(args, log, cpu_baseline=True):
    """The universal library's column reduce (madi_reduce_f64) at production
    size: 50,000 walkers x 129 Y samples x 31,125 columns, Y resident."""
    import torch
    from custom_madi_amd import _hip, workloads
    _hip.require_gpu()
    lib = _hip.load_library()
    if args.copy_probe:
        # a 1 GiB device copy: its FETCH_SIZE / WRITE_SIZE scale the reduce's
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    Y, cols = workloads.universal_reduce()
    n, G, _ = Y.shape
    n_cols = len(cols.phase_coef)
    dev = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).cuda()  # noqa: E731
    dY = dev(Y, np.float64)
    jd, jD, js = dev(cols.j_delta, np.int32), dev(cols.j_Delta, np.int32), dev(cols.j_sum, np.int32)
    pc = dev(cols.phase_coef, np.float64)
    cs = torch.empty(n_cols, dtype=torch.float64, device="cuda")
    sn = torch.empty_like(cs)
    stream = torch.cuda.current_stream()

    def launch():
        _hip.check(lib.madi_reduce_f64(_hip.dptr(dY), n, G, _hip.dptr(jd), _hip.dptr(jD), _hip.dptr(js),
                                       _hip.dptr(pc), n_cols, _hip.dptr(cs), _hip.dptr(sn), _hip.stream_ptr()),
                   "madi_reduce_f64")

    steps = max(5, args.steps)
    launch()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    t0 = time.perf_counter()
    for k in range(steps):
        starts[k].record(stream)
        launch()
        ends[k].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = [a.elapsed_time(b) for a, b in zip(starts, ends)]
    kavg = float(np.mean(kern_ms)) * 1e-3
    sample_cols = 3.0 * n * n_cols
    achieved = sample_cols * F_REDUCE / kavg / 1e12
    traffic = None
    pmc = ROOT / "profiles" / "reduce_f64_universal_traffic.json"
    if pmc.is_file():
        traffic = json.loads(pmc.read_text()).get("bytes_per_launch")
    out = {
        "workload": f"universal reduce: {n} walkers x {G} Y samples x {n_cols} columns ({cols.n_pairs} "
                    f"(delta, Delta) pairs x {cols.n_b} b), FP64, Y resident ({Y.nbytes / 1e6:.0f} MB)",
        "kernel_ms": kavg * 1e3, "kernel_ms_all": kern_ms, "ms_per_call": 1e3 * wall / steps,
        "sample_columns_per_s": sample_cols / kavg,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic, "kernel_ms": kavg * 1e3,
                     "flops_per_sample_column": F_REDUCE},
    }
    log(f"universal reduce: {kavg * 1e3:.2f} ms per launch, {achieved:.1f} TFLOP/s")
    if cpu_baseline:
        # the reference's CPU reduce (_reduce_cpu: numpy cos/sin over all
        # walkers and axes, one column at a time) on an evenly spread subset
        from oracle import oracle as O
        got_c = cs.cpu().numpy()
        sub = np.arange(0, n_cols, 7)
        t0 = time.perf_counter()
        done = 0
        for c in sub:
            O.reduce_reference(Y, cols.j_delta[c:c + 1], cols.j_Delta[c:c + 1], cols.j_sum[c:c + 1],
                               cols.phase_coef[c:c + 1])
            done += 1
            if time.perf_counter() - t0 > args.cpu_seconds:
                break
        dt = time.perf_counter() - t0
        ocs, _ = O.reduce_reference(Y, cols.j_delta[sub[:8]], cols.j_Delta[sub[:8]], cols.j_sum[sub[:8]],
                                    cols.phase_coef[sub[:8]])
        out["cpu_baseline"] = {
            "value": 3.0 * n * done / dt, "unit": "sample-columns/s", "cores": 1, "kind": "port",
            "sample": f"{done} of the {n_cols} columns over all {n} walkers x 3 axes with the reference's CPU "
                      f"reduce arithmetic (numpy, one process) in {dt:.1f} s on {cpu_model()}",
            "max_rel_diff_vs_gpu": float(np.max(np.abs(got_c[sub[:8]] - ocs) / np.maximum(np.abs(ocs), 1.0))),
        }
    return out


def run_reduce(args):
    log = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)  # noqa: E731
    r = measure_reduce(args, log, cpu_baseline=not args.no_cpu_baseline)
    out = {
        "metric": "sample-columns/s", "value": r["sample_columns_per_s"], "unit": "sample-columns/s", "n_gpus": 1,
        "steps": max(5, args.steps), "warmup": 1, "ms_per_step": r["ms_per_call"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": r["workload"]}, "roofline": r["roofline"],
    }
    if "cpu_baseline" in r:
        out["cpu_baseline"] = r["cpu_baseline"]
    print(json.dumps(out), flush=True)
    return 0


# The reference's published walk-plus-reduction times on a full NVIDIA
# A100-SXM4-80GB, one ensemble x one k_io at the production grid's low,
# centre and high rho (docs/v5_fast_classifier_launch_readiness.md:210-219).
PUBLISHED_A100_S = {"exact": {"low": 34.825, "center": 52.957, "high": 87.590},
                    "exact_cached": {"low": 3.533, "center": 5.105, "high": 16.260}}


def production_coordinate(which: str):
    """The reference benchmark's coordinate (analysis/v5_runtime_benchmark.py:101-123):
    rho index 0 / 21 / last of the production log grid, the admissible V whose
    v_i is nearest (in log) to the geometric centre of the grid's v_i range."""
    from custom_madi_amd.library import make_remediation_log_grid
    grid = make_remediation_log_grid()
    idx = {"low": 0, "center": 21, "high": len(grid.rhos) - 1}[which]
    target = float(np.sqrt(grid.vi_min * grid.vi_max))
    rho = float(grid.rhos[idx])
    cands = [(float(V), rho * float(V) * 1e-6) for V in grid.Vs if grid.vi_min <= rho * float(V) * 1e-6 <= grid.vi_max]
    V, vi = min(cands, key=lambda c: abs(np.log(c[1] / target)))
    return rho, V, vi


def measure_production(args, log, sweep=False, exact_coords=("low", "center", "high")):
    """The reference's own production benchmark, through the drop-in API:
    _walk_and_reduce_one_ensemble of one Eq. S8 ensemble (50,000 walkers x
    128,000 steps of 1 us, D0 3.0) at k_io 20 /s, reduced onto the 31,125
    production columns, at the three density coordinates; geometry and the
    first launch are outside the timed region, as in the reference's timing.
    Three ways: the FP64 exact walk (reference semantics), the FP64
    exact_cached walk (graph labels, bitwise the exact one) and the FP32
    production walk; each against the published A100 times.  The FP32 |dS|
    is set beside the FP64 walk's own seed-to-seed |dS|.  With sweep, also
    the library build's 51-k_io sweep of the ensemble in FP64 exact_cached
    (its walks run together, madi_walk_f64_multi).  The exact walk runs at
    exact_coords only; FP32 is compared with it there, else with
    exact_cached (bitwise the same walk)."""
    from dataclasses import replace
    from custom_madi_amd import _hip, signal, walker_gpu
    from custom_madi_amd.config import SimConfig
    from custom_madi_amd.library import make_remediation_log_grid
    _hip.require_gpu()
    torch = _hip.torch_mod()
    base = SimConfig(n_ensembles=1)
    cols = signal.build_columns(base)
    args_cols = (cols.j_delta, cols.j_Delta, cols.j_sum, cols.phase_coef)
    modes = {"exact": base, "exact_cached": replace(base, classifier_mode="exact_cached"),
             "fp32": replace(base, precision="fp32")}
    want = [m.strip() for m in args.prod_modes.split(",") if m.strip()]
    coords = [c.strip() for c in args.prod_coords.split(",") if c.strip()]
    if not want or any(m not in modes for m in want) or not coords or \
            any(c not in ("low", "center", "high") for c in coords):
        raise SystemExit(f"--prod-modes takes {sorted(modes)}, --prod-coords low, center, high")
    modes = {m: modes[m] for m in modes if m in want}
    out = {}
    for which in coords:
        rho, V, vi = production_coordinate(which)
        t0 = time.perf_counter()
        ens = walker_gpu._build_ensembles_for_entry(rho, V, base, seed=args.prod_seed, verbose=False)[0]
        ens.device_geometry()
        geom_s = time.perf_counter() - t0
        n_pop = int(len(ens.seeds))
        log(f"production {which}: rho {rho:.4g}/uL, V {V:.4g} pL, v_i {vi:.3f}: {n_pop} seeds, geometry {geom_s:.1f} s")
        pp = walker_gpu._checked_pp(walker_gpu.kio_to_pp(20.0, ens.mean_AV, base))
        seed = walker_gpu._walk_seed(args.prod_seed, 0)
        row = {"rho_per_uL": rho, "V_pL": V, "target_vi": vi, "n_seeds_pop": n_pop, "geometry_s": geom_s, "pp": pp}
        res = {}
        for mode, cfg in modes.items():
            if mode == "exact" and which not in exact_coords:
                continue
            # first launch of each kernel outside the timed region
            walker_gpu._walk_and_reduce_one_ensemble(ens, pp, replace(cfg, n_walkers=64), *args_cols, seed)
            torch.cuda.synchronize()
            t = time.perf_counter()
            r = walker_gpu._walk_and_reduce_one_ensemble(ens, pp, cfg, *args_cols, seed)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t
            res[mode] = r
            steps = cfg.n_walkers * cfg.n_steps
            entry = {"walk_and_reduction_s": dt, "walker_steps_per_s": steps / dt, "n_escaped": int(r.n_escaped)}
            if mode in PUBLISHED_A100_S:
                pub = PUBLISHED_A100_S[mode][which]
                entry["published_a100_s"] = pub
                entry["speedup_over_published_a100"] = pub / dt
            entry["speedup_over_published_a100_cached"] = PUBLISHED_A100_S["exact_cached"][which] / dt
            row[mode] = entry
            log(f"production {which} {mode}: {dt:.3f} s ({steps / dt:.3e} walker-steps/s)")
        a, b = res.get("exact"), res.get("exact_cached")
        if a is not None and b is not None:
            row["exact_cached_bitwise_equal_to_exact"] = bool(
                np.array_equal(a.cos_sum, b.cos_sum) and np.array_equal(a.sin_sum, b.sin_sum)
                and np.array_equal(a.occupancy_counts, b.occupancy_counts))
        n3 = 3.0 * base.n_walkers
        ref64 = a if a is not None else b
        if ref64 is not None and "fp32" in res:
            d32 = (res["fp32"].cos_sum - ref64.cos_sum) / n3
            row["fp32_vs_fp64_max_abs_dS"] = float(np.max(np.abs(d32)))
            row["fp32_vs_fp64_rms_dS"] = float(np.sqrt(np.mean(d32 ** 2)))
        if b is not None:
            # the Monte Carlo noise of S at this size: the same FP64 walk with another seed
            r2 = walker_gpu._walk_and_reduce_one_ensemble(ens, pp, modes["exact_cached"], *args_cols,
                                                          walker_gpu._walk_seed(args.prod_seed, 1))
            dseed = (r2.cos_sum - b.cos_sum) / n3
            row["fp64_seed_to_seed_max_abs_dS"] = float(np.max(np.abs(dseed)))
            row["fp64_seed_to_seed_rms_dS"] = float(np.sqrt(np.mean(dseed ** 2)))
        if "fp32_vs_fp64_max_abs_dS" in row and "fp64_seed_to_seed_max_abs_dS" in row:
            log(f"production {which}: |dS| fp32 vs fp64 max {row['fp32_vs_fp64_max_abs_dS']:.2e} rms "
                f"{row['fp32_vs_fp64_rms_dS']:.2e}; fp64 seed to seed max {row['fp64_seed_to_seed_max_abs_dS']:.2e} "
                f"rms {row['fp64_seed_to_seed_rms_dS']:.2e}")
        if sweep and b is not None:
            # the library build's sweep: every production k_io of this ensemble, FP64 exact_cached
            kios = make_remediation_log_grid().kios
            pps = [walker_gpu._checked_pp(walker_gpu.kio_to_pp(float(k), ens.mean_AV, base)) for k in kios]
            torch.cuda.synchronize()
            t = time.perf_counter()
            rs = walker_gpu._walk_and_reduce_kio_sweep(ens, pps, modes["exact_cached"], *args_cols, seed)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t
            k20 = int(np.flatnonzero(kios == 20.0)[0])
            one = row["exact_cached"]["walk_and_reduction_s"]
            row["kio_sweep"] = {
                "n_kio": len(kios), "s": dt, "s_per_kio": dt / len(kios), "speedup_over_one_walk": one * len(kios) / dt,
                "speedup_over_published_a100_cached": PUBLISHED_A100_S["exact_cached"][which] * len(kios) / dt,
                "kio_20_bitwise_equal_to_single_walk": bool(
                    np.array_equal(rs[k20].cos_sum, b.cos_sum) and np.array_equal(rs[k20].sin_sum, b.sin_sum)
                    and np.array_equal(rs[k20].occupancy_counts, b.occupancy_counts)),
            }
            log(f"production {which}: k_io sweep of {len(kios)} walks {dt:.2f} s ({dt / len(kios):.3f} s per k_io)")
            del rs
        out[which] = row
        del ens
    return {"workload": "production: the reference's walk-plus-reduction benchmark (one Eq. S8 ensemble, 50,000 "
                        "walkers x 128,000 steps of 1 us, D0 3.0, k_io 20/s, 31,125 columns) at the production "
                        "grid's low / centre / high rho", "coordinates": out}


def run_production(args):
    log = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)  # noqa: E731
    r = measure_production(args, log, sweep=True)
    c = r["coordinates"]
    if "high" not in c or "exact_cached" not in c["high"]:
        print(json.dumps({"metric": "production walks (partial run)", "coordinates": c}), flush=True)
        return 0
    hi = c["high"]["exact_cached"]
    out = {
        "metric": "walk + reduction seconds per ensemble x k_io (the reference's production benchmark)",
        "value": hi["walk_and_reduction_s"], "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": 1e3 * hi["walk_and_reduction_s"], "higher_is_better": False, "scaling": "weak",
        "vs_baseline": hi["speedup_over_published_a100"], "dtype": "f64 (reference semantics)", "data": "synthetic",
        "config": {"workload": r["workload"], "headline": "high rho, FP64 exact_cached"},
        "coordinates": c,
    }
    print(json.dumps(out), flush=True)
    return 0


# The reference's published production task times on a full A100: one
# (rho, V) group of 40 ensembles x 51 k_io in FP64 exact_cached, planned from
# its one-ensemble, one-k_io times (docs/v5_fast_classifier_launch_readiness.md:210-219).
PUBLISHED_A100_GROUP_H = {"low": 2.00, "center": 2.89, "high": 9.21}


def run_group(args):
    """One production task of the reference's library build: the (rho, V)
    group of 40 Eq. S8 ensembles x the 51 production k_io values (50,000
    walkers x 128,000 steps each, 31,125 columns), FP64 exact_cached, through
    run_simulation_multi_kio_reduced -- each ensemble's geometry built (the
    next one's CPU part while the current one walks) and its k_io walks run
    in one sweep launch -- at each of --prod-coords.  Geometry is inside the
    timed region; the published A100 group time is its walks alone."""
    import contextlib
    from dataclasses import replace
    from custom_madi_amd import _hip, signal, walker_gpu
    from custom_madi_amd.config import SimConfig
    from custom_madi_amd.library import make_remediation_log_grid
    log = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)  # noqa: E731
    _hip.require_gpu()
    torch = _hip.torch_mod()
    # the SI v_i acceptance tolerance is 0.01 instead of 0.005: with seed 0
    # one of the low-rho group's 40 finite packings (ensemble 37) lands 0.0058
    # from its target and the reference's acceptance test rejects it; the
    # tolerance does not change the work of a walk
    base = SimConfig(n_ensembles=int(args.group_ensembles), classifier_mode="exact_cached",
                     geometry_vi_tolerance=0.01)
    cols = signal.build_columns(base)
    cargs = (cols.j_delta, cols.j_Delta, cols.j_sum, cols.phase_coef)
    kios = [float(k) for k in make_remediation_log_grid().kios]
    coords = [c.strip() for c in args.prod_coords.split(",") if c.strip()]
    out = {}
    for which in coords:
        rho, V, vi = production_coordinate(which)
        # first launch of the kernels outside the timed region
        with contextlib.redirect_stdout(sys.stderr):
            walker_gpu.run_simulation_multi_kio_reduced(rho, V, kios[:2], *cargs,
                                                        cfg=replace(base, n_walkers=64, n_ensembles=1),
                                                        seed=args.prod_seed, verbose=False)
        torch.cuda.synchronize()
        log(f"group {which}: rho {rho:.4g}/uL, V {V:.4g} pL: {base.n_ensembles} ensembles x {len(kios)} k_io")
        t = time.perf_counter()
        with contextlib.redirect_stdout(sys.stderr):  # per-ensemble progress
            res = walker_gpu.run_simulation_multi_kio_reduced(rho, V, kios, *cargs, cfg=base, seed=args.prod_seed,
                                                              verbose=True)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        pub = PUBLISHED_A100_GROUP_H[which] if base.n_ensembles == 40 else None
        out[which] = {"rho_per_uL": rho, "V_pL": V, "target_vi": vi, "n_ensembles": base.n_ensembles,
                      "n_kio": len(kios), "walkers_per_ensemble": base.n_walkers, "s": dt, "h": dt / 3600.0,
                      "s_per_ensemble": dt / base.n_ensembles,
                      "walker_steps_per_s": base.n_ensembles * len(kios) * base.n_walkers * base.n_steps / dt,
                      "n_walkers_merged": int(res[kios[0]].n_walkers)}
        if pub is not None:
            out[which]["published_a100_h"] = pub
            out[which]["speedup_over_published_a100"] = pub * 3600.0 / dt
        log(f"group {which}: {dt:.1f} s ({dt / 3600.0:.3f} h" +
            (f", published A100 {pub} h: {pub * 3600.0 / dt:.1f}x)" if pub else ")"))
    head = out.get("high") or next(iter(out.values()))
    print(json.dumps({
        "metric": "production task time (one (rho, V) group: 40 ensembles x 51 k_io, FP64 exact_cached)",
        "value": head["h"], "unit": "h", "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": 1e3 * head["s"],
        "higher_is_better": False, "scaling": "weak", "vs_baseline": head.get("speedup_over_published_a100"),
        "dtype": "f64 (reference semantics)", "data": "synthetic",
        "config": {"workload": "group: the reference's production task through run_simulation_multi_kio_reduced "
                               "(Eq. S8 geometry per ensemble inside the timed region)",
                   "coordinates": coords, "seed": args.prod_seed,
                   "geometry_vi_tolerance": base.geometry_vi_tolerance},
        "groups": out,
    }), flush=True)
    return 0


def run_library(args):
    """The drop-in library build (build_library_from_triplets) at the default
    production configuration, group by group and on the batched scheduler."""
    import tempfile
    from custom_madi_amd import _hip, library, serve
    from custom_madi_amd.config import SimConfig
    log = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)  # noqa: E731
    # ensemble construction for the batched build runs in worker processes,
    # spawned before this process touches the GPU
    pool = serve.make_pool(args.workers) if args.workers > 0 else None
    if pool is not None:
        pool.map(_noop, range(args.workers))
    _hip.require_gpu()
    cfg = SimConfig(precision="fp32")
    groups = [(5.0e5, 1.2), (2.0e5, 3.0), (1.0e5, 6.0)][:max(1, args.library_groups)]
    kios = (5.0, 22.0, 60.0)[:max(1, args.library_kios)]
    triplets = [(k, r, v) for r, v in groups for k in kios]
    n_walks = len(triplets) * cfg.n_ensembles
    steps = n_walks * cfg.n_walkers * cfg.n_steps
    with tempfile.TemporaryDirectory() as d:
        times = {}
        for mode in ("sequential", "batched"):
            log(f"library: {len(triplets)} entries, {mode}")
            t0 = time.perf_counter()
            library.build_library_from_triplets(triplets, cfg=cfg, seed=11, save_path=f"{d}/{mode}.npz",
                                                verbose=False, batched=(mode == "batched"),
                                                pool=pool if mode == "batched" else None)
            times[mode] = time.perf_counter() - t0
            log(f"library: {mode} {times[mode]:.1f} s")
        a, b = np.load(f"{d}/sequential.npz"), np.load(f"{d}/batched.npz")
        same = sorted(a.files) == sorted(b.files) and all(
            np.array_equal(a[k], b[k], equal_nan=a[k].dtype.kind in "fc") for k in a.files)
    if pool is not None:
        pool.close()
        pool.join()
    eph = {m: len(triplets) * 3600.0 / t for m, t in times.items()}
    out = {
        "metric": "library entries/h", "value": eph["batched"], "unit": "entries/h", "n_gpus": 1, "steps": 1,
        "warmup": 0, "ms_per_step": 1e3 * times["batched"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 (fp64 sums)", "data": "synthetic",
        "config": {"workload": f"library: build_library_from_triplets, {len(triplets)} entries ({len(groups)} "
                               f"(rho, V) groups x {len(kios)} k_io), default configuration: {cfg.n_ensembles} "
                               f"ensembles x {cfg.n_walkers} walkers x {cfg.n_steps} steps, "
                               f"{len(cfg.delta_pairs()) * len(cfg.b_values)} columns (v5)",
                   "walks": n_walks, "walker_steps": steps},
        "sequential": {"wall_s": times["sequential"], "entries_per_hour": eph["sequential"],
                       "walker_steps_per_s": steps / times["sequential"]},
        "batched": {"wall_s": times["batched"], "entries_per_hour": eph["batched"],
                    "walker_steps_per_s": steps / times["batched"]},
        "batched_over_sequential": times["sequential"] / times["batched"],
        "files_equal": bool(same),
        "host_workers": {"sequential": 0, "batched": int(args.workers)},
    }
    print(json.dumps(out), flush=True)
    return 0


def _prepared_lookup(preps):
    table = {p.req.key: p for p in preps}
    return lambda req: table[req.key]


def measure_c3(args, pool, log, duration):
    """Serving: Poisson arrivals of entry requests, continuous batching, p50/p99 latency."""
    from custom_madi_amd import _hip, batch, serve, workloads
    reqs, arrivals = workloads.c3_requests(args.rate, duration)
    spec = workloads.acquisition()
    _hip.require_gpu()
    warm = batch.library_requests(2, seed=99, walkers=(20_000, 20_000), cells=(60, 80))
    batch.run_plain(warm, spec)
    knobs, knob_src = _tuning(args, spec, log)
    sch = batch.DeviceScheduler(spec, tier_cap_um=args.tier_cap, max_tiles_per_generation=args.gen_tiles, **knobs)
    stream = serve.LibraryStream(args.library_out)
    srv = serve.Server(sch, pool=pool, stream=stream)
    offered = sum(r.n_walkers * int(round(r.T_max_ms / r.ts)) for r in reqs) / duration
    log(f"{len(reqs)} requests over {duration:.0f} s, offered {offered:.3e} walker-steps/s")
    rep = serve.serve_poisson(srv, reqs, arrivals, progress=log)
    sch.close()
    stream.close()
    hbm = rep.pop("hbm_used_bytes_samples")
    return {"workload": f"c3: Poisson {args.rate}/s for {duration:.0f} s, entries as c2 with 1e5-1e6 walkers, "
                        "480 signals", "offered_walker_steps_per_s": offered,
            "hbm_used_gib_max": max((b for _, b in hbm), default=0) / 2 ** 30,
            "hbm_used_gib_samples": [[round(t, 1), round(b / 2 ** 30, 2)] for t, b in hbm], **rep,
            "tuning": {**knobs, "source": knob_src}, "classifier": sch.counter_dict()}


def run_c3(args):
    from custom_madi_amd import serve
    pool = serve.make_pool(args.workers)
    pool.map(_noop, range(args.workers))
    log = lambda m: print(f"[bench c3] {m}", file=sys.stderr, flush=True)  # noqa: E731
    out = measure_c3(args, pool, log, args.duration)
    pool.close()
    pool.join()      # workers exit on their own (no SIGTERM at interpreter exit)
    print(json.dumps({"metric": "p50/p99 entry latency", "unit": "s", "n_gpus": 1, "higher_is_better": False,
                      **out}), flush=True)
    return 0


def run_c4(args):
    """Full-range library sweep (14,105 entries): a seeded sample batched, FP64-oracle error on a sub-sample."""
    from custom_madi_amd import batch, engine, serve
    from custom_madi_amd.ensemble import DeviceGeometry, create_periodic_ensemble, geometry_arrays
    allreqs = batch.sweep_requests()
    rng = np.random.default_rng(4)
    n = len(allreqs) if args.entries <= 0 else min(args.entries, len(allreqs))
    pick = np.sort(rng.choice(len(allreqs), size=n, replace=False))
    reqs = [allreqs[i] for i in pick]
    check = [reqs[i] for i in np.sort(rng.choice(len(reqs), size=min(args.fp64_entries, len(reqs)), replace=False))]
    spec = batch.acquisition_spec(24)
    pool = serve.make_pool(args.workers)
    pool.map(_noop, range(args.workers))
    from custom_madi_amd import _hip
    _hip.require_gpu()
    log = lambda m: print(f"[bench c4] {m}", file=sys.stderr, flush=True)  # noqa: E731
    walker_steps = sum(r.n_walkers * int(round(r.T_max_ms / r.ts)) for r in reqs)
    log(f"{len(reqs)} of {len(allreqs)} entries, {walker_steps:.3e} walker-steps")
    knobs, knob_src = _tuning(args, spec, log, sample=reqs[:: max(1, len(reqs) // 6)][:6], tag="c4")
    sch = batch.DeviceScheduler(spec, tier_cap_um=args.tier_cap, max_tiles_per_generation=args.gen_tiles, **knobs)
    srv = serve.Server(sch, pool=pool)
    t0 = time.perf_counter()
    for r in reqs:
        srv.submit(r)
    srv.drain(progress=lambda: log(f"{len(srv.results)} of {len(reqs)} entries done, "
                                   f"{time.perf_counter() - t0:.0f} s"))
    wall = time.perf_counter() - t0
    sch.close()
    pool.close()
    pool.join()      # workers exit on their own (no SIGTERM at interpreter exit)
    res = {r.key: r for r in srv.results}
    log(f"batched: {wall:.1f} s")
    errs = []
    for r in check:
        ens = create_periodic_ensemble(r.n_cells, r.rho, r.rho * r.V * 1e-6, r.config(), seed=r.seed)
        dg = DeviceGeometry(geometry_arrays(ens))
        p32 = res[r.key]
        p = engine.walk_params(r.n_walkers, r.config(), p32.pp, r.walk_seed)
        f64 = engine.walk_fused(dg, p, spec, precision=64)
        S32, S64 = p32.cos / r.n_walkers, f64.cos / r.n_walkers
        se = np.sqrt(np.maximum(1.0 - S64 ** 2, 0.0) / r.n_walkers)
        dS = np.abs(S32 - S64)
        errs.append({"key": list(r.key), "cells": r.n_cells, "ts_ms": r.ts, "max_abs_dS": float(dS.max()),
                     "max_abs_dS_over_3se": float((dS / (3 * se + 1e-300)).max()),
                     "per_Delta_b_max_abs_dS": dS.max(axis=2).tolist(), "fp64_escaped": f64.n_escaped})
        log(f"fp64 check {r.key}: max |dS| {dS.max():.2e}")
    esc = len(srv.failed)  # failed entries (escape or classifier failure)
    lat = [r.latency_s for r in srv.results]
    out = {"metric": "library entries/h", "value": len(reqs) / wall * 3600.0, "unit": "entries/h", "n_gpus": 1,
           "higher_is_better": True,
           "workload": f"c4: full-range sweep, {len(reqs)} of {len(allreqs)} entries (seeded sample; V 0.01-206 pL, "
                       "v_i 0.5/0.747/0.994, rho 4.4e3-5.8e7, k_io 0-130 x 65, 32-8000 cells, 5e5 walkers, "
                       "dt 0.01 ms below 0.5 pL else 0.02 ms, 480 signals), FP32 walk, FP64 sums",
           "wall_s": wall, "walker_steps_per_s": walker_steps / wall, "escapes_or_errors": esc,
           "latency_s_p50": float(np.percentile(lat, 50)), "latency_s_p99": float(np.percentile(lat, 99)),
           "full_sweep_walker_steps": sum(r.n_walkers * int(round(r.T_max_ms / r.ts)) for r in allreqs),
           "fp64_check": {"entries": len(errs), "max_abs_dS": max((e["max_abs_dS"] for e in errs), default=None),
                          "per_entry": errs},
           "tuning": {**knobs, "source": knob_src}, "classifier": sch.counter_dict()}
    print(json.dumps(out), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--walkers", type=int, default=1_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64-check", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target seconds per CPU baseline sample")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run a 1 GiB device copy first (reference kernel for HBM counter scaling)")
    ap.add_argument("--workload", choices=["c1", "c2", "c3", "c4", "map", "reduce", "library", "production", "group"],
                    default="c1")
    ap.add_argument("--group-ensembles", type=int, default=40,
                    help="group: ensembles of the (rho, V) group (the reference's production task has 40)")
    ap.add_argument("--prod-seed", type=int, default=0, help="production: geometry and walk seed")
    ap.add_argument("--prod-coords", default="low,center,high",
                    help="production: comma list of the density coordinates to run (low, center, high)")
    ap.add_argument("--prod-modes", default="exact,exact_cached,fp32",
                    help="production: comma list of walks to time (exact, exact_cached, fp32); the FP32 |dS| "
                         "needs exact and fp32, the seed-to-seed |dS| and the sweep exact_cached")
    ap.add_argument("--voxels", type=int, default=1_000_000, help="map: voxels matched per step")
    ap.add_argument("--library-entries", type=int, default=14_105, help="map: library entries")
    ap.add_argument("--entries", type=int, default=256)
    ap.add_argument("--plain-entries", type=int, default=256)
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--grid-per-cu", type=int, default=0)
    ap.add_argument("--tier-cap", type=float, default=0.0)
    ap.add_argument("--scan-rounds", type=int, default=0,
                    help="FP32 walk: classification rounds per iteration, 1-2 (0 default)")
    ap.add_argument("--classify-mode", type=int, default=0,
                    help="FP32 walk: 1 per-lane state machine per tile, 3 state machine with lane refill, "
                         "4 walker waves with classifier waves (0 default)")
    ap.add_argument("--gen-tiles", type=int, default=65536)
    ap.add_argument("--no-tune", action="store_true", help="c2/c3/c4: default knobs instead of the autotuner")
    ap.add_argument("--fp64-entries", type=int, default=16, help="c4: entries checked against the FP64 walk")
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--duration", type=float, default=60.0, help="c3 workload: seconds of arrivals")
    ap.add_argument("--c3-duration", type=float, default=60.0, help="default line: seconds of c3 arrivals")
    ap.add_argument("--no-c2", action="store_true", help="default line without the c2 object")
    ap.add_argument("--no-c3", action="store_true", help="default line without the c3 object")
    ap.add_argument("--no-reduce", action="store_true", help="default line without the universal_reduce object")
    ap.add_argument("--no-production", action="store_true", help="default line without the production object")
    ap.add_argument("--library-groups", type=int, default=2, help="library: (rho, V) groups (at most 3)")
    ap.add_argument("--library-kios", type=int, default=3, help="library: k_io values per group (at most 3)")
    ap.add_argument("--library-out", default=None)
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: library entries are "
              "independent, so N GPUs would each run their own share of the entries with no collective",
              file=sys.stderr)
        return 2
    if args.workload == "c2":
        return run_c2(args)
    if args.workload == "c3":
        return run_c3(args)
    if args.workload == "c4":
        return run_c4(args)
    if args.workload == "map":
        return run_map(args)
    if args.workload == "reduce":
        return run_reduce(args)
    if args.workload == "group":
        return run_group(args)
    if args.workload == "production":
        return run_production(args)
    if args.workload == "library":
        return run_library(args)

    log = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)  # noqa: E731
    pool = None
    if not (args.no_c2 and args.no_c3 and args.no_cpu_baseline):
        from custom_madi_amd import serve
        pool = serve.make_pool(args.workers)       # spawned before the GPU is initialised
        pool.map(_noop, range(args.workers))
    out, (ens, cfg, pp) = measure_c1(args, log)
    if not args.no_c2:
        log("c2: 256 uneven entries, plain and batched")
        out["c2"] = measure_c2(args, pool, lambda m: log(f"c2: {m}"))
    if not args.no_c3:
        log(f"c3: Poisson arrivals at {args.rate}/s for {args.c3_duration:.0f} s")
        out["c3"] = measure_c3(args, pool, lambda m: log(f"c3: {m}"), args.c3_duration)
    if not args.no_reduce:
        out["universal_reduce"] = measure_reduce(args, log, cpu_baseline=False)
    if not args.no_production:
        log("production: the reference's walk-plus-reduction benchmark at its three densities")
        # the exact walk (reference semantics, ~20 s a coordinate) at low rho only;
        # --workload production runs it at all three
        out["production"] = measure_production(args, log, sweep=True, exact_coords=("low",))
    if not args.no_cpu_baseline:
        procs, vis = max(1, int(args.workers)), visible_cpus()
        single, multi = cpu_baseline_reference(ens, cfg, pp, pool, procs, args.cpu_seconds, log)
        threads = max(1, min(procs, vis))
        rate_c, dt_c, n_c = cpu_baseline_c_oracle(ens, cfg, pp, threads, target_s=args.cpu_seconds,
                                                  max_walkers=args.walkers)
        cpu = cpu_model()
        out["cpu_baseline"] = {
            "value": multi["value"], "unit": "walker-steps/s", "cores": multi["cores"], "kind": "port",
            "sample": f"the reference CPU walk algorithm (numpy/scipy restatement of _walk_cpu + classify_cpu: "
                      f"cKDTree nearest and ball queries, per-point facet loop; FP64) on the same entry's "
                      f"image-padded seeds, {multi['cores']} processes x {multi['walkers'] // multi['cores']} "
                      f"walkers x {cfg.n_steps} steps in {multi['wall_s']:.1f} s on {cpu} ({vis} CPUs visible; "
                      f"walker-steps/s is linear in walkers, so the sample rate is the entry's)",
            "single_process": {**single, "unit": "walker-steps/s",
                               "sample": f"1 process x {single['walkers']} walkers x {cfg.n_steps} steps in "
                                         f"{single['wall_s']:.1f} s"},
            "c_oracle": {"value": rate_c, "unit": "walker-steps/s", "cores": threads,
                         "sample": f"{n_c} walkers x {cfg.n_steps} steps, FP64 C restatement "
                                   f"(oracle/madi_oracle.c, OpenMP, {threads} threads) in {dt_c:.1f} s"},
        }
    if pool is not None:
        pool.close()
        pool.join()      # workers exit on their own (no SIGTERM at interpreter exit)
    print(json.dumps(out), flush=True)
    return 0


def measure_c1(args, log):
    """The C1 entry: headline walker-steps/s, roofline and the FP64 check."""
    import torch
    from custom_madi_amd import _hip, engine, workloads
    _hip.require_gpu()
    info = _hip.device_info()
    if args.copy_probe:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    cfg, ens, pp, spec = workloads.c1_entry(args.walkers)
    dg = ens.device_geometry()
    N = args.walkers
    runner = engine.FusedRunner(N, spec, cfg.n_grid)
    params = engine.walk_params(N, cfg, pp, seed=0, tier_cap_um=args.tier_cap, scan_rounds=args.scan_rounds,
                                classify_mode=args.classify_mode)
    log(f"device {info['name']} {info['arch']} CUs={info['cu_count']}; workload c1 walkers={N} "
        f"steps={cfg.n_steps} pp={pp:.6g}")

    for _ in range(args.warmup):
        runner.launch(dg, params, 32)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record(stream)
        runner.launch(dg, params, 32)
        ends[k].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    res = runner.result(N)
    kern_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    ms_per_step = 1e3 * wall / args.steps
    walker_steps = N * cfg.n_steps
    value = walker_steps * args.steps / wall
    if res.n_escaped or res.n_error:
        raise RuntimeError(f"walk reported escapes={res.n_escaped} errors={res.n_error}")

    # madi_walk_fused zeroes the counters: they hold the last launch's values
    facets = res.counters["facets"]
    flops = walker_steps * F_STEP + facets * F_FACET + N * F_WALKER
    kavg = float(np.mean(kern_ms)) * 1e-3
    achieved = flops / kavg / 1e12
    traffic = None
    pmc = ROOT / "profiles" / "split_f32_c1_traffic.json"
    if pmc.is_file():
        traffic = json.loads(pmc.read_text()).get("bytes_per_launch")

    out = {
        "metric": "walker-steps/s", "value": value, "unit": "walker-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 (fp64 final sums)", "data": "synthetic",
        "config": {"workload": "c1: periodic 1000-cell box, rho 5e5/uL, V 1.2 pL, k_io 22/s, D 2.0, dt 0.02 ms "
                               "x 2300 steps, 1e6 walkers, 4 Delta x 5 b x 24 dirs = 480 signals",
                   "walkers": N, "steps_per_walk": cfg.n_steps, "cells": 1000, "signals": 480},
        "c1_kernel_ms_p50": float(np.percentile(kern_ms, 50)),
        "c1_kernel_ms_p99": float(np.percentile(kern_ms, 99)),
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "frac_nonpacked_peak": achieved / PEAK_FP32_NONPACKED_TFLOPS,
                     "kernel_ms": kavg * 1e3, "algorithmic_tflop_per_launch": flops / 1e12},
        "classifier": dict(res.counters),
    }

    if not args.no_fp64_check:
        # the FP64 exact-classifier walk on the same seeds, with the 4b
        # columns that give each signal's standard error
        specv, idx, idx4 = workloads.acquisition_with_variance()
        t64 = time.perf_counter()
        f64 = engine.walk_fused(dg, params, specv, precision=64)
        t64 = time.perf_counter() - t64
        S64, se = workloads.signal_se(f64.cos, N, idx, idx4)
        dS = np.abs(res.cos / N - S64)
        ratio = dS / np.maximum(se, 1e-300)
        out["fp64_check"] = {
            "max_abs_dS": float(dS.max()), "max_abs_dS_over_se": float(ratio.max()),
            "signals_within_3se": f"{int(np.count_nonzero(dS <= 3.0 * se))}/{dS.size}",
            "per_Delta_b_max_abs_dS": dS.max(axis=2).tolist(), "fp64_escaped": f64.n_escaped,
            "occupancy_max_abs_d": float(np.abs(res.occupancy - f64.occupancy).max() / N),
            "fp64_walk_s": t64, "fp64_kd_nodes": f64.counters["kd_nodes"],
        }
        # the same FP64 walk with graph labels (the reference's exact_cached
        # mode, classify_mode 5): bitwise the exact walk's results
        pg = engine.walk_params(N, cfg, pp, params.seed, classify_mode=engine.MADI_CLASSIFY_FP64_CACHED)
        tg = time.perf_counter()
        fg = engine.walk_fused(dg, pg, specv, precision=64)
        tg = time.perf_counter() - tg
        out["fp64_check"]["fp64_graph_labels"] = {
            "walk_s": tg, "walker_steps_per_s": N * cfg.n_steps / tg, "speedup_over_exact": t64 / tg,
            "bitwise_equal_to_exact": bool(np.array_equal(fg.cos, f64.cos) and np.array_equal(fg.sin, f64.sin)
                                           and np.array_equal(fg.occupancy, f64.occupancy)),
            "kd_classifications": fg.counters["fp64_recheck"], "graph_scans": fg.counters["classify"],
            "ball_answers": fg.counters["safe_hit"],
        }
        log(f"fp64 check: max |dS| {dS.max():.3e}, max |dS|/SE {ratio.max():.3f}; graph-labelled FP64 walk "
            f"{tg:.3f} s vs {t64:.3f} s exact")
    return out, (ens, cfg, pp)


if __name__ == "__main__":
    sys.exit(main())
