"""Device walk engine: one call per C-ABI entry point, torch tensors as buffers.

Every function here enqueues HIP kernels of libmadi_hip.so on the current
torch stream and synchronises before returning host arrays.
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _hip
from .config import GAMMA_RAD, SimConfig
from .ensemble import DeviceGeometry

AXIS_DIRECTIONS = np.eye(3)


@dataclass
class PhaseSpec:
    """Fusable acquisition: pairs given as Y-grid indices, pc[pair, b], unit dirs."""

    j_delta: np.ndarray  # (n_pairs,)
    j_Delta: np.ndarray
    j_sum: np.ndarray
    pc: np.ndarray       # (n_pairs, n_b)
    dirs: np.ndarray     # (n_dirs, 3)

    @property
    def shape(self):
        return (len(self.j_delta), self.pc.shape[1], len(self.dirs))

    def to_ctypes(self) -> _hip.PhasePlan:
        n_pairs, n_b, n_dirs = self.shape
        if not (1 <= n_pairs <= _hip.MADI_MAX_PAIRS and 1 <= n_b <= _hip.MADI_MAX_B
                and 1 <= n_dirs <= _hip.MADI_MAX_DIRS):
            raise ValueError("phase plan exceeds the fused kernel's limits")
        p = _hip.PhasePlan()
        p.n_pairs, p.n_b, p.n_dirs = n_pairs, n_b, n_dirs
        for k in range(n_pairs):
            p.j_delta[k] = int(self.j_delta[k])
            p.j_Delta[k] = int(self.j_Delta[k])
            p.j_sum[k] = int(self.j_sum[k])
        flat_pc = np.asarray(self.pc, dtype=np.float64).ravel()
        for k, v in enumerate(flat_pc):
            p.pc[k] = float(v)
        flat_d = np.asarray(self.dirs, dtype=np.float64).ravel()
        for k, v in enumerate(flat_d):
            p.dirs[k] = float(v)
        return p


def phase_coef(b_s_mm2: float, delta_ms: float, Delta_ms: float) -> float:
    """gamma * G(b, delta, Delta) * 1e-9 (madi/signal.py:38-45, 114-115)."""
    b_si = b_s_mm2 * 1e6
    d_si = delta_ms * 1e-3
    tD = (Delta_ms - delta_ms / 3.0) * 1e-3
    if b_si <= 0 or tD <= 0:
        return 0.0
    G = math.sqrt(b_si / ((GAMMA_RAD * d_si) ** 2 * tD))
    return GAMMA_RAD * G * 1e-9


def spec_from_columns(j_delta, j_Delta, j_sum, pcs, dirs=AXIS_DIRECTIONS):
    """Group reference columns (jd, jD, js, pc) into a PhaseSpec.

    Returns (spec, col_pair, col_b) mapping every input column to its
    (pair, b) slot, or None when the columns do not fit the fused kernel.
    """
    jd = np.asarray(j_delta, dtype=np.int64)
    jD = np.asarray(j_Delta, dtype=np.int64)
    js = np.asarray(j_sum, dtype=np.int64)
    pcs = np.asarray(pcs, dtype=np.float64)
    keys = {}
    col_pair = np.empty(len(pcs), dtype=np.int64)
    col_b = np.empty(len(pcs), dtype=np.int64)
    per_pair: list[list[float]] = []
    for c in range(len(pcs)):
        key = (int(jd[c]), int(jD[c]), int(js[c]))
        if key not in keys:
            keys[key] = len(keys)
            per_pair.append([])
        p = keys[key]
        col_pair[c] = p
        col_b[c] = len(per_pair[p])
        per_pair[p].append(float(pcs[c]))
    if not per_pair:
        return None
    n_b = max(len(v) for v in per_pair)
    if len(per_pair) > _hip.MADI_MAX_PAIRS or n_b > _hip.MADI_MAX_B or len(dirs) > _hip.MADI_MAX_DIRS:
        return None
    pc = np.zeros((len(per_pair), n_b))
    for p, vals in enumerate(per_pair):
        pc[p, : len(vals)] = vals
    triples = sorted(keys, key=keys.get)
    spec = PhaseSpec(
        j_delta=np.array([t[0] for t in triples]), j_Delta=np.array([t[1] for t in triples]),
        j_sum=np.array([t[2] for t in triples]), pc=pc, dirs=np.asarray(dirs, dtype=np.float64),
    )
    return spec, col_pair, col_b


MADI_CLASSIFY_FP64_CACHED = 5


def walk_params(n_walkers: int, cfg: SimConfig, pp: float, seed: int, first_walker: int = 0,
                tier_cap_um: float = 0.0, scan_rounds: int = 0, classify_mode: int = 0) -> _hip.WalkParams:
    """madi_walk_params of one walk.  An FP64 configuration with
    classifier_mode "exact_cached" (the reference's cached classifier) selects
    the FP64 walk's graph labels with their certified ball (labels equal to
    the exact classifier's), unless the caller picks a classify_mode."""
    if (classify_mode == 0 and getattr(cfg, "classifier_mode", "exact") == "exact_cached"
            and getattr(cfg, "precision", "fp64") != "fp32"):
        classify_mode = MADI_CLASSIFY_FP64_CACHED
    p = _hip.WalkParams()
    p.tier_cap_um = float(tier_cap_um)
    p.scan_rounds = int(scan_rounds)
    p.classify_mode = int(classify_mode)
    p.n_walkers = int(n_walkers)
    p.n_steps = int(cfg.n_steps)
    p.steps_per_h = int(cfg.steps_per_h)
    p.n_grid = int(cfg.n_grid)
    p.sigma = float(cfg.sigma)
    p.ts = float(cfg.ts)
    p.pp = float(pp)
    p.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    p.first_walker = int(first_walker)
    return p


def rng_states(seed: int, first: int, n: int) -> np.ndarray:
    """Device-computed xoroshiro128+ states (n x 2 uint64)."""
    torch = _hip.require_gpu()
    out = torch.empty(2 * n, dtype=torch.int64, device="cuda")
    lib = _hip.load_library()
    _hip.check(lib.madi_rng_init(int(seed) & 0xFFFFFFFFFFFFFFFF, int(first), int(n), _hip.dptr(out),
                                 _hip.stream_ptr()), "madi_rng_init")
    torch.cuda.synchronize()
    return out.cpu().numpy().view(np.uint64).reshape(n, 2)


@dataclass
class WalkY:
    Y: object            # device tensor (n, n_grid, 3) float64
    inside_trace: Optional[np.ndarray]  # (n, n_grid) int8 (None when only the occupancy was asked for)
    escaped: np.ndarray
    err: np.ndarray
    counters: dict
    occupancy: Optional[np.ndarray] = None  # (n_grid,) int64 inside counts, counted on the device


def walk_y(geom: DeviceGeometry, params: _hip.WalkParams, stream=None) -> WalkY:
    """FP64 reference-semantics walk writing Y(t) on the device."""
    torch = _hip.require_gpu()
    n, G = int(params.n_walkers), int(params.n_grid)
    Y = torch.empty((n, G, 3), dtype=torch.float64, device="cuda")
    itr = torch.empty((n, G), dtype=torch.int8, device="cuda")
    esc = torch.empty(n, dtype=torch.int8, device="cuda")
    err = torch.empty(n, dtype=torch.int8, device="cuda")
    cnt = torch.zeros(_hip.N_COUNTERS, dtype=torch.int64, device="cuda")
    gi = gd = gu = None
    if stream is not None:
        gi = _hip.to_device(np.asarray(stream.initial_positions, dtype=np.float64))
        gd = _hip.to_device(np.asarray(stream.increments, dtype=np.float64))
        gu = _hip.to_device(np.asarray(stream.acceptance_uniforms, dtype=np.float64))
    lib = _hip.load_library()
    rc = lib.madi_walk_f64(
        geom.handle, C.byref(params), _hip.dptr(gi) if gi is not None else None,
        _hip.dptr(gd) if gd is not None else None, _hip.dptr(gu) if gu is not None else None,
        _hip.dptr(Y), _hip.dptr(itr), _hip.dptr(esc), _hip.dptr(err), _hip.dptr(cnt), _hip.stream_ptr(),
    )
    _hip.check(rc, "madi_walk_f64")
    torch.cuda.synchronize()
    return WalkY(Y, itr.cpu().numpy(), esc.cpu().numpy(), err.cpu().numpy(),
                 dict(zip(_hip.COUNTER_NAMES, cnt.cpu().numpy().tolist())))


def walk_y_multi(geom: DeviceGeometry, params: _hip.WalkParams, pps, traces: bool = True) -> List[WalkY]:
    """Several FP64 walks that differ only in p_p, in one launch
    (madi_walk_f64_multi): walk k equals walk_y with params.pp = pps[k], bit
    for bit.  traces=False leaves the inside traces on the device and
    returns each walk's occupancy, counted there (madi_count_inside),
    instead."""
    torch = _hip.require_gpu()
    K = len(pps)
    if not 1 <= K <= _hip.MADI_MAX_MULTI_WALKS:
        raise ValueError(f"1 to {_hip.MADI_MAX_MULTI_WALKS} walks per launch, got {K}")
    n, G = int(params.n_walkers), int(params.n_grid)
    Y = torch.empty((K, n, G, 3), dtype=torch.float64, device="cuda")
    itr = torch.empty((K, n, G), dtype=torch.int8, device="cuda")
    esc = torch.empty((K, n), dtype=torch.int8, device="cuda")
    err = torch.empty((K, n), dtype=torch.int8, device="cuda")
    cnt = torch.zeros((K, _hip.N_COUNTERS), dtype=torch.int64, device="cuda")
    pp = np.ascontiguousarray(pps, dtype=np.float64)
    lib = _hip.load_library()
    rc = lib.madi_walk_f64_multi(geom.handle, C.byref(params), K, pp.ctypes.data, _hip.dptr(Y), _hip.dptr(itr),
                                 _hip.dptr(esc), _hip.dptr(err), _hip.dptr(cnt), _hip.stream_ptr())
    _hip.check(rc, "madi_walk_f64_multi")
    occ_h = None
    if not traces:
        occ = torch.empty((K, G), dtype=torch.int64, device="cuda")
        _hip.check(lib.madi_count_inside(_hip.dptr(itr), K, n, G, _hip.dptr(occ), _hip.stream_ptr()),
                   "madi_count_inside")
        occ_h = occ.cpu().numpy()
    torch.cuda.synchronize()
    esc_h, err_h, cnt_h = esc.cpu().numpy(), err.cpu().numpy(), cnt.cpu().numpy()
    itr_h = itr.cpu().numpy() if traces else None
    return [WalkY(Y[k], None if itr_h is None else itr_h[k], esc_h[k], err_h[k],
                  dict(zip(_hip.COUNTER_NAMES, cnt_h[k].tolist())), None if occ_h is None else occ_h[k])
            for k in range(K)]


@dataclass
class WalkYF32:
    Y: object              # device tensor (n, n_grid, 3) float64: the FP32 walk's Y widened
    occupancy: np.ndarray  # (n_grid,) int64 inside counts
    n_escaped: int
    n_error: int
    counters: dict


def walk_y_f32(geom: DeviceGeometry, params: _hip.WalkParams) -> WalkYF32:
    """The FP32 production walk writing Y(t) on the device (madi_walk_y_f32):
    the same walk as walk_fused at precision 32, for any column set."""
    torch = _hip.require_gpu()
    n, G = int(params.n_walkers), int(params.n_grid)
    lib = _hip.load_library()
    nbytes = C.c_int64()
    _hip.check(lib.madi_walk_y_f32_workspace(n, C.byref(nbytes)), "madi_walk_y_f32_workspace")
    ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device="cuda")
    Y = torch.empty((n, G, 3), dtype=torch.float64, device="cuda")
    occ = torch.empty(G, dtype=torch.int64, device="cuda")
    flags = torch.empty(_hip.N_FLAGS, dtype=torch.int64, device="cuda")
    cnt = torch.empty(_hip.N_COUNTERS, dtype=torch.int64, device="cuda")
    rc = lib.madi_walk_y_f32(geom.handle, C.byref(params), _hip.dptr(ws), int(nbytes.value), _hip.dptr(Y),
                             _hip.dptr(occ), _hip.dptr(flags), _hip.dptr(cnt), _hip.stream_ptr())
    _hip.check(rc, "madi_walk_y_f32")
    torch.cuda.synchronize()
    fl = flags.cpu().numpy()
    return WalkYF32(Y, occ.cpu().numpy(), int(fl[_hip.MADI_FLAG_ESCAPED]), int(fl[_hip.MADI_FLAG_CLASSIFIER_ERR]),
                    dict(zip(_hip.COUNTER_NAMES, cnt.cpu().numpy().tolist())))


def single_column_spec() -> "PhaseSpec":
    """A one-column phase plan (pair (0, 0, 0), pc 0, one direction): the plan
    of walks whose output is Y, whose single sum nobody reads."""
    return PhaseSpec(j_delta=np.array([0]), j_Delta=np.array([0]), j_sum=np.array([0]), pc=np.zeros((1, 1)),
                     dirs=np.array([[1.0, 0.0, 0.0]]))


def reduce_columns(Y_dev, j_delta, j_Delta, j_sum, pcs):
    """Sigma cos / sigma sin per column from a device Y (FP64)."""
    n, G, _ = Y_dev.shape
    n_cols = len(pcs)
    if not (len(j_delta) == len(j_Delta) == len(j_sum) == n_cols):
        raise ValueError("j_delta, j_Delta, j_sum and the phase coefficients need one entry per column")
    if n_cols > _hip.MADI_MAX_REDUCE_COLS:
        raise ValueError(f"{n_cols} columns exceed one reduce call's {_hip.MADI_MAX_REDUCE_COLS}")
    for name, j in (("j_delta", j_delta), ("j_Delta", j_Delta), ("j_sum", j_sum)):
        j = np.asarray(j)
        if j.size and (j.min() < 0 or j.max() >= G):
            raise ValueError(f"{name} holds grid indices outside [0, {G})")
    torch = _hip.require_gpu()
    jd = _hip.to_device(np.asarray(j_delta, dtype=np.int32))
    jD = _hip.to_device(np.asarray(j_Delta, dtype=np.int32))
    js = _hip.to_device(np.asarray(j_sum, dtype=np.int32))
    pc = _hip.to_device(np.asarray(pcs, dtype=np.float64))
    cs = torch.empty(n_cols, dtype=torch.float64, device="cuda")
    sn = torch.empty(n_cols, dtype=torch.float64, device="cuda")
    lib = _hip.load_library()
    rc = lib.madi_reduce_f64(_hip.dptr(Y_dev), int(n), int(G), _hip.dptr(jd), _hip.dptr(jD), _hip.dptr(js),
                             _hip.dptr(pc), int(n_cols), _hip.dptr(cs), _hip.dptr(sn), _hip.stream_ptr())
    _hip.check(rc, "madi_reduce_f64")
    torch.cuda.synchronize()
    return cs.cpu().numpy(), sn.cpu().numpy()


@dataclass
class FusedResult:
    cos: np.ndarray        # (n_pairs, n_b, n_dirs) sums over walkers
    sin: np.ndarray
    occupancy: np.ndarray  # (n_grid,) int64
    n_escaped: int
    n_error: int
    counters: dict
    n_walkers: int


class FusedRunner:
    """Reusable device buffers for repeated fused walks of one shape."""

    def __init__(self, n_walkers: int, spec: PhaseSpec, n_grid: int):
        torch = _hip.require_gpu()
        self.lib = _hip.load_library()
        self.spec = spec
        self.plan = spec.to_ctypes()
        need = C.c_int64()
        _hip.check(self.lib.madi_walk_fused_workspace(int(n_walkers), C.byref(self.plan), C.byref(need)),
                   "madi_walk_fused_workspace")
        self.ws_bytes = int(need.value)
        self.ws = torch.empty(self.ws_bytes, dtype=torch.uint8, device="cuda")
        n_cols = int(np.prod(spec.shape))
        self.cos = torch.empty(n_cols, dtype=torch.float64, device="cuda")
        self.sin = torch.empty(n_cols, dtype=torch.float64, device="cuda")
        self.occ = torch.empty(n_grid, dtype=torch.int64, device="cuda")
        self.flags = torch.empty(_hip.N_FLAGS, dtype=torch.int64, device="cuda")
        self.cnt = torch.zeros(_hip.N_COUNTERS, dtype=torch.int64, device="cuda")
        self.n_walkers = int(n_walkers)

    def launch(self, geom: DeviceGeometry, params: _hip.WalkParams, precision: int = 32, stream: int = None):
        if int(params.n_walkers) > self.n_walkers:
            raise ValueError("runner was sized for fewer walkers")
        rc = self.lib.madi_walk_fused(
            geom.handle, C.byref(params), C.byref(self.plan), int(precision), _hip.dptr(self.ws), self.ws_bytes,
            _hip.dptr(self.cos), _hip.dptr(self.sin), _hip.dptr(self.occ), _hip.dptr(self.flags),
            _hip.dptr(self.cnt), _hip.stream_ptr() if stream is None else stream,
        )
        _hip.check(rc, "madi_walk_fused")

    def result(self, n_walkers: int) -> FusedResult:
        torch = _hip.torch_mod()
        torch.cuda.synchronize()
        shape = self.spec.shape
        flags = self.flags.cpu().numpy()
        return FusedResult(
            cos=self.cos.cpu().numpy().reshape(shape), sin=self.sin.cpu().numpy().reshape(shape),
            occupancy=self.occ.cpu().numpy().copy(), n_escaped=int(flags[0]), n_error=int(flags[1]),
            counters=dict(zip(_hip.COUNTER_NAMES, self.cnt.cpu().numpy().tolist())), n_walkers=int(n_walkers),
        )


def walk_fused_verify(geom: DeviceGeometry, params: _hip.WalkParams, spec: PhaseSpec):
    """The FP32 split walk's verification build: (FusedResult, verify counts).

    Every label the walk takes (plane tier, certified exit, classifier wave)
    is checked against the exact FP64 KD classifier on the same represented
    point; the counts name how many of each kind were checked and how many
    disagree.  The geometry needs its KD tree and its Delaunay graph."""
    torch = _hip.require_gpu()
    runner = FusedRunner(int(params.n_walkers), spec, int(params.n_grid))
    ver = torch.empty(_hip.N_VERIFY, dtype=torch.int64, device="cuda")
    lib = runner.lib
    rc = lib.madi_walk_fused_verify(
        geom.handle, C.byref(params), C.byref(runner.plan), _hip.dptr(runner.ws), runner.ws_bytes,
        _hip.dptr(runner.cos), _hip.dptr(runner.sin), _hip.dptr(runner.occ), _hip.dptr(runner.flags),
        _hip.dptr(runner.cnt), _hip.dptr(ver), _hip.stream_ptr())
    _hip.check(rc, "madi_walk_fused_verify")
    res = runner.result(int(params.n_walkers))
    return res, dict(zip(_hip.VERIFY_NAMES, ver.cpu().numpy().tolist()))


def walk_fused(geom: DeviceGeometry, params: _hip.WalkParams, spec: PhaseSpec, precision: int = 32) -> FusedResult:
    runner = FusedRunner(int(params.n_walkers), spec, int(params.n_grid))
    runner.launch(geom, params, precision)
    return runner.result(int(params.n_walkers))
