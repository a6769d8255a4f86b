"""Writes walk_vectors.npz: the reference's own CPU walk (_walk_cpu,
madi/walker_gpu.py:698-840) and CPU column reduce (_reduce_cpu,
madi/walker_gpu.py:954-968) on one seeded Eq. S2 geometry with an injected
random stream (make_walk_random_stream, madi/walker_gpu.py:115-127).

The reference's committed CPU/GPU golden file is missing from its tree, so
this fixture is the walk's only anchor to the reference's own output on a
given geometry and stream.  The geometry is the reference's population for
generator seed 21 (rho = 2e6 cells/uL, Omega_sim side 40 um, v_i = 0.7 for
alpha*); 32 walkers x 160 steps of 0.02 ms with p_p = 0.3, so membranes are
crossed and crossings accepted and rejected.  The stream itself is not
stored: the tests regenerate it from its seed and check its SHA-256.
Run in a container that holds the reference tree at /root/reference.
"""

from __future__ import annotations

import hashlib
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, "/root/reference")
from madi import ensemble as RE  # noqa: E402
from madi import walker_gpu as RW  # noqa: E402
from madi.config import SimConfig  # noqa: E402

OUT = Path(__file__).resolve().parent / "walk_vectors.npz"
RHO, SIDE, GEOM_SEED, VI = 2.0e6, 40.0, 21, 0.7
N_WALKERS, TS, T_MAX, H, D0 = 32, 0.02, 3.2, 0.2, 2.0
STREAM_SEED, PP = 5, 0.3
# columns (j_delta, j_Delta, j_sum, phase coefficient) over the 17 grid points
COLS = ([2, 2, 3, 4], [6, 8, 10, 12], [8, 10, 13, 16], [0.0, 0.35, 0.8, 1.7])


def stream_digest(stream) -> str:
    h = hashlib.sha256()
    for a in (stream.initial_positions, stream.increments, stream.acceptance_uniforms):
        h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    return h.hexdigest()


def main() -> None:
    cfg = SimConfig(D0=D0, ts=TS, T_max_ms=T_MAX, h_ms=H, n_walkers=N_WALKERS, small_deltas=[0.4],
                    big_deltas=[1.2], b_values=[0.0, 1000.0])
    cfg.assert_grid_alignment()
    rng = np.random.default_rng(GEOM_SEED)
    seeds, _, nn, _ = RE._make_population(RE._rho_um3(RHO), SIDE, cfg, rng)
    alpha = RE.alpha_star_from_vi(VI, RHO)
    annulus = np.minimum(alpha, cfg.kappa * nn / 2.0)
    nodes = RE._build_gpu_kdtree(seeds)
    source_side = cfg.source_fraction * SIDE
    lo = 0.5 * (SIDE - source_side)
    ens = RE.Ensemble(seeds=seeds, annulus=annulus, rho=RHO, V=float("nan"), vi=float("nan"), alpha_star=alpha,
                      L=SIDE, source_lo=lo, source_hi=lo + source_side, mean_AV=1.0, geometry=None,
                      rho_requested=RHO, V_requested=float("nan"), kd_node_seed=nodes[0], kd_node_axis=nodes[1],
                      kd_node_left=nodes[2], kd_node_right=nodes[3], kd_node_parent=nodes[4])
    stream = RW.make_walk_random_stream(cfg, STREAM_SEED, ensemble=ens)
    Y, escaped, inside_trace, _, _ = RW._walk_cpu(ens, PP, cfg, random_stream=stream)
    assert not escaped.any(), "choose a case without escapes (their Y is path-specific)"
    cos, sin = RW._reduce_cpu(Y, *(np.asarray(c) for c in COLS))
    print(f"{len(seeds)} seeds, n_steps {cfg.n_steps}, n_grid {cfg.n_grid}, steps_per_h {cfg.steps_per_h}, "
          f"inside fraction {inside_trace.mean():.3f}")
    np.savez_compressed(
        OUT, seeds=seeds, annulus=annulus, kd_seed=nodes[0], kd_axis=nodes[1], kd_left=nodes[2],
        kd_right=nodes[3], kd_parent=nodes[4],
        scalars=np.array([SIDE, lo, lo + source_side, PP, D0, TS, T_MAX, H, cfg.sigma, float(cfg.n_steps),
                          float(cfg.steps_per_h), float(cfg.n_grid), float(N_WALKERS), float(STREAM_SEED), RHO, VI]),
        stream_sha256=np.array(stream_digest(stream)), Y=Y, inside_trace=inside_trace, escaped=escaped,
        cols_jd=np.asarray(COLS[0]), cols_jD=np.asarray(COLS[1]), cols_js=np.asarray(COLS[2]),
        cols_pc=np.asarray(COLS[3]), cos=cos, sin=sin)
    print(f"wrote {OUT}")


def time_cpu_walks() -> None:
    """With --time: the reference's _walk_cpu beside this repository's
    restatement (oracle.walk_cpu_reference, the benchmark's CPU baseline) on a
    larger seeded case, same stream; prints both rates and whether Y agrees."""
    import time

    sys.path.insert(1, str(Path(__file__).resolve().parents[2]))
    from oracle import oracle as O

    cfg = SimConfig(D0=2.0, ts=0.02, T_max_ms=4.0, h_ms=0.2, n_walkers=400, small_deltas=[0.4], big_deltas=[1.2],
                    b_values=[0.0])
    rho, side, vi = 5.0e5, 60.0, 0.6
    seeds, _, nn, _ = RE._make_population(RE._rho_um3(rho), side, cfg, np.random.default_rng(3))
    annulus = np.minimum(RE.alpha_star_from_vi(vi, rho), cfg.kappa * nn / 2.0)
    nodes = RE._build_gpu_kdtree(seeds)
    src = cfg.source_fraction * side
    lo = 0.5 * (side - src)
    ens = RE.Ensemble(seeds=seeds, annulus=annulus, rho=rho, V=float("nan"), vi=float("nan"), alpha_star=0.0,
                      L=side, source_lo=lo, source_hi=lo + src, mean_AV=1.0, geometry=None, rho_requested=rho,
                      V_requested=float("nan"), kd_node_seed=nodes[0], kd_node_axis=nodes[1], kd_node_left=nodes[2],
                      kd_node_right=nodes[3], kd_node_parent=nodes[4])
    stream = RW.make_walk_random_stream(cfg, 9, ensemble=ens)
    t = time.perf_counter()
    Y, _, itr, _, _ = RW._walk_cpu(ens, 0.02, cfg, random_stream=stream)
    t_ref = time.perf_counter() - t
    t = time.perf_counter()
    Y2, itr2, _ = O.walk_cpu_reference(seeds, annulus, side, lo, lo + src, cfg.n_walkers, cfg.n_steps,
                                       cfg.steps_per_h, cfg.n_grid, cfg.sigma, cfg.ts, 0.02,
                                       stream=(stream.initial_positions, stream.increments,
                                               stream.acceptance_uniforms))
    t_port = time.perf_counter() - t
    ws = cfg.n_walkers * cfg.n_steps
    print(f"{len(seeds)} seeds, {ws} walker-steps: reference _walk_cpu {t_ref:.2f} s ({ws / t_ref:.3e}/s), "
          f"restatement {t_port:.2f} s ({ws / t_port:.3e}/s); Y equal {np.array_equal(Y, Y2)}, "
          f"traces equal {np.array_equal(itr, itr2)}")


if __name__ == "__main__":
    if "--time" in sys.argv:
        time_cpu_walks()
    else:
        main()
