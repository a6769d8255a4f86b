"""Host-side logic on the CPU: reference known-answer tests, geometry
preparation, the column grid, the library format and the C ABI exports.

KAT sources (reference tree): tests/physics_audit/test_reference_path.py,
test_geometry_diagnostics.py, test_p0_remediation.py,
test_v5_library_schema.py.
"""

from __future__ import annotations

import math
import re
from pathlib import Path

import numpy as np
import pytest
from scipy.spatial import cKDTree

from custom_madi_amd import engine, library, signal
from custom_madi_amd.config import D0_UM2_MS, ENSEMBLE_MEAN_SUBSET_B_VALUES_S_MM2, GAMMA_RAD, SimConfig
from custom_madi_amd.ensemble import (
    _make_population, alpha_star_from_vi, build_kdtree, geometry_arrays, governing_mean_A_over_V,
    si_domain_side_um, create_dummy_ensemble,
)
from custom_madi_amd.walker_gpu import (
    ReducedResult, _membrane_transition_counts, _merge_results, kio_to_pp, pp_to_kio, pp_to_kio_eq5,
)
from oracle import oracle as O
from tests import _geom

ROOT = Path(__file__).resolve().parents[1]


def _s7a(x):
    return -0.570444 * x ** 3 + 10.6047 * x ** 2 - 5.81457 * x + 1.0


# ----------------------------------------------------------- reference KATs

def test_constants_and_step_convention():
    cfg = SimConfig()
    assert D0_UM2_MS == 3.0 and cfg.ts == 1e-3
    assert np.isclose(cfg.sigma, math.sqrt(2.0 * 3.0 * 1e-3))
    assert np.isclose(cfg.ls_rms, 0.1341640786499874)
    assert np.isclose(GAMMA_RAD, 2.675222e8)


def test_eq5_round_trip_and_rejection():
    cfg = SimConfig()
    pp = kio_to_pp(12.0, 0.5, cfg)
    assert np.isclose(pp_to_kio(pp, 0.5, cfg), 12.0)
    assert np.isclose(pp, (12.0 / 1000.0) / (math.sqrt(3.0 / 0.006) * 0.5))
    assert pp_to_kio_eq5(0.005, 0.5, cfg) == pytest.approx(0.005 * np.sqrt(3.0 / 0.006) * 0.5 * 1000.0)
    with pytest.raises(ValueError, match=r"\[0, 1\]"):
        kio_to_pp(1.0e7, 0.5, cfg)


def test_gradient_round_trip():
    g = signal.G_from_b(4000.0, 20.0, 50.0)
    b = (GAMMA_RAD * g * 20.0 * 1e-3) ** 2 * ((50.0 - 20.0 / 3.0) * 1e-3) / 1e6
    assert np.isclose(b, 4000.0)
    assert engine.phase_coef(4000.0, 20.0, 50.0) == pytest.approx(GAMMA_RAD * g * 1e-9, rel=1e-14)


@pytest.mark.parametrize("vi", [0.40, 0.784, 0.90, 0.99])
def test_s7a_inversion(vi):
    a = alpha_star_from_vi(vi, 5.0e5)
    x = (5.0e5 / 1.0e9) ** (1.0 / 3.0) * a
    assert _s7a(x) == pytest.approx(vi, abs=2e-12)
    if vi == pytest.approx(0.784):
        assert a == pytest.approx(0.5048514755, abs=2e-8)


def test_s7a_high_vi_distinct():
    vals = np.array([alpha_star_from_vi(v, 5.0e5) for v in (0.80, 0.85, 0.90, 0.94, 0.99)])
    assert np.all(np.diff(vals) < 0.0)


def test_governing_reference_interpolation_and_scaling(uncertified_reference):
    strict = SimConfig(geometry_reference_path=str(uncertified_reference), small_deltas=[1.0], big_deltas=[1.0],
                       T_max_ms=2.0)
    with pytest.raises(RuntimeError, match="not SI-certified"):
        governing_mean_A_over_V(0.80, 5.0e5, strict)
    cfg = SimConfig(geometry_reference_path=str(uncertified_reference), allow_uncertified_geometry_reference=True)
    low, ref = governing_mean_A_over_V(0.80, 1.0e5, cfg)
    high, _ = governing_mean_A_over_V(0.80, 8.0e5, cfg)
    assert low == pytest.approx((4.0 - 2.0 * 0.80) * (1.0e5 / 1.0e9) ** (1.0 / 3.0), rel=1e-12)
    assert high / low == pytest.approx(2.0, rel=1e-12)


def test_certified_table_value_at_c1():
    av, _ = governing_mean_A_over_V(0.6, 5.0e5, SimConfig())
    assert av == pytest.approx(0.6296, abs=5e-4)


def test_s8_domain_sides():
    cfg = SimConfig(T_max_ms=128.0)
    assert si_domain_side_um(5.0e5, cfg) == pytest.approx(554.2562584, rel=1e-9)
    assert si_domain_side_um(1.0e7, cfg) == pytest.approx(430.8869380063766)
    assert create_dummy_ensemble(cfg).L == pytest.approx(554.2562584220407)


def test_membrane_truth_table():
    m = _membrane_transition_counts(np.array([1, 1, 1, 1, 1]), np.array([True, False, True, True, False]),
                                    np.array([1, 1, 2, 1, 1]), np.array([False, True, True, True, False]))
    assert np.array_equal(m, [1, 1, 2, 0, 0])


def test_merge_keeps_ensemble_order():
    def red(c, s):
        return ReducedResult(cos_sum=c, sin_sum=s, n_walkers=2, n_escaped=0,
                             occupancy_counts=np.zeros(2, dtype=np.int64), ensemble_cos_means=[c / 6.0],
                             ensemble_sin_means=[s / 6.0])
    merged = _merge_results([red(np.array([6.0, 3.0]), np.array([0.0, 1.5])),
                             red(np.array([3.0, 0.0]), np.array([1.5, 3.0]))])
    cfg = SimConfig(small_deltas=[1.0], big_deltas=[1.0], b_values=[0.0, 500.0])
    payload = signal._assemble(merged, signal.build_columns(cfg))
    assert np.array_equal(payload["ensemble_S"].reshape(2, -1), [[1.0, 0.5], [0.5, 0.0]])
    assert np.array_equal(payload["ensemble_S_imag"].reshape(2, -1), [[0.0, 0.25], [0.25, 0.5]])


def test_columns_pair_major_then_b():
    cfg = SimConfig(small_deltas=[1.0, 2.0], big_deltas=[3.0, 4.0], b_values=[0.0, 500.0, 1000.0], T_max_ms=8.0)
    cols = signal.build_columns(cfg)
    assert cols.n_pairs == 4 and cols.n_b == 3
    assert cols.delta_pairs[1] == (1.0, 4.0)
    assert (cols.j_delta[3], cols.j_Delta[3], cols.j_sum[3]) == (1, 4, 5)
    assert cols.phase_coef[0] == 0.0
    assert cols.phase_coef[5] == pytest.approx(GAMMA_RAD * signal.G_from_b(1000.0, 1.0, 4.0) * 1e-9)


def test_fused_spec_groups_reference_columns():
    cfg = SimConfig(small_deltas=[6.0], big_deltas=[15.0, 25.0], b_values=[0.0, 1000.0, 2500.0], T_max_ms=31.0)
    cols = signal.build_columns(cfg)
    spec, cp, cb = engine.spec_from_columns(cols.j_delta, cols.j_Delta, cols.j_sum, cols.phase_coef)
    assert spec.shape == (2, 3, 3)
    assert np.array_equal(spec.pc[cp, cb], cols.phase_coef)
    assert np.array_equal(spec.j_Delta, [15, 25])


# --------------------------------------------------------------- geometry

def test_kdtree_structure_and_median_splits():
    pts = np.random.default_rng(3).uniform(0, 100, size=(1500, 3))
    seed, axis, left, right, parent = build_kdtree(pts)
    n = len(pts)
    assert np.array_equal(np.sort(seed), np.arange(n))
    assert parent[0] == -1
    for node in range(n):
        for child in (left[node], right[node]):
            if child >= 0:
                assert parent[child] == node
    # pre-order: the left child of a node with a left subtree is node + 1
    has_left = left >= 0
    assert np.array_equal(left[has_left], np.nonzero(has_left)[0] + 1)

    def subtree(node):
        out, stack = [], [node]
        while stack:
            k = stack.pop()
            if k >= 0:
                out.append(seed[k])
                stack += [left[k], right[k]]
        return np.array(out)

    for node in (0, 1, left[1], right[0]):
        ax = axis[node]
        split = pts[seed[node], ax]
        lo = subtree(left[node])
        hi = subtree(right[node])
        assert np.all(pts[lo, ax] <= split) and np.all(pts[hi, ax] >= split)
        size = 1 + len(lo) + len(hi)
        assert len(lo) == size // 2
        sub = pts[np.concatenate(([seed[node]], lo, hi))]
        assert ax == int(np.argmax(sub.max(axis=0) - sub.min(axis=0)))


def test_population_certificate_and_annulus(uncertified_reference):
    cfg = SimConfig(L=60.0, geometry_reference_path=str(uncertified_reference),
                    allow_uncertified_geometry_reference=True)
    rng = np.random.default_rng(5)
    seeds, nn, cert = _make_population(5.0e5 * 1e-9, 60.0, cfg, rng)
    assert cert.s10_pass and cert.s11_pass
    # the native nearest-neighbour distances are bitwise those of a KD tree
    d2 = cKDTree(seeds).query(seeds, k=2)[0][:, 1]
    assert np.array_equal(nn, d2)
    # the S10 bound is the reference's plane-by-plane grid query
    # (madi/ensemble.py:566-582), restated with a KD tree
    tree = cKDTree(seeds)
    m = int(np.ceil(60.0 / cfg.population_certification_spacing_um))
    h = 60.0 / m
    coords = (np.arange(m) + 0.5) * h
    yy, zz = np.meshgrid(coords, coords, indexing="ij")
    worst = max(float(np.max(tree.query(np.column_stack((np.full(yy.size, x), yy.ravel(), zz.ravel())), k=1)[0]))
                for x in coords)
    assert cert.s10_distance_upper_um == worst + 0.5 * np.sqrt(3.0) * h
    assert cert.certification_spacing_um == h


def _graph_classify_np(arr, pts):
    """FP64 numpy restatement of the Delaunay-graph classifier (greedy walk + neighbour facets)."""
    tree = cKDTree(arr.cell_xyz, boxsize=arr.L if arr.periodic else None)
    out_c = np.empty(len(pts), np.int64)
    out_i = np.empty(len(pts), bool)
    wrap = np.mod(pts, arr.L) if arr.periodic else pts
    start = tree.query(wrap)[1]
    for i, p in enumerate(wrap):
        a = int(start[i])
        q = p - arr.cell_xyz[a]
        while True:
            lo, hi = arr.nb_offsets[a], arr.nb_offsets[a + 1]
            e = arr.nb_vec[lo:hi]
            d2 = np.sum((q - e) ** 2, axis=1)
            da = q @ q
            k = int(np.argmin(d2))
            if d2[k] < da:
                q = q - e[k]
                a = int(arr.nb_ids[lo + k])
                continue
            sep = np.linalg.norm(e, axis=1)
            out_c[i] = a
            out_i[i] = bool(np.all((d2 - da) / (2 * sep) >= arr.cell_annulus[a]))
            break
    return out_c, out_i


@pytest.mark.parametrize("periodic", [False, True])
def test_delaunay_graph_classification_equals_brute_force(periodic):
    ens = _geom.periodic(n_cells=80, seed=2) if periodic else _geom.cellular(L=40.0, seed=4)
    arr = geometry_arrays(ens)
    # graph symmetry: every edge a->k has k->a with the opposite vector
    edges = {}
    for a in range(len(arr.cell_xyz)):
        for k in range(arr.nb_offsets[a], arr.nb_offsets[a + 1]):
            edges[(a, int(arr.nb_ids[k]), tuple(np.round(arr.nb_vec[k], 9)))] = True
    for (a, b, v) in list(edges)[:2000]:
        assert (b, a, tuple(np.round(-np.asarray(v), 9))) in edges
    pts = np.random.default_rng(9).uniform(0, ens.L, size=(1500, 3))
    c, i = _graph_classify_np(arr, pts)
    if periodic:
        bc, bi = O.periodic_brute(ens.seeds, ens.annulus, ens.L, pts)
    else:
        bc, bi = O.classify_brute(ens.seeds, ens.annulus, pts)
    assert np.array_equal(c, bc) and np.array_equal(i, bi)


# ---------------------------------------------------------------- library

def _diag_cfg(n_ens):
    return SimConfig(n_walkers=2, n_ensembles=n_ens, small_deltas=[5.0, 7.0, 10.0, 12.0, 15.0, 20.0, 25.0, 30.0],
                     big_deltas=[15.0, 25.0, 30.0, 36.0, 40.0, 50.0, 60.0, 80.0],
                     b_values=list(ENSEMBLE_MEAN_SUBSET_B_VALUES_S_MM2))


def test_v5_variance_is_sample_variance_and_subset_indexed():
    cfg = _diag_cfg(3)
    cols = signal.build_columns(cfg)
    real = np.stack([np.full((cols.n_pairs, cols.n_b), v) for v in (0.2, 0.4, 0.8)])
    res = {"ensemble_S": real, "ensemble_S_imag": np.zeros_like(real), "S_imag": np.zeros((cols.n_pairs, cols.n_b))}
    imag, var, sub, check = library._v5_diagnostics_from_result(res, cols, cfg)
    assert imag.dtype == var.dtype == sub.dtype == np.float32
    assert np.allclose(var, np.var([0.2, 0.4, 0.8], ddof=1))
    assert sub.shape == (3, library.ENSEMBLE_MEAN_SUBSET_N_COLUMNS)
    idx = library.ensemble_mean_subset_column_indices(cols)
    assert np.allclose(sub, real.reshape(3, -1)[:, idx])
    assert check["max_abs_standardized_deviation"] == 0.0


def test_library_save_load_round_trip(tmp_path):
    cfg = SimConfig(small_deltas=[1.0], big_deltas=[2.0, 3.0], b_values=[0.0, 500.0], T_max_ms=4.0)
    cols = signal.build_columns(cfg)
    lib = [library.LibraryEntry(10.0, 5.0e5, 1.4, np.asarray([1.0, 0.7, 1.0, 0.6]), vi=0.70, kio_nominal=10.0,
                                pp=0.01, kio_analytic_eq5=10.0)]
    out = tmp_path / "lib.npz"
    library._save_library(lib, str(out), cfg=cfg, columns=cols, build_seed=3)
    back = library.load_library(str(out))
    meta = library.load_library_meta(str(out))
    assert len(back) == 1 and np.array_equal(back[0].vector, lib[0].vector)
    assert back[0].pp == 0.01 and back[0].vi == 0.70
    assert meta["format"] == "v2" and meta["delta_pairs"] == [(1.0, 2.0), (1.0, 3.0)]
    assert meta["library_schema"] == "madi-library-v4"
    with np.load(out, allow_pickle=False) as d:
        assert "kio_measured_se" not in d.files


def test_log_grid_weights_and_free_water():
    grid = library.make_remediation_log_grid(n_rho=9, n_V=10, kios=np.asarray([0.0, 1.0, 5.0, 20.0, 80.0, 130.0]))
    triplets, weights = grid.triplets_and_weights()
    assert triplets[0] == (0.0, 0.0, 0.0) and weights[("free_water", 0.0, 0.0)] > 0.0
    cell = np.asarray([t for t in triplets if t[1] > 0.0])
    vi = cell[:, 1] * cell[:, 2] * 1e-6
    assert np.all((vi >= 0.40) & (vi <= 0.99))
    assert np.allclose(library._cell_widths(np.asarray([0.0, 1.0, 5.0]), 0.0, 5.0), [0.5, 2.5, 2.0])


# ------------------------------------------------------------------ C ABI

def test_production_grid_points_and_benchmark_coordinates():
    """The production log grid's rho and V at indices the reference's
    replicate subset lists with their values
    (data/madi_v5_replicate_entry_subset.json:17-37), and the benchmark's
    density coordinates derived from it (analysis/v5_runtime_benchmark.py:101-123:
    rho index 0 / 21 / last, the admissible V whose v_i is nearest in log to
    the geometric centre of the v_i range)."""
    import bench
    from custom_madi_amd.library import make_remediation_log_grid
    grid = make_remediation_log_grid()
    assert (len(grid.rhos), len(grid.Vs), grid.vi_min, grid.vi_max) == (64, 64, 0.4, 0.99)
    for ri, vj, rho, V in [(0, 53, 10000.0, 41.52664275536708), (3, 54, 13894.954943731389, 48.59562131612258),
                           (9, 47, 26826.957952797275, 16.169810020515047), (18, 41, 71968.56730011514,
                                                                            6.2962652107424715),
                           (21, 42, 100000.0, 7.368062997280774), (42, 29, 1000000.0, 0.9546390696098913),
                           (63, 14, 10000000.0, 0.09032010701366826)]:
        assert grid.rhos[ri] == pytest.approx(rho, rel=1e-14) and grid.Vs[vj] == pytest.approx(V, rel=1e-14)
    centre = np.sqrt(0.4 * 0.99)
    for which, ri in (("low", 0), ("center", 21), ("high", 63)):
        rho, V, vi = bench.production_coordinate(which)
        assert rho == grid.rhos[ri] and vi == pytest.approx(rho * V * 1e-6)
        admissible = [float(v) for v in grid.Vs if 0.4 <= rho * float(v) * 1e-6 <= 0.99]
        assert V in admissible
        assert all(abs(np.log(vi / centre)) <= abs(np.log(rho * v * 1e-6 / centre)) for v in admissible)
    assert bench.production_coordinate("center")[1] == grid.Vs[41]  # 6.296 pL, v_i 0.630


def test_hip_library_exports_every_declared_symbol():
    import ctypes
    from custom_madi_amd import _hip
    header = (ROOT / "include" / "madi_hip.h").read_text()
    declared = set(re.findall(r"^(?:int|const char\*)\s+(madi_\w+)\(", header, flags=re.M))
    assert {"madi_classify", "madi_walk_f64", "madi_walk_fused", "madi_rng_init"} <= declared
    lib = ctypes.CDLL(str(_hip.LIB_PATH))
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert set(_hip.SIGNATURES) >= declared


def test_y_walk_and_reduce_abi_argument_checks():
    """madi_walk_y_f32's workspace size and argument errors, and the column
    reduce's limits, which return before any device call."""
    import ctypes as C
    from custom_madi_amd import _hip
    lib = _hip.load_library()
    b0, b1 = C.c_int64(), C.c_int64()
    assert lib.madi_walk_y_f32_workspace(0, C.byref(b0)) == _hip.MADI_OK and b0.value > 0
    assert lib.madi_walk_y_f32_workspace(1000, C.byref(b1)) == _hip.MADI_OK and b1.value > b0.value
    assert lib.madi_walk_y_f32_workspace(-1, C.byref(b0)) == _hip.MADI_ERR_ARG
    assert lib.madi_walk_y_f32_workspace(10, None) == _hip.MADI_ERR_ARG
    p = _hip.WalkParams()
    p.n_walkers, p.n_steps, p.steps_per_h, p.n_grid = 10, 20, 10, 3
    fake = 4096  # never dereferenced: each call below returns before touching memory
    assert lib.madi_walk_y_f32(None, C.byref(p), fake, b1.value, fake, fake, fake, None, None) == _hip.MADI_ERR_ARG
    assert lib.madi_walk_y_f32(fake, C.byref(p), fake, b1.value, None, fake, fake, None, None) == _hip.MADI_ERR_ARG
    assert lib.madi_walk_y_f32(fake, C.byref(p), fake, 16, fake, fake, fake, None, None) == _hip.MADI_ERR_ARG
    assert "smaller than" in _hip.last_error()
    cs = C.c_double()
    assert lib.madi_reduce_f64(fake, 10, 3, fake, fake, fake, fake, _hip.MADI_MAX_REDUCE_COLS + 1, fake, fake,
                               None) == _hip.MADI_ERR_ARG
    assert lib.madi_reduce_f64(fake, 10, 3, fake, fake, fake, fake, 0, C.byref(cs), C.byref(cs), None) == _hip.MADI_OK


def test_walk_f64_multi_abi_argument_checks():
    """madi_walk_f64_multi's argument errors (no geometry, walk counts outside
    1..MADI_MAX_MULTI_WALKS, missing or invalid p_p, missing outputs) and
    madi_count_inside's, which return before any device call or geometry
    access."""
    import ctypes as C
    from custom_madi_amd import _hip
    lib = _hip.load_library()
    p = _hip.WalkParams()
    p.n_walkers, p.n_steps, p.steps_per_h, p.n_grid, p.pp = 10, 20, 10, 3, 0.5
    fake = 4096  # never dereferenced: each call below returns before touching memory
    K = _hip.MADI_MAX_MULTI_WALKS

    def call(g, k, pps, Y=fake):
        arr = None if pps is None else C.cast((C.c_double * len(pps))(*pps), C.c_void_p)
        return lib.madi_walk_f64_multi(g, C.byref(p), k, arr, Y, fake, fake, fake, None, None)

    assert call(None, 2, [0.1, 0.2]) == _hip.MADI_ERR_ARG
    assert call(fake, 0, [0.1]) == _hip.MADI_ERR_ARG
    assert call(fake, K + 1, [0.1] * (K + 1)) == _hip.MADI_ERR_ARG
    assert f"1..{K}" in _hip.last_error()
    assert call(fake, 2, None) == _hip.MADI_ERR_ARG
    assert call(fake, 2, [0.1, 0.2], Y=None) == _hip.MADI_ERR_ARG
    assert call(fake, 2, [0.1, float("nan")]) == _hip.MADI_ERR_ARG
    assert "walk 1" in _hip.last_error()
    assert call(fake, 3, [0.1, 0.2, 1.5]) == _hip.MADI_ERR_ARG
    assert call(fake, 2, [-0.1, 0.2]) == _hip.MADI_ERR_ARG
    p.pp = 2.0  # the params' own p_p is checked too
    assert call(fake, 2, [0.1, 0.2]) == _hip.MADI_ERR_ARG
    # the device occupancy count of stacked walks
    assert lib.madi_count_inside(fake, -1, 10, 3, fake, None) == _hip.MADI_ERR_ARG
    assert lib.madi_count_inside(fake, 2, 10, 3, None, None) == _hip.MADI_ERR_ARG
    assert lib.madi_count_inside(None, 2, 10, 3, fake, None) == _hip.MADI_ERR_ARG
    assert lib.madi_count_inside(None, 0, 10, 3, None, None) == _hip.MADI_OK  # nothing to count


def test_matcher_and_finalize_abi_argument_checks():
    """Workspace sizes and argument errors of the matcher and tile-sum entry
    points, which return before any device call."""
    import ctypes as C
    from custom_madi_amd import _hip
    lib = _hip.load_library()
    b = C.c_int64()
    assert lib.madi_map_workspace(10, 3, C.byref(b)) == _hip.MADI_OK and b.value == 13 * 8
    assert lib.madi_map_workspace(0, 0, C.byref(b)) == _hip.MADI_OK and b.value == 8
    assert lib.madi_map_workspace(-1, 3, C.byref(b)) == _hip.MADI_ERR_ARG
    fake = 4096  # never dereferenced: each call below returns before touching memory
    args = dict(meas=fake, n_vox=10, lib=fake, n_lib=3, nf=4, fit=0, ws=fake, wsb=13 * 8, idx=fake, res=fake, s0=None)

    def call(**kw):
        a = {**args, **kw}
        return lib.madi_map_match(a["meas"], a["n_vox"], a["lib"], a["n_lib"], a["nf"], a["fit"], a["ws"], a["wsb"],
                                  a["idx"], a["res"], a["s0"], None)

    assert call(n_lib=0) == _hip.MADI_ERR_ARG
    assert call(fit=2) == _hip.MADI_ERR_ARG
    assert call(fit=1) == _hip.MADI_ERR_ARG            # free S0 needs the S0 output
    assert call(ws=None) == _hip.MADI_ERR_ARG
    assert call(ws=fake + 4) == _hip.MADI_ERR_ARG      # 8-byte alignment
    assert call(nf=-1) == _hip.MADI_ERR_ARG
    assert call(wsb=12 * 8) == _hip.MADI_ERR_ARG
    assert "smaller than" in _hip.last_error()
    assert call(n_vox=0, idx=None, res=None, wsb=8) == _hip.MADI_ERR_ARG  # the library norms still need 3 x 8 bytes
    assert call(n_vox=0, idx=None, res=None, wsb=3 * 8) == _hip.MADI_OK  # nothing to match
    plan = _hip.PhasePlan()
    plan.n_pairs, plan.n_b, plan.n_dirs = 4, 5, 24
    assert lib.madi_tiles_finalize_workspace(C.byref(plan), 100, C.byref(b)) == _hip.MADI_OK
    assert b.value == 4 * 480 * 2 * 8                  # ceil(100 / 32) chunks x 480 columns x (cos, sin)
    plan.n_pairs = 9
    assert lib.madi_tiles_finalize_workspace(C.byref(plan), 100, C.byref(b)) == _hip.MADI_ERR_ARG
    plan.n_pairs = 4
    assert lib.madi_tiles_finalize_ws(C.byref(plan), fake, 100, fake, fake, fake, 4 * 480 * 2 * 8 - 8,
                                      None) == _hip.MADI_ERR_ARG


def test_ctypes_structs_match_the_header(tmp_path):
    """Every ctypes mirror of a header struct has the C compiler's size and field offsets."""
    import ctypes
    import shutil
    import subprocess
    from custom_madi_amd import _hip
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    pairs = {"madi_device_info": _hip.DeviceInfo, "madi_geom_desc": _hip.GeomDesc, "madi_walk_params": _hip.WalkParams,
             "madi_phase_plan": _hip.PhasePlan, "madi_work": _hip.Work, "madi_tune": _hip.Tune}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "madi_hip.h"', "int main(void) {"]
    for cname, py in pairs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for ln in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        name, key, val = ln.split()
        got[(name, key)] = int(val)
    for cname, py in pairs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


def test_coincident_seeds_are_a_classifier_failure_before_the_walk():
    """A seed on top of another has no Delaunay facets, so no label is defined
    near it: the reference's classifier failure (madi/walker_gpu.py:938-940),
    raised when the device geometry is prepared, in periodic and
    non-periodic boxes alike; the failing entry fails alone in the plain and
    served forms."""
    from custom_madi_amd import batch, serve
    from custom_madi_amd.ensemble import build_kdtree, geometry_arrays
    for ens in (_geom.periodic(n_cells=64, seed=0), _geom.cellular(L=40.0, seed=2)):
        ens.seeds[5] = ens.seeds[4]
        (ens.kd_node_seed, ens.kd_node_axis, ens.kd_node_left, ens.kd_node_right,
         ens.kd_node_parent) = build_kdtree(ens.seeds)
        with pytest.raises(RuntimeError, match="classification failed"):
            geometry_arrays(ens)

    req = batch.library_requests(1, seed=3, walkers=(1000, 1000), cells=(60, 60), T_max_ms=12.0)[0]

    def bad_prepare(r):
        raise RuntimeError("GPU full-facet classification failed: seeds coincide")

    spec = batch.acquisition_spec(n_dirs=3, delta=2.0, Deltas=(5.0,), b_values=(0.0, 1000.0))
    failed = []
    assert batch.run_plain([req], spec, prepare=bad_prepare, failed=failed) == []
    assert len(failed) == 1 and failed[0].key == req.key and "classification failed" in failed[0].error
    with pytest.raises(RuntimeError, match="classification failed"):
        batch.run_plain([req], spec, prepare=bad_prepare)

    class _Sch:  # a scheduler stand-in: the request never reaches a device
        spec = batch.acquisition_spec(n_dirs=3, delta=2.0, Deltas=(5.0,), b_values=(0.0, 1000.0))

    srv = serve.Server(_Sch(), prepare=bad_prepare)
    srv.submit(req)
    assert len(srv.failed) == 1 and not srv.results and not srv.preparing


def test_reduce_columns_checks_its_column_lists_before_any_device_call():
    """Column lists of unequal length or grid indices outside Y are refused on
    the host (the kernel would read past them)."""
    from custom_madi_amd import engine
    Y = np.zeros((4, 5, 3))
    with pytest.raises(ValueError, match="one entry per column"):
        engine.reduce_columns(Y, [1, 1, 2], [2, 3, 3], [3, 4, 4], [0.0, 0.05, 0.3, 0.1])
    with pytest.raises(ValueError, match="outside"):
        engine.reduce_columns(Y, [1], [2], [5], [0.1])
    with pytest.raises(ValueError, match="outside"):
        engine.reduce_columns(Y, [-1], [2], [3], [0.1])
