"""MAP matcher on the CPU: the oracle restatement of the reference's GPU
matchers (madi/fitters_gpu.py:96-154) pinned against the reference's own
known-answer tests and the numpy form of its CPU path, and the host logic
of the matcher API (column resolution, candidate filtering)."""

from __future__ import annotations

import warnings

import numpy as np
import pytest

from custom_madi_amd.library import LibraryEntry
from custom_madi_amd.match import _build_candidate_lib_matrix, _grid_columns, resolve_grid_columns
from oracle import oracle as O


def _small_library():
    # tests/physics_audit/test_reference_path.py:56-62
    return [
        LibraryEntry(11.0, 1_000_000.0, 0.80, np.array([0.8, 0.5, 0.3])),
        LibraryEntry(22.0, 1_000_000.0, 0.81, np.array([0.8, 0.5, 0.3])),
        LibraryEntry(33.0, 1_000_000.0, 0.82, np.array([0.7, 0.4, 0.2])),
    ]


def test_reference_known_answer_first_minimum_wins():
    """tests/physics_audit/test_reference_path.py:65-88: an exact tie between
    entries 0 and 1 resolves to the lower index."""
    lib = _small_library()
    R = np.array([e.vector for e in lib])
    M = np.array([[0.8, 0.5, 0.3], [0.71, 0.39, 0.19]])
    idx, res, _ = O.map_match(M, R)
    assert [lib[i].kio for i in idx] == [11.0, 33.0]
    assert [lib[i].V for i in idx] == [0.80, 0.82]
    assert np.all(res >= -1e-14)
    # the same call twice gives the same bits
    idx2, res2, _ = O.map_match(M, R)
    assert np.array_equal(idx, idx2) and np.array_equal(res, res2)
    # a last-minimum scan would return entry 1 for the tied row: the test pins the rule
    d = np.sum((M[:, None, :] - R[None, :, :]) ** 2, axis=2)
    assert (len(R) - 1 - np.argmin(d[:, ::-1], axis=1))[0] == 1


def test_pair_major_b_major_column_order():
    # tests/physics_audit/test_reference_path.py:91-100
    cols = _grid_columns([(20.0, 50.0, 0.0), (20.0, 50.0, 1_000.0)], [(1.0, 1.0), (20.0, 50.0)],
                         [0.0, 500.0, 1_000.0], n_b=3)
    assert cols.tolist() == [3, 5]


def test_column_snaps_are_reported_and_bounded():
    pairs, bvals = [(20.0, 50.0), (10.0, 30.0)], [0.0, 1000.0, 2000.0]
    cols, ev = resolve_grid_columns([(20.0, 50.0, 1000.0), (10.5, 30.0, 2010.0), (10.5, 30.0, 2010.0)], pairs,
                                    bvals, n_b=3)
    assert cols.tolist() == [1, 5, 5]
    assert [e["axis"] for e in ev] == ["timing", "b"]  # one event per distinct snap, none for exact columns
    assert ev[0]["used"] == {"delta_ms": 10.0, "Delta_ms": 30.0}
    assert ev[1]["used"] == 2000.0 and ev[1]["abs_offset"] == pytest.approx(10.0)
    with pytest.raises(ValueError, match="timing snap tolerance"):
        resolve_grid_columns([(12.0, 30.0, 0.0)], pairs, bvals, n_b=3)
    with pytest.raises(ValueError, match="not within"):
        resolve_grid_columns([(20.0, 50.0, 1500.0)], pairs, bvals, n_b=3)
    with pytest.raises(ValueError, match="no stored b-values"):
        resolve_grid_columns([(20.0, 50.0, 0.0)], pairs, None, n_b=3)


def test_candidate_filter():
    vec = np.arange(6, dtype=float)
    lib = [LibraryEntry(1.0, 1e5, 5.0, vec),            # v_i 0.5
           LibraryEntry(2.0, 1e6, 0.96, vec + 1),       # v_i 0.96: above vi_max
           LibraryEntry(3.0, 2e6, 0.3, vec + 2),        # v_i 0.6, rho 2e6
           LibraryEntry(4.0, 0.0, 0.0, vec + 3, is_free_water=True)]
    args = ([(20.0, 50.0)], [0.0, 500.0, 1000.0, 1500.0, 2000.0, 2500.0], 6, [(20.0, 50.0, 500.0),
                                                                               (20.0, 50.0, 2500.0)])
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        mat, kio, rho, V = _build_candidate_lib_matrix(lib, *args[:3], 0.5, 0.95, None, args[3])
    assert kio.tolist() == [1.0, 3.0] and any("pass the filter" in str(x.message) for x in w)
    assert mat.tolist() == [[1.0, 5.0], [3.0, 7.0]]
    mat, kio, *_ = _build_candidate_lib_matrix(lib, *args[:3], 0.5, 0.95, 1e6, args[3])
    assert kio.tolist() == [1.0]
    mat, kio, *_ = _build_candidate_lib_matrix(lib, *args[:3], 0.5, 0.95, None, args[3], include_free_water=True)
    assert kio.tolist() == [1.0, 3.0, 4.0]
    with pytest.raises(ValueError, match="No library entries survive"):
        _build_candidate_lib_matrix(lib, *args[:3], 0.97, 0.99, None, args[3])


@pytest.mark.parametrize("fit_s0", [False, True])
@pytest.mark.parametrize("shape", [(1, 1, 1), (40, 7, 3), (300, 200, 20), (65, 130, 9)])
def test_oracle_agrees_with_the_reference_cpu_path(shape, fit_s0):
    """Same argmin as the BLAS-ordered CPU formula wherever its best residual
    leads the runner-up by more than the rounding of the two orders, and the
    same residuals within that rounding."""
    nv, nl, nf = shape
    rng = np.random.default_rng(nv * 1000 + nl + nf)
    R = rng.uniform(0.05, 1.0, (nl, nf))
    M = R[rng.integers(0, nl, nv)] * (rng.uniform(0.7, 1.3, (nv, 1)) if fit_s0 else 1.0)
    M = M + rng.normal(0.0, 0.02, M.shape)
    idx, res, s0 = O.map_match(M, R, fit_s0)
    ridx, rres, rs0 = O.map_match_cpu_reference(M, R, fit_s0)
    scale = np.sum(M * M, axis=1) + np.max(np.sum(R * R, axis=1))
    tol = 64 * np.finfo(float).eps * scale * max(nf, 1)
    assert np.all(np.abs(res - rres) <= tol)
    # margins of the runner-up under the CPU formula
    if fit_s0:
        MR = M @ R.T
        rr = np.maximum(np.sum(R * R, axis=1), 1e-30)
        full = np.where(MR / rr > 0, np.sum(M * M, axis=1)[:, None] - MR ** 2 / rr, np.inf)
    else:
        full = np.sum(M * M, axis=1)[:, None] + np.sum(R * R, axis=1)[None, :] - 2.0 * M @ R.T
    srt = np.sort(full, axis=1)
    clear = (srt[:, 1] - srt[:, 0] > 2 * tol) if nl > 1 else np.ones(nv, bool)
    assert np.array_equal(idx[clear], ridx[clear])
    assert clear.mean() > 0.9
    if fit_s0:
        assert np.allclose(s0[clear], rs0[clear], rtol=1e-9)


def test_exact_and_scaled_copies_have_zero_residual():
    """A voxel equal to a library row (fixed S0) or to 4x a row (free S0)
    matches that row with residual exactly 0 and S0 exactly 4; a sign error
    in the score (+2 m.r) would not pick the copy."""
    rng = np.random.default_rng(7)
    R = rng.uniform(0.1, 1.0, (50, 12))
    M = R[[3, 17, 42]].copy()
    idx, res, _ = O.map_match(M, R)
    assert idx.tolist() == [3, 17, 42] and np.all(res == 0.0)
    wrong = np.argmin(np.sum(R * R, axis=1)[None, :] + 2.0 * M @ R.T, axis=1)
    assert wrong.tolist() != [3, 17, 42]
    idx, res, s0 = O.map_match(4.0 * M, R, fit_s0=True)
    assert idx.tolist() == [3, 17, 42] and np.all(res == 0.0) and np.all(s0 == 4.0)


def test_free_s0_masks_negative_scales():
    rng = np.random.default_rng(8)
    R = rng.uniform(0.1, 1.0, (20, 5))
    M = -R[:3]  # every candidate has S0 < 0
    idx, res, s0 = O.map_match(M, R, fit_s0=True)
    assert idx.tolist() == [0, 0, 0] and np.all(np.isinf(res)) and np.all(s0 == 0.0)
    # one admissible row among negatives is found
    R2 = np.vstack([-R, R[5:6]])
    idx, res, s0 = O.map_match(R[5:6] * 2.0, R2, fit_s0=True)
    assert idx.tolist() == [20] and s0[0] == 2.0 and res[0] == 0.0
