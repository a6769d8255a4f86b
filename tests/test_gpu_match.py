"""MAP matcher on an MI355X: madi_map_match through the C-ABI, bitwise
against the oracle restatement of the reference's GPU matchers
(madi/fitters_gpu.py:96-154), and the matcher API on the reference's
known-answer test (tests/physics_audit/test_reference_path.py:65-88)."""

from __future__ import annotations

import sys

import numpy as np
import pytest

from custom_madi_amd.library import LibraryEntry
from custom_madi_amd.match import DeviceLibrary, match_voxels_batch, match_voxels_batch_fits0
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _same(a, b):
    """Bitwise equality of FP64 arrays (NaN and inf included)."""
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


@pytest.mark.parametrize("fit_s0", [False, True])
@pytest.mark.parametrize("shape", [(1, 1, 1), (7, 3, 3), (129, 65, 9), (300, 200, 20), (257, 130, 0), (130, 70, 37),
                                   (200, 129, 24), (150, 64, 25),
                                   (64, 1000, 480), (2000, 3001, 20)])
def test_kernel_equals_oracle_bitwise(shape, fit_s0):
    nv, nl, nf = shape
    say(f"map match shape={shape} fit_s0={fit_s0}")
    rng = np.random.default_rng(nv + 7 * nl + 13 * nf + int(fit_s0))
    R = rng.uniform(0.05, 1.0, (nl, nf))
    M = R[rng.integers(0, nl, nv)] * (rng.uniform(0.5, 2.0, (nv, 1)) if fit_s0 else 1.0)
    M = M + rng.normal(0.0, 0.05, M.shape)
    idx, res, s0 = DeviceLibrary(R).match(M, fit_s0)
    oidx, ores, os0 = O.map_match(M, R, fit_s0)
    assert np.array_equal(idx, oidx)
    assert _same(res, ores)
    if fit_s0:
        assert _same(s0, os0)


@pytest.mark.parametrize("fit_s0", [False, True])
def test_ties_nans_and_masked_rows(fit_s0):
    """Duplicate library rows (ties resolve to the first), NaN voxels, and
    (free S0) voxels whose every candidate has a negative scale."""
    say(f"map match edge cases fit_s0={fit_s0}")
    rng = np.random.default_rng(5)
    base = rng.uniform(0.1, 1.0, (90, 11))
    R = np.vstack([base, base[::-1], base[:7]])          # every row appears 2-3 times
    M = np.vstack([base[[0, 5, 89]], rng.uniform(0.1, 1.0, (200, 11)), np.full((2, 11), np.nan),
                   -base[:4]])
    idx, res, s0 = DeviceLibrary(R).match(M, fit_s0)
    oidx, ores, os0 = O.map_match(M, R, fit_s0)
    assert np.array_equal(idx, oidx) and _same(res, ores)
    assert idx[0] == 0 and idx[1] == 5 and idx[2] == 89
    assert idx[203] == 0 and idx[204] == 0
    if fit_s0:
        assert _same(s0, os0)
        assert np.all(np.isinf(res[-4:])) and np.all(idx[-4:] == 0) and np.all(s0[-4:] == 0.0)


def test_device_batches_and_workspace_reuse():
    """Two batches in a row through one workspace, the second smaller, equal
    the oracle each time."""
    rng = np.random.default_rng(9)
    R = rng.uniform(0.0, 1.0, (700, 20))
    lib = DeviceLibrary(R)
    ws = lib.workspace(1000)
    import torch
    for n in (1000, 333):
        M = rng.uniform(0.0, 1.0, (n, 20))
        idx, res, _ = lib.match_device(torch.from_numpy(M).to("cuda"), workspace=ws)
        oidx, ores, _ = O.map_match(M, R)
        assert np.array_equal(idx.cpu().numpy(), oidx) and _same(res.cpu().numpy(), ores)


def _small_library():
    # tests/physics_audit/test_reference_path.py:56-62
    return [
        LibraryEntry(11.0, 1_000_000.0, 0.80, np.array([0.8, 0.5, 0.3])),
        LibraryEntry(22.0, 1_000_000.0, 0.81, np.array([0.8, 0.5, 0.3])),
        LibraryEntry(33.0, 1_000_000.0, 0.82, np.array([0.7, 0.4, 0.2])),
    ]


def test_reference_matcher_known_answer_through_the_api():
    library = _small_library()
    triples = [(20.0, 50.0, b) for b in (500.0, 1_000.0, 1_500.0)]
    measured = np.array([[0.8, 0.5, 0.3], [0.71, 0.39, 0.19]])
    kw = dict(lib_delta_pairs=[(20.0, 50.0)], lib_b_values=[500.0, 1_000.0, 1_500.0], n_b=3, fit_triples=triples)
    first = match_voxels_batch(measured, library, **kw)
    second = match_voxels_batch(measured, library, **kw)
    for a, b in zip(first, second):
        assert np.array_equal(a, b)
    kio, rho, volume, residual = first
    assert kio.tolist() == [11.0, 33.0]
    assert rho.tolist() == [1_000_000.0, 1_000_000.0]
    assert volume.tolist() == [0.80, 0.82]
    assert np.all(residual >= -1e-14)
    with pytest.raises(ValueError, match="no CPU path"):
        match_voxels_batch(measured, library, use_gpu=False, **kw)


def test_api_log_space_and_free_s0_equal_the_oracle():
    rng = np.random.default_rng(3)
    pairs, bvals = [(20.0, 50.0), (10.0, 30.0)], [0.0, 1000.0, 2000.0, 3000.0]
    lib = [LibraryEntry(float(k), 1e6, 0.5 + 0.4 * rng.random(), np.exp(-rng.uniform(0.0, 2.0, 8)))
           for k in range(120)]
    triples = [(20.0, 50.0, 1000.0), (20.0, 50.0, 3000.0), (10.0, 30.0, 2000.0), (10.0, 30.0, 0.0)]
    cols = [1, 3, 6, 4]
    cand = [e for e in lib if 0.5 <= e.realised_vi <= 0.95]
    R = np.array([e.vector for e in cand])[:, cols]
    S = R[rng.integers(0, len(R), 500)] * rng.uniform(0.9, 1.1, (500, 4))
    kio, _, _, res = match_voxels_batch(S, lib, pairs, bvals, 4, triples, log_space=True)
    oidx, ores, _ = O.map_match(np.log(np.clip(S, 1e-3, 1.0)), np.log(np.clip(R, 1e-3, 1.0)))
    assert np.array_equal(kio, np.array([e.kio for e in cand])[oidx]) and _same(res, ores)
    raw = S * 1234.5
    kio, _, _, res, s0 = match_voxels_batch_fits0(raw, lib, pairs, bvals, 4, triples)
    oidx, ores, os0 = O.map_match(raw, R, fit_s0=True)
    assert np.array_equal(kio, np.array([e.kio for e in cand])[oidx]) and _same(res, ores) and _same(s0, os0)


@pytest.mark.parametrize("fit_s0", [False, True])
@pytest.mark.parametrize("nf", [6, 30])
def test_many_exact_ties_resolve_to_the_first_row(nf, fit_s0):
    """Quantised inputs (values k/2, k in 0..2) make many different rows tie
    exactly, across threads, waves and library tiles of a workgroup; the
    kernel (LDS-resident variant for 6 columns, staged for 30) must return
    the oracle's first minimum for every voxel."""
    say(f"map match quantised ties nf={nf} fit_s0={fit_s0}")
    rng = np.random.default_rng(100 + nf + int(fit_s0))
    R = rng.integers(0, 3, (700, nf)) * 0.5
    M = rng.integers(0, 3, (400, nf)) * 0.5
    idx, res, s0 = DeviceLibrary(R).match(M, fit_s0)
    oidx, ores, os0 = O.map_match(M, R, fit_s0)
    assert np.array_equal(idx, oidx) and _same(res, ores)
    if fit_s0:
        assert _same(s0, os0)
    # ties really occur: several rows share the best score (65 % of voxels at
    # 6 columns, 18 % at 30)
    if not fit_s0:
        full = np.sum(R * R, axis=1)[None, :] - 2.0 * M @ R.T
        n_best = (full == full.min(axis=1, keepdims=True)).sum(axis=1)
        assert np.mean(n_best > 1) > 0.1
