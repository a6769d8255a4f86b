"""Shard merge of v5 libraries (scripts/merge_shards.py:54-132) on the CPU:
entries keep shard order, a (k_io, rho, V) key already taken is skipped
(first occurrence wins; the free-water atom is one key), every shard must
carry the first one's grid, schema, seed and uncertainty contract, and shard
ids are checked when asked."""

from __future__ import annotations

import numpy as np
import pytest

from custom_madi_amd import library
from custom_madi_amd.config import SimConfig


def _entry(rng, n_cols, kio, rho, V, free=False):
    return library.LibraryEntry(
        kio=kio, rho=rho, V=V, vector=rng.uniform(0, 1, n_cols), signal_imag=rng.normal(0, 1e-3, n_cols),
        signal_variance=rng.uniform(0, 1e-4, n_cols),
        ensemble_means_subset=rng.uniform(0, 1, (2, library.ENSEMBLE_MEAN_SUBSET_N_COLUMNS)), vi=rho * V * 1e-6,
        kio_nominal=kio, rho_nominal=rho, V_nominal=V, pp=0.01, kio_analytic_eq5=kio, is_free_water=free,
        metadata={"nominal": {"kio": kio, "rho": rho, "V": V}})


@pytest.fixture(scope="module")
def shards(tmp_path_factory):
    d = tmp_path_factory.mktemp("shards")
    cfg = SimConfig(n_ensembles=2)
    n_cols = len(cfg.delta_pairs()) * len(cfg.b_values)
    rng = np.random.default_rng(4)
    a, b, c = _entry(rng, n_cols, 20.0, 1e5, 6.0), _entry(rng, n_cols, 40.0, 1e5, 6.0), _entry(rng, n_cols, 10.0, 2e5, 3.0)
    f = _entry(rng, n_cols, 0.0, 0.0, 0.0, free=True)
    b2 = _entry(rng, n_cols, 40.0, 1e5, 6.0)                  # same key as b, other values
    f2 = _entry(rng, n_cols, 0.0, 0.0, 0.0, free=True)
    p0, p1 = d / "lib.shard0.npz", d / "lib.shard1.npz"
    library._save_library([a, b, f], str(p0), cfg=cfg, build_seed=3)
    library._save_library([b2, c, f2], str(p1), cfg=cfg, build_seed=3)
    return d, cfg, (a, b, c, f), (p0, p1)


def test_merge_dedupes_first_occurrence_wins(shards):
    d, cfg, (a, b, c, f), (p0, p1) = shards
    out = d / "merged.npz"
    merged = library.merge_library_shards([str(p0), str(p1)], str(out), require_shards=2, verbose=False)
    assert [(e.kio, e.rho, e.V) for e in merged] == [(20.0, 1e5, 6.0), (40.0, 1e5, 6.0), (0.0, 0.0, 0.0),
                                                     (10.0, 2e5, 3.0)]
    back = library.load_library(str(out))
    for got, want in zip(back, (a, b, f, c)):
        assert np.array_equal(got.vector, want.vector)
        assert np.array_equal(got.signal_imag, want.signal_imag.astype(np.float32))
        assert np.array_equal(got.ensemble_means_subset, want.ensemble_means_subset.astype(np.float32))
        assert got.is_free_water == want.is_free_water
    meta = library.load_library_meta(str(out))
    m0 = library.load_library_meta(str(p0))
    assert meta["library_schema"] == "madi-library-v5" and meta["has_v5_diagnostics"]
    assert meta["build_metadata"] == m0["build_metadata"]
    assert meta["delta_pairs"] == m0["delta_pairs"] and meta["b_values"] == m0["b_values"]


def test_merge_refuses_mismatched_shards(shards):
    d, cfg, (a, b, c, f), (p0, p1) = shards
    other = d / "lib.shard2.npz"
    library._save_library([c], str(other), cfg=cfg, build_seed=4)   # another build seed
    with pytest.raises(ValueError, match="mismatch"):
        library.merge_library_shards([str(p0), str(other)], str(d / "bad.npz"), verbose=False)
    with pytest.raises(ValueError, match="coverage"):
        library.merge_library_shards([str(p0), str(p1)], str(d / "bad.npz"), require_shards=3, verbose=False)
    with pytest.raises(ValueError, match="shard id"):
        library.merge_library_shards([str(p0), str(d / "merged_x.npz")], str(d / "bad.npz"), require_shards=2,
                                     verbose=False)
