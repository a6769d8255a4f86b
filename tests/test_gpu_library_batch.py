"""Batched library builds and library serving (custom_madi_amd.library_batch)
against the sequential group-by-group build (madi/library.py:546-629).

* build_library_from_triplets with batched=True gives the sequential
  build's entries bit for bit, in the same order, and the same v5 file.
  FP32, v5 grid (so the scheduler's Y mode), two ensembles, two walker
  chunks per walk; also with the ensembles' CPU geometry prepared apart
  from their GPU v_i measurement (the worker-pool split).
* A group with a failed walk raises the sequential path's RuntimeError
  ("fatal boundary condition") after the groups before it are
  checkpointed, as the sequential loop does.
* Served groups (continuous batching, completion order) append v5 entries
  equal to the built ones, and the served file loads.
"""

from __future__ import annotations

import sys
from dataclasses import replace

import numpy as np
import pytest

from custom_madi_amd import library, library_batch, signal
from custom_madi_amd.config import ENSEMBLE_MEAN_SUBSET_B_VALUES_S_MM2, SimConfig

pytestmark = pytest.mark.gpu

TRIPLETS = [(20.0, 1.0e5, 6.0), (40.0, 1.0e5, 6.0), (0.0, 0.0, 0.0)]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _cfg():
    return SimConfig(n_walkers=2048, walker_chunk=1024, n_ensembles=2, T_max_ms=110.0, D0=2.0, ts=0.05,
                     small_deltas=[5.0, 7.0, 10.0, 12.0, 15.0, 20.0, 25.0, 30.0],
                     big_deltas=[15.0, 25.0, 30.0, 36.0, 40.0, 50.0, 60.0, 80.0],
                     b_values=list(ENSEMBLE_MEAN_SUBSET_B_VALUES_S_MM2), geometry_validation_points=20_000,
                     geometry_vi_tolerance=0.03, precision="fp32")


@pytest.fixture(scope="module")
def sequential(tmp_path_factory):
    d = tmp_path_factory.mktemp("libbatch")
    say("[libbatch] sequential FP32 v5 build")
    lib = library.build_library_from_triplets(TRIPLETS, cfg=_cfg(), seed=7, save_path=str(d / "seq.npz"),
                                              verbose=False, batched=False)
    return d, lib


def _same_entry(a, b):
    fields = ("kio", "rho", "V", "vi", "pp", "kio_analytic_eq5", "kio_nominal", "rho_nominal", "V_nominal")
    num = lambda e: np.array([np.nan if getattr(e, f) is None else getattr(e, f) for f in fields])  # noqa: E731
    assert np.array_equal(num(a), num(b), equal_nan=True)   # free water's k_io is NaN
    assert a.is_free_water == b.is_free_water
    for name in ("vector", "signal_imag", "signal_variance", "ensemble_means_subset"):
        assert np.array_equal(np.asarray(getattr(a, name)), np.asarray(getattr(b, name))), name
    assert a.metadata == b.metadata


def test_batched_build_equals_sequential(sequential):
    d, seq = sequential
    cols = signal.build_columns(_cfg())
    assert library_batch.batchable(_cfg(), cols)
    say("[libbatch] batched FP32 v5 build")
    bat = library.build_library_from_triplets(TRIPLETS, cfg=_cfg(), seed=7, save_path=str(d / "bat.npz"),
                                              verbose=False, batched=True)
    assert len(bat) == len(seq) == 3
    for a, b in zip(seq, bat):
        _same_entry(a, b)
    f_seq, f_bat = np.load(d / "seq.npz"), np.load(d / "bat.npz")
    assert sorted(f_seq.files) == sorted(f_bat.files)
    for k in f_seq.files:
        a, b = f_seq[k], f_bat[k]
        assert np.array_equal(a, b, equal_nan=a.dtype.kind in "fc"), k   # NaN: free water's k_io
    back = library.load_library(str(d / "bat.npz"))
    assert library.load_library_meta(str(d / "bat.npz"))["has_v5_diagnostics"]
    for a, b in zip(back, library.load_library(str(d / "seq.npz"))):
        _same_entry(a, b)


class _InlinePool:
    """apply_async runs the job at once in this process: the worker/GPU split
    of ensemble construction (prepare_ensemble_host in a worker,
    finish_prepared here) without spawning processes."""

    class _Done:
        def __init__(self, value):
            self.value = value

        def ready(self):
            return True

        def get(self):
            return self.value

    def apply_async(self, fn, args):
        return self._Done(fn(*args))


def test_batched_build_with_worker_prepared_geometry_equals_sequential(sequential):
    d, seq = sequential
    say("[libbatch] batched build, ensembles prepared by the worker split")
    bat = library.build_library_from_triplets(TRIPLETS, cfg=_cfg(), seed=7, save_path=str(d / "pool.npz"),
                                              verbose=False, batched=True, pool=_InlinePool())
    assert len(bat) == len(seq)
    for a, b in zip(seq, bat):
        _same_entry(a, b)


def test_a_failed_group_raises_after_the_earlier_groups_are_saved(tmp_path):
    cfg = _cfg()
    cols = signal.build_columns(cfg)
    calls = []

    def prepare(rho, V, c, seed):
        calls.append((rho, V))
        if len(calls) == 2:      # the second group in a 3 um box: its walkers leave it
            c = replace(c, L=3.0)
        return library_batch.prepare_group(rho, V, c, seed)

    cp = library_batch.LibraryCheckpoint(str(tmp_path / "fail.npz"), cfg, cols, build_seed=1)
    say("[libbatch] failing second group")
    with pytest.raises(RuntimeError, match="fatal boundary condition"):
        library_batch.build_groups_batched([((0.0, 0.0), [0.0]), ((0.0, 0.0), [5.0]), ((0.0, 0.0), [9.0])], cfg,
                                           cols, 1, cp, prepare=prepare)
    back = library.load_library(str(tmp_path / "fail.npz"))
    assert len(back) == 1 and back[0].is_free_water and back[0].kio_nominal == 0.0


def test_served_groups_append_v5_entries_equal_to_the_built_ones(sequential, tmp_path):
    d, seq = sequential
    cfg = _cfg()
    cols = signal.build_columns(cfg)
    b = library_batch.LibraryBuilder(cfg, cols, seed=7, ordered=False)
    cp = library_batch.LibraryCheckpoint(str(tmp_path / "served.npz"), cfg, cols, build_seed=7)
    arrivals = [(0.0, 1.0e5, 6.0, [40.0]), (0.05, 0.0, 0.0, [0.0]), (0.1, 1.0e5, 6.0, [20.0])]
    say("[libbatch] serving three groups")
    try:
        done, failed = library_batch.serve_library(b, arrivals, cp)
    finally:
        b.close()
    assert not failed and len(done) == 3
    by_key = {library._entry_key(e.kio_nominal, e.rho_nominal, e.V_nominal): e for e in seq}
    served = library.load_library(str(tmp_path / "served.npz"))
    assert len(served) == 3 and cp.writes == 3
    for e in served:
        _same_entry(e, by_key[library._entry_key(e.kio_nominal, e.rho_nominal, e.V_nominal)])
    assert all(gr.latency_s > 0.0 for gr in done)
