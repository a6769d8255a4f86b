"""The allocation-free fused walk (madi_walk_fused_static) against madi_walk_fused.

Both are called through the C-ABI on the same geometry and parameters; the
results (cos/sin sums, occupancy, flags, counters) must be equal bit for
bit, on two calls in a row with the second smaller than the first (the
shared scratch keeps the first call's contents), for the production FP32
walk in its default and per-tile variants and for the FP64 walk.  An entry
whose workspace exceeds the scratch is rejected before anything runs.
"""

from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

from custom_madi_amd import _hip, batch, engine
from custom_madi_amd.config import SimConfig
from custom_madi_amd.ensemble import DeviceGeometry, create_periodic_ensemble, geometry_arrays

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def entry():
    cfg = SimConfig(D0=2.0, ts=0.02, T_max_ms=46.0, h_ms=1.0, n_walkers=1, small_deltas=[6.0],
                    big_deltas=[15.0, 25.0, 30.0, 40.0], b_values=[0.0, 1000.0, 2500.0, 4000.0, 6000.0],
                    periodic=True, precision="fp32")
    ens = create_periodic_ensemble(300, 5.0e5, 0.6, cfg, seed=7)
    dg = DeviceGeometry(geometry_arrays(ens))
    spec = batch.acquisition_spec(24)
    return cfg, dg, spec


def _outputs(torch, n_cols, n_grid):
    return (torch.empty(n_cols, dtype=torch.float64, device="cuda"),
            torch.empty(n_cols, dtype=torch.float64, device="cuda"),
            torch.empty(n_grid, dtype=torch.int64, device="cuda"),
            torch.empty(_hip.N_FLAGS, dtype=torch.int64, device="cuda"),
            torch.empty(_hip.N_COUNTERS, dtype=torch.int64, device="cuda"))


def _host(outs):
    return [o.cpu().numpy().copy() for o in outs]


def _run_both(dg, params, plan, precision, n_cols, n_grid):
    """(madi_walk_fused outputs, madi_walk_fused_static outputs) on the host."""
    torch = _hip.require_gpu()
    lib = _hip.load_library()
    s = _hip.stream_ptr()
    need = C.c_int64()
    _hip.check(lib.madi_walk_fused_workspace(params.n_walkers, C.byref(plan), C.byref(need)),
               "madi_walk_fused_workspace")
    ws = torch.empty(int(need.value), dtype=torch.uint8, device="cuda")
    a = _outputs(torch, n_cols, n_grid)
    rc = lib.madi_walk_fused(dg.handle, C.byref(params), C.byref(plan), precision, _hip.dptr(ws), int(need.value),
                             *[_hip.dptr(o) for o in a], s)
    _hip.check(rc, "madi_walk_fused")
    b = _outputs(torch, n_cols, n_grid)
    rc = lib.madi_walk_fused_static(dg.handle, C.byref(params), C.byref(plan), precision,
                                    *[_hip.dptr(o) for o in b], s)
    _hip.check(rc, "madi_walk_fused_static")
    torch.cuda.synchronize()
    return _host(a), _host(b)


# counters fixed by each walker's own work; the iteration counters
# (wave_iters, lane_iters, class_iters, class_busy) depend on how the waves
# were scheduled and differ from run to run
_WORK_COUNTERS = [_hip.COUNTER_NAMES.index(k) for k in (
    "classify", "safe_hit", "graph_iters", "fp64_recheck", "facets", "kd_nodes", "crossings", "rejects", "exits")]


def _assert_same(a, b, what):
    names = ("cos", "sin", "occupancy", "flags")
    for name, x, y in zip(names, a, b):
        assert np.array_equal(x, y), f"{what}: {name} differs"
    assert np.array_equal(a[4][_WORK_COUNTERS], b[4][_WORK_COUNTERS]), f"{what}: work counters differ"
    assert a[3][0] == 0 and a[3][1] == 0, f"{what}: escapes or classifier failures"


@pytest.mark.parametrize("classify_mode, precision", [(0, 32), (1, 32), (0, 64)])
def test_static_equals_workspace_walk_twice(entry, classify_mode, precision):
    cfg, dg, spec = entry
    plan = spec.to_ctypes()
    n_cols = int(np.prod(spec.shape))
    # the second call is smaller than the first and starts from the first's scratch
    for n, seed in ((40_000, 11), (9_000, 12)):
        if precision == 64:
            n //= 8
        say(f"[no-alloc] classify_mode={classify_mode} precision={precision} walkers={n} seed={seed}")
        params = engine.walk_params(n, cfg, 0.01, seed, classify_mode=classify_mode)
        a, b = _run_both(dg, params, plan, precision, n_cols, cfg.n_grid)
        _assert_same(a, b, f"walkers={n}")
        assert a[2][0] > 0  # some walkers start inside a cell


def test_static_rejects_entries_beyond_its_scratch(entry):
    cfg, dg, spec = entry
    torch = _hip.require_gpu()
    lib = _hip.load_library()
    plan = spec.to_ctypes()
    mx = C.c_int64()
    _hip.check(lib.madi_walk_fused_static_max_bytes(C.byref(mx)), "madi_walk_fused_static_max_bytes")
    assert mx.value == 128 << 20
    n = 4_000_000  # about 320 MB of workspace at 480 columns
    need = C.c_int64()
    _hip.check(lib.madi_walk_fused_workspace(n, C.byref(plan), C.byref(need)), "madi_walk_fused_workspace")
    assert need.value > mx.value
    outs = _outputs(torch, int(np.prod(spec.shape)), cfg.n_grid)
    params = engine.walk_params(n, cfg, 0.01, 3)
    say("[no-alloc] oversized entry")
    rc = lib.madi_walk_fused_static(dg.handle, C.byref(params), C.byref(plan), 32, *[_hip.dptr(o) for o in outs],
                                    _hip.stream_ptr())
    assert rc == _hip.MADI_ERR_ARG
    assert "smaller than" in _hip.last_error()
    torch.cuda.synchronize()
