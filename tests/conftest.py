"""Shared test setup: repo root on sys.path, the gpu marker, small fixtures."""

from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP kernels of libmadi_hip.so)")


def gpu_ready() -> bool:
    try:
        import torch
        return bool(torch.cuda.is_available())
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if gpu_ready():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def uncertified_reference(tmp_path_factory):
    """Schema-valid 26-row <A/V> table with <A/V>_norm = 4 - 2 v_i.

    Mirrors the reference's unit-test fixture
    (tests/physics_audit/_si_test_support.py:14-47): linear in v_i so
    interpolation and rho^(1/3) scaling have closed forms.
    """
    from custom_madi_amd.ensemble import alpha_star_from_vi

    path = tmp_path_factory.mktemp("ref") / "si_geometry_reference_test_fixture.npz"
    vi = np.linspace(0.40, 1.00, 26)
    ax = np.asarray([0.0 if v >= 1.0 else alpha_star_from_vi(v, 1.0e9) for v in vi])
    meta = {"schema": "madi-si-geometry-test-fixture-v1", "kappa": 0.90, "n_single_cells": 16, "n_alpha": 26,
            "mean_estimator": "untrimmed_arithmetic_mean_A_over_V",
            "contraction_rule": "all_shifted_voronoi_facets_S2"}
    np.savez(path, vi=vi, alpha_x=ax, mean_A_over_V_norm=4.0 - 2.0 * vi,
             metadata_json=np.array(json.dumps(meta, sort_keys=True)))
    return path
