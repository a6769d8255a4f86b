"""MAP library matching on the GPU (the matchers of madi/library.py:1163-1522).

``match_voxels_batch`` (S0 fixed: data already divided by the measured b=0
signal) and ``match_voxels_batch_fits0`` (S0 a free per-voxel scale) keep the
reference's signatures and host logic: requested (delta, Delta, b) columns
snap to the nearest stored column inside the acquisition tolerances (never
interpolated), candidates are filtered by realised v_i, rho and free water,
and the fixed-S0 matcher may work on log signals.  The exhaustive search, every
voxel against every candidate entry with the first minimum kept on ties, runs
in the FP64 HIP kernel ``madi_map_match`` (the reference's GPU path,
madi/fitters_gpu.py:96-154 and :414-463).  There is no CPU path here:
``use_gpu=False`` raises.
"""

from __future__ import annotations

import ctypes as C
import warnings
from typing import Optional

import numpy as np

from . import _hip

# Acquisition snap tolerances of madi/library.py:1171-1172: a request inside
# them uses the nearest stored coordinate (and is reported), outside them it
# is rejected.
B_SNAP_TOL_S_MM2 = 30.0
TIMING_SNAP_TOL_MS = 1.5


def _nearest_pair_index(delta: float, Delta: float, lib_pairs: np.ndarray) -> int:
    """Stored (delta, Delta) pair nearest in the plane, first on ties (madi/library.py:1174-1176)."""
    d2 = (lib_pairs[:, 0] - delta) ** 2 + (lib_pairs[:, 1] - Delta) ** 2
    return int(np.argmin(d2))


def resolve_grid_columns(fit_triples, lib_delta_pairs, lib_b_values, n_b, b_tol=B_SNAP_TOL_S_MM2,
                         timing_tol=TIMING_SNAP_TOL_MS):
    """Columns of the flat library vector (pair-major, then b) for each
    requested (delta, Delta, b), and one snap event per distinct inexact
    timing pair or b-value (madi/library.py:1179-1257).  Raises ValueError
    outside the tolerances or without stored b-values."""
    if lib_b_values is None:
        raise ValueError("Library has no stored b-values. Rebuild the library with the current _save_library, "
                         "or pass lib_b_values explicitly.")
    pairs = np.asarray(lib_delta_pairs, dtype=float)
    bvals = np.asarray(lib_b_values, dtype=float)
    cols = np.empty(len(fit_triples), dtype=int)
    events: list = []
    timing_seen: set = set()
    b_seen: set = set()
    for n, (delta, Delta, b) in enumerate(fit_triples):
        p = _nearest_pair_index(delta, Delta, pairs)
        used_d, used_D = (float(x) for x in pairs[p])
        off_d, off_D = abs(used_d - delta), abs(used_D - Delta)
        if off_d > timing_tol or off_D > timing_tol:
            raise ValueError(f"(δ={delta:g}, Δ={Delta:g}) ms is outside the {timing_tol:g} ms timing snap "
                             f"tolerance; nearest stored pair is (δ={used_d:g}, Δ={used_D:g}) ms. "
                             "No interpolation is implemented.")
        key = (float(delta), float(Delta), used_d, used_D)
        if (off_d > 1e-12 or off_D > 1e-12) and key not in timing_seen:
            timing_seen.add(key)
            events.append({
                "axis": "timing",
                "requested": {"delta_ms": float(delta), "Delta_ms": float(Delta)},
                "used": {"delta_ms": used_d, "Delta_ms": used_D},
                "abs_offset": {"delta_ms": off_d, "Delta_ms": off_D},
                "rel_offset": {"delta": off_d / max(abs(float(delta)), 1e-12),
                               "Delta": off_D / max(abs(float(Delta)), 1e-12)},
            })
        q = int(np.argmin(np.abs(bvals - b)))
        used_b = float(bvals[q])
        off_b = abs(used_b - float(b))
        if off_b > b_tol:
            raise ValueError(f"b = {b} s/mm² not within {b_tol:g} of any library b-value {sorted(lib_b_values)}.")
        bkey = (float(b), used_b)
        if off_b > 1e-12 and bkey not in b_seen:
            b_seen.add(bkey)
            events.append({"axis": "b", "requested": float(b), "used": used_b, "abs_offset": off_b,
                           "rel_offset": off_b / max(abs(float(b)), 1e-12)})
        cols[n] = p * n_b + q
    return cols, events


def _grid_columns(fit_triples, lib_delta_pairs, lib_b_values, n_b, b_tol=B_SNAP_TOL_S_MM2,
                  pair_warn_tol=TIMING_SNAP_TOL_MS):
    """Columns only (madi/library.py:1260-1273)."""
    return resolve_grid_columns(fit_triples, lib_delta_pairs, lib_b_values, n_b, b_tol=b_tol,
                                timing_tol=pair_warn_tol)[0]


def _build_candidate_lib_matrix(library, lib_delta_pairs, lib_b_values, n_b, vi_min, vi_max, rho_max, fit_triples,
                                *, include_free_water=False, return_weights=False, require_weights=False):
    """Candidate entries (realised v_i in [vi_min, vi_max], rho <= rho_max,
    free water only when asked) and their requested columns
    (madi/library.py:1276-1330): (lib_mat, kios, rhos, Vs[, weights])."""
    vi = np.array([e.realised_vi for e in library])
    rho = np.array([e.rho for e in library])
    free = np.array([e.is_free_water for e in library], dtype=bool)
    keep = (vi >= vi_min) & (vi <= vi_max)
    keep = (keep | free) if include_free_water else (keep & ~free)
    if rho_max is not None:
        keep &= rho <= rho_max
    n = int(keep.sum())
    if n == 0:
        raise ValueError(f"No library entries survive vi in [{vi_min}, {vi_max}] and rho <= {rho_max}.")
    if n < 50:
        warnings.warn(f"Only {n} library entries pass the filter.")
    chosen = [e for e, k in zip(library, keep) if k]
    cols = _grid_columns(fit_triples, lib_delta_pairs, lib_b_values, n_b)
    lib_mat = np.array([e.vector for e in chosen])[:, cols]
    kios = np.array([e.kio for e in chosen])
    rhos = np.array([e.rho for e in chosen])
    Vs = np.array([e.V for e in chosen])
    if not return_weights:
        return lib_mat, kios, rhos, Vs
    w = np.array([np.nan if e.weight is None else float(e.weight) for e in chosen], dtype=float)
    if require_weights and not np.all(np.isfinite(w) & (w > 0.0)):
        raise ValueError("Bayesian fitting requires positive per-entry quadrature weights. This library is "
                         "legacy/unweighted; rebuild it with the remediation grid.")
    return lib_mat, kios, rhos, Vs, w


class DeviceLibrary:
    """A candidate matrix (entries x features, FP64) resident in HBM, matched
    against batches of voxels with ``madi_map_match``."""

    def __init__(self, lib_mat):
        torch = _hip.require_gpu()
        self._lib = _hip.load_library()
        R = np.ascontiguousarray(lib_mat, dtype=np.float64)
        if R.ndim != 2 or R.shape[0] < 1:
            raise ValueError("the library matrix must be 2-D with at least one entry")
        self.n_entries, self.n_features = R.shape
        self.rows = torch.from_numpy(R).to("cuda")

    def workspace(self, n_vox: int):
        torch = _hip.torch_mod()
        need = C.c_int64()
        _hip.check(self._lib.madi_map_workspace(int(n_vox), self.n_entries, C.byref(need)), "madi_map_workspace")
        return torch.empty(-(-int(need.value) // 8), dtype=torch.float64, device="cuda")

    def match_device(self, measured, fit_s0: bool = False, workspace=None, out=None):
        """Match a device tensor (n_vox x n_features, FP64, contiguous): returns
        device tensors (row index, residual, s0 or None), queued on the current
        stream."""
        torch = _hip.torch_mod()
        if measured.dtype != torch.float64 or measured.dim() != 2 or not measured.is_contiguous():
            raise ValueError("measured must be a contiguous 2-D float64 tensor")
        if measured.shape[1] != self.n_features:
            raise ValueError(f"measured has {measured.shape[1]} features, the library {self.n_features}")
        n = int(measured.shape[0])
        ws = workspace if workspace is not None else self.workspace(n)
        if out is None:
            out = (torch.empty(n, dtype=torch.int64, device="cuda"), torch.empty(n, dtype=torch.float64, device="cuda"),
                   torch.empty(n, dtype=torch.float64, device="cuda") if fit_s0 else None)
        idx, res, s0 = out
        rc = self._lib.madi_map_match(_hip.dptr(measured), n, _hip.dptr(self.rows), self.n_entries, self.n_features,
                                      int(bool(fit_s0)), _hip.dptr(ws), ws.numel() * 8, _hip.dptr(idx),
                                      _hip.dptr(res), _hip.dptr(s0) if s0 is not None else None, _hip.stream_ptr())
        _hip.check(rc, "madi_map_match")
        return idx, res, s0

    def match(self, measured, fit_s0: bool = False):
        """Host arrays in and out: (row index, residual, s0 or None)."""
        torch = _hip.torch_mod()
        M = np.ascontiguousarray(measured, dtype=np.float64)
        if M.ndim != 2:
            raise ValueError("measured must be 2-D (voxels x features)")
        idx, res, s0 = self.match_device(torch.from_numpy(M).to("cuda"), fit_s0)
        return idx.cpu().numpy(), res.cpu().numpy(), (s0.cpu().numpy() if s0 is not None else None)


def _require_gpu_path(use_gpu) -> None:
    if use_gpu is False:
        raise ValueError("this build matches in its HIP kernel only: use_gpu=False has no CPU path")


def match_voxels_batch(measured_batch, library, lib_delta_pairs, lib_b_values, n_b, fit_triples, log_space=False,
                       s_floor=1e-3, vi_min=0.5, vi_max=0.95, rho_max=None, include_free_water=False,
                       use_gpu: Optional[bool] = None):
    """Nearest library entry per voxel with S0 fixed (madi/library.py:1402-1450):
    argmin of |m - r|^2 over candidates, optionally on log(clip(x, s_floor, 1)).
    Returns (kio, rho, V, residual) per voxel."""
    _require_gpu_path(use_gpu)
    lib_mat, kios, rhos, Vs = _build_candidate_lib_matrix(
        library, lib_delta_pairs, lib_b_values, n_b, vi_min, vi_max, rho_max, fit_triples,
        include_free_water=include_free_water)
    if log_space:
        measured = np.log(np.clip(measured_batch, s_floor, 1.0))
        lib_m = np.log(np.clip(lib_mat, s_floor, 1.0))
    else:
        measured, lib_m = measured_batch, lib_mat
    idx, res, _ = DeviceLibrary(lib_m).match(measured)
    return kios[idx], rhos[idx], Vs[idx], res


def match_voxels_batch_fits0(raw_signal, library, lib_delta_pairs, lib_b_values, n_b, fit_triples, vi_min=0.5,
                             vi_max=0.95, rho_max=None, include_free_water=False, use_gpu: Optional[bool] = None):
    """Nearest library entry per voxel with S0 a free scale (madi/library.py:1457-1522):
    S0 = m.r / r.r, residual |m|^2 - (m.r)^2 / r.r, candidates with S0 <= 0
    excluded.  Returns (kio, rho, V, residual, S0) per voxel; a voxel with no
    positive-S0 candidate gets entry 0, residual +inf and S0 0, as the
    reference's GPU kernel gives."""
    _require_gpu_path(use_gpu)
    lib_mat, kios, rhos, Vs = _build_candidate_lib_matrix(
        library, lib_delta_pairs, lib_b_values, n_b, vi_min, vi_max, rho_max, fit_triples,
        include_free_water=include_free_water)
    idx, res, s0 = DeviceLibrary(lib_mat.astype(np.float64)).match(np.asarray(raw_signal, dtype=np.float64),
                                                                    fit_s0=True)
    return kios[idx], rhos[idx], Vs[idx], res, s0
