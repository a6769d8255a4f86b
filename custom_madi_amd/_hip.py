"""ctypes binding of libmadi_hip.so (the C ABI declared in include/madi_hip.h).

PyTorch-ROCm tensors provide device memory and the current stream; every
computation runs in the HIP kernels of the library.  A missing library or a
missing GPU raises: there is no CPU fallback behind these functions.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path
from typing import Optional

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("MADI_HIP_LIB", HERE / "libmadi_hip.so"))

MADI_OK = 0
MADI_ERR_ARG = 1
MADI_ERR_HIP = 2
MADI_ERR_OOM = 3
MADI_ERR_CLASSIFIER = 4
MADI_ERR_ESCAPE = 5
MADI_ERR_BUSY = 6

MADI_MAX_PAIRS = 8
MADI_MAX_B = 16
MADI_MAX_DIRS = 64
MADI_MAX_REDUCE_COLS = 2_000_000
MADI_MAX_MULTI_WALKS = 64
N_COUNTERS = 13
N_FLAGS = 2
MADI_FLAG_ESCAPED = 0
MADI_FLAG_CLASSIFIER_ERR = 1
N_VERIFY = 7
VERIFY_NAMES = ("tier_checked", "tier_mismatch", "exit_checked", "exit_mismatch", "classifier_checked",
                "classifier_mismatch", "exact_failed")
COUNTER_NAMES = (
    "classify", "safe_hit", "graph_iters", "fp64_recheck", "facets", "kd_nodes", "crossings", "rejects",
    "wave_iters", "lane_iters", "class_iters", "class_busy", "exits",
)

_p = C.c_void_p
_i8p = C.POINTER(C.c_int8)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_f64p = C.POINTER(C.c_double)


class GeomDesc(C.Structure):
    _fields_ = [
        ("n_kd", C.c_int32), ("kd_xyz", _f64p), ("kd_annulus", _f64p), ("kd_label", _i32p),
        ("node_seed", _i32p), ("node_axis", _i8p), ("node_left", _i32p), ("node_right", _i32p),
        ("node_parent", _i32p),
        ("n_cells", C.c_int32), ("cell_xyz", _f64p), ("cell_annulus", _f64p),
        ("n_edges", C.c_int64), ("nb_offsets", _i64p), ("nb_ids", _i32p), ("nb_vec", _f64p),
        ("hint_n", C.c_int32), ("hint", _i32p),
        ("L", C.c_double), ("source_lo", C.c_double), ("source_hi", C.c_double),
        ("periodic", C.c_int32), ("free_water", C.c_int32),
        ("cell_incomplete", _i8p),
    ]


class WalkParams(C.Structure):
    _fields_ = [
        ("n_walkers", C.c_int64), ("n_steps", C.c_int32), ("steps_per_h", C.c_int32), ("n_grid", C.c_int32),
        ("tier_cap_um", C.c_float), ("sigma", C.c_double), ("ts", C.c_double), ("pp", C.c_double),
        ("seed", C.c_uint64), ("first_walker", C.c_uint64), ("scan_rounds", C.c_int32),
        ("classify_mode", C.c_int32),
    ]


class Work(C.Structure):
    _fields_ = [
        ("geom", C.c_void_p), ("walk", WalkParams), ("tile_lo", C.c_int64), ("n_tiles", C.c_int64),
        ("d_partial", C.c_void_p), ("d_occupancy", C.c_void_p), ("d_flags", C.c_void_p),
        ("d_moments", C.c_void_p), ("d_Y", C.c_void_p),
    ]


class Tune(C.Structure):
    _fields_ = [("grid_per_cu", C.c_int32), ("tier_cap_um", C.c_float), ("scan_rounds", C.c_int32),
                ("classify_mode", C.c_int32)]


TILE_WALKERS = 256


class PhasePlan(C.Structure):
    _fields_ = [
        ("n_pairs", C.c_int32), ("n_b", C.c_int32), ("n_dirs", C.c_int32), ("reserved", C.c_int32),
        ("j_delta", C.c_int32 * MADI_MAX_PAIRS), ("j_Delta", C.c_int32 * MADI_MAX_PAIRS),
        ("j_sum", C.c_int32 * MADI_MAX_PAIRS),
        ("pc", C.c_double * (MADI_MAX_PAIRS * MADI_MAX_B)),
        ("dirs", C.c_double * (MADI_MAX_DIRS * 3)),
    ]


class DeviceInfo(C.Structure):
    _fields_ = [
        ("device", C.c_int32), ("cu_count", C.c_int32), ("clock_khz", C.c_int32), ("xcc_count", C.c_int32),
        ("total_mem", C.c_int64), ("free_mem", C.c_int64), ("name", C.c_char * 64), ("arch", C.c_char * 32),
    ]


# name -> (restype, argtypes) for every symbol the header declares
SIGNATURES = {
    "madi_abi_version": (C.c_int, []),
    "madi_last_error": (C.c_char_p, []),
    "madi_get_device_info": (C.c_int, [C.POINTER(DeviceInfo)]),
    "madi_geom_create": (C.c_int, [C.POINTER(GeomDesc), C.POINTER(_p)]),
    "madi_geom_destroy": (C.c_int, [_p]),
    "madi_geom_destroy_async": (C.c_int, [_p, _p]),
    "madi_geom_device_bytes": (C.c_int, [_p, _i64p]),
    "madi_classify": (C.c_int, [_p, _p, C.c_int64, _p, _p, _p, _p]),
    "madi_classify_graph": (C.c_int, [_p, _p, C.c_int64, _p, _p, _p, _p, _p]),
    "madi_rng_init": (C.c_int, [C.c_uint64, C.c_uint64, C.c_int64, _p, _p]),
    "madi_walk_f64": (C.c_int, [_p, C.POINTER(WalkParams), _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    "madi_walk_f64_multi": (C.c_int, [_p, C.POINTER(WalkParams), C.c_int32, _p, _p, _p, _p, _p, _p, _p]),
    "madi_count_inside": (C.c_int, [_p, C.c_int64, C.c_int64, C.c_int32, _p, _p]),
    "madi_reduce_f64": (C.c_int, [_p, C.c_int64, C.c_int32, _p, _p, _p, _p, C.c_int32, _p, _p, _p]),
    "madi_walk_fused_workspace": (C.c_int, [C.c_int64, C.POINTER(PhasePlan), _i64p]),
    "madi_walk_fused": (C.c_int, [_p, C.POINTER(WalkParams), C.POINTER(PhasePlan), C.c_int32, _p, C.c_int64,
                                  _p, _p, _p, _p, _p, _p]),
    "madi_walk_fused_verify": (C.c_int, [_p, C.POINTER(WalkParams), C.POINTER(PhasePlan), _p, C.c_int64, _p, _p,
                                         _p, _p, _p, _p, _p]),
    "madi_walk_fused_static_max_bytes": (C.c_int, [_i64p]),
    "madi_walk_fused_static": (C.c_int, [_p, C.POINTER(WalkParams), C.POINTER(PhasePlan), C.c_int32, _p, _p, _p,
                                         _p, _p, _p]),
    "madi_moments_bytes": (C.c_int, [C.c_int64, C.POINTER(PhasePlan), _i64p]),
    "madi_kd_build": (C.c_int, [_p, C.c_int64, _p, _p, _p, _p, _p]),
    "madi_nearest_points": (C.c_int, [_p, C.c_int64, _p, C.c_int64, C.c_double, C.c_int32, C.c_int32, _p, _p]),
    "madi_voronoi_graph": (C.c_int, [_p, C.c_int64, C.c_int64, C.c_double, C.c_int32, _p, _p, C.c_int64, _p, _p,
                                     _p]),
    "madi_walk_y_f32_workspace": (C.c_int, [C.c_int64, _i64p]),
    "madi_walk_y_f32": (C.c_int, [_p, C.POINTER(WalkParams), _p, C.c_int64, _p, _p, _p, _p, _p]),
    "madi_batch_workspace": (C.c_int, [C.c_int32, _i64p]),
    "madi_batch_launch": (C.c_int, [C.POINTER(Work), C.c_int32, C.POINTER(PhasePlan), C.POINTER(Tune), _p, C.c_int64,
                                    _p, _p]),
    "madi_batch_graph_create": (C.c_int, [C.POINTER(PhasePlan), C.c_int32, C.POINTER(Tune), _p, C.POINTER(_p)]),
    "madi_batch_graph_launch": (C.c_int, [_p, C.POINTER(Work), C.c_int32, _p]),
    "madi_batch_graph_destroy": (C.c_int, [_p]),
    "madi_tiles_finalize": (C.c_int, [C.POINTER(PhasePlan), _p, C.c_int64, _p, _p, _p]),
    "madi_tiles_finalize_workspace": (C.c_int, [C.POINTER(PhasePlan), C.c_int64, _i64p]),
    "madi_tiles_finalize_ws": (C.c_int, [C.POINTER(PhasePlan), _p, C.c_int64, _p, _p, _p, C.c_int64, _p]),
    "madi_fill_i64": (C.c_int, [_p, C.c_int64, C.c_int64, _p]),
    "madi_map_workspace": (C.c_int, [C.c_int64, C.c_int64, _i64p]),
    "madi_map_match": (C.c_int, [_p, C.c_int64, _p, C.c_int64, C.c_int32, C.c_int32, _p, C.c_int64, _p, _p, _p, _p]),
}

_lib: Optional[C.CDLL] = None


def load_library() -> C.CDLL:
    """Load libmadi_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.is_file():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()' or python custom_madi_amd/build.py)"
        )
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    msg = load_library().madi_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str) -> None:
    if rc == MADI_OK:
        return
    msg = f"{what}: {last_error()}"
    if rc == MADI_ERR_ARG:
        raise ValueError(msg)
    if rc == MADI_ERR_OOM:
        raise MemoryError(msg)
    raise RuntimeError(msg)


# ----------------------------------------------------------------- torch glue

def torch_mod():
    import torch
    return torch


def require_gpu():
    """Return torch after checking a ROCm device is usable."""
    torch = torch_mod()
    if not torch.cuda.is_available():
        raise RuntimeError("a ROCm GPU is required: the MADI walk runs only in its HIP kernels")
    load_library()
    return torch


def stream_ptr() -> int:
    torch = torch_mod()
    return int(torch.cuda.current_stream().cuda_stream)


def dptr(t) -> int:
    return int(t.data_ptr())


def to_device(a: np.ndarray, dtype=None):
    torch = require_gpu()
    arr = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype=dtype))
    return torch.from_numpy(arr).to("cuda")


def empty(shape, dtype):
    torch = require_gpu()
    return torch.empty(shape, dtype=dtype, device="cuda")


def zeros(shape, dtype):
    torch = require_gpu()
    return torch.zeros(shape, dtype=dtype, device="cuda")


def host_ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


def device_info() -> dict:
    require_gpu()
    info = DeviceInfo()
    check(load_library().madi_get_device_info(C.byref(info)), "madi_get_device_info")
    return {
        "device": info.device, "cu_count": info.cu_count, "clock_khz": info.clock_khz,
        "total_mem": info.total_mem, "free_mem": info.free_mem,
        "name": info.name.decode(errors="replace"), "arch": info.arch.decode(errors="replace"),
    }
