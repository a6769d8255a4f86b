"""MI355X-native MADI Monte Carlo water-diffusion walk.

Modules mirror the reference package's walk path (config, ensemble,
walker_gpu, signal, library); the walk runs in HIP kernels for gfx950
(libmadi_hip.so, C ABI in include/madi_hip.h).
"""

__all__ = ["config", "ensemble", "engine", "walker_gpu", "signal", "library"]
