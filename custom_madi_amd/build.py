"""Build libmadi_hip.so for gfx950 in-tree with hipcc.

The FP64 reference-semantics kernels (parity.hip) and the FP64 library
matcher (match.hip) are compiled without FP contraction, so their arithmetic
order is the source's (the matcher fuses only where it calls fma()); the
production kernels and the ABI use hipcc's defaults.  Every translation unit disables
packed FP32 instructions.
"""

from __future__ import annotations

import os
import sys

# Run as a script, this directory would shadow the standard library's
# ``signal`` module with the package's signal.py.
if sys.path and os.path.abspath(sys.path[0]) == os.path.dirname(os.path.abspath(__file__)):
    sys.path.pop(0)

import shutil  # noqa: E402
import subprocess  # noqa: E402
from pathlib import Path  # noqa: E402

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libmadi_hip.so"
ARCH = os.environ.get("MADI_OFFLOAD_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
UNITS = {
    "parity.hip": ["-ffp-contract=off"],
    "prod.hip": [],
    "abi.hip": [],
    "match.hip": ["-ffp-contract=off"],
    "geom_host.cpp": ["-ffp-contract=off"],
}


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(path).exists():
        raise RuntimeError("hipcc not found: the MADI HIP library cannot be built")
    return path


def build(verbose: bool = False) -> Path:
    """Compile every unit and link the shared library; returns its path."""
    out_dir = HERE / "_build"
    out_dir.mkdir(exist_ok=True)
    objs = []
    for unit, extra in UNITS.items():
        obj = out_dir / (unit + ".o")
        cmd = [hipcc(), *COMMON, *extra, "-c", str(CSRC / unit), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(str(obj))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
