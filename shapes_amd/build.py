"""Build the gfx950 shared library libshapes_gpu.so in-tree with hipcc.

Device code is compiled without floating-point contraction (the drop-in
solver must round every multiply and add separately, like the reference's
SSE2 doubles) and without packed-FP32 instructions.  The sources compile
in parallel to objects under shapes_amd/_build, which are then linked.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libshapes_gpu.so")
SOURCES = ["shp_engine.hip", "shp_solver.hip", "shp_colored.hip", "shp_mlp.hip", "shp_grid.hip"]
HEADERS = ["shp_engine.h", "shp_geom.h", "shp_trig.h", "shp_solve_common.h", "shp_mlp.h",
           os.path.join("..", "..", "include", "shapes_gpu.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = ["--offload-arch=gfx950"]
FLAGS = ARCH + [
    "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def build(verbose: bool = False) -> str:
    """Compile every HIP source into libshapes_gpu.so (always rebuilds)."""
    os.makedirs(OBJ, exist_ok=True)
    procs = []
    for src in SOURCES:
        obj = os.path.join(OBJ, src.replace(".hip", ".o"))
        cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, obj, subprocess.Popen(cmd)))
    failed = [src for src, _, p in procs if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, f"hipcc {' '.join(failed)}")
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + ARCH + ["-shared", "-fPIC", "-o", tmp] + [o for _, o, _ in procs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
