"""shapes_amd - MI355X (gfx950) implementation of the contact-physics step
of the Haskell engine `shapes` (Physics.Engine.Main.updateWorld).

`shapes_amd.engine` mirrors the reference's engine API over the C ABI of
libshapes_gpu.so (include/shapes_gpu.h); `shapes_amd.scenes` builds the
seeded workloads.
"""
from . import scenes  # noqa: F401

__all__ = ["scenes", "engine", "build"]
