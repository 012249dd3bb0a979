"""fledge_amd — MI355X (gfx950) implementation of fledge's ray-tracing hot
path: radix-tree LBVH build, closest-hit traversal with the watertight
Pluecker test, the Halton sampler and the volume-only path integrator with
its path-replay adjoint, behind a C ABI (include/fledge_hip.h).

The Python package only wraps the native library (ctypes + torch device
buffers); it has no compute path of its own.
"""
from . import _native  # noqa: F401

__all__ = ["_native"]
