"""ajisai_render_amd — MI355X-native (gfx950) path-tracing hot path of Ajisai.

PLOC BVH construction, exact-order ray/BVH traversal and the wavefront MIS
path tracer run as hand-written HIP kernels in libajisai_gpu.so behind the
C-ABI of include/ajisai_gpu.h. This package holds the Python host mirror of
the reference's Aggregate / Renderer interfaces and the seeded scene
generators used by tests and the benchmark.
"""
from . import _capi, scenes  # noqa: F401
from ._capi import AjgError, load  # noqa: F401

__all__ = ["AjgError", "load", "scenes"]
