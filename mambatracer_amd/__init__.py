"""mambatracer_amd: MI355X (gfx950) wavefront path-tracing core.

The product is libmambart.so (HIP kernels behind the C ABI in
include/mambart.h); this package builds it (build.py), binds it (_lib.py)
and mirrors the reference's host drivers for the hot path (api.py).
"""
from . import _lib  # noqa: F401

__all__ = ["_lib"]
