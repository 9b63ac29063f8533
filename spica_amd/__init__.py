"""spica_amd: MI355X-native wavefront path tracer behind spica's PathIntegrator interface.

The render path is libspk.so (hand-written gfx950 HIP kernels behind a C ABI,
include/spk.h); this package is the host-side mirror of spica's scene
description and plugin interface.  There is no CPU fallback: if libspk.so is
missing, creating a GPU scene raises.
"""
from . import _abi  # noqa: F401
from .scene import Camera, Scene, conductor, dielectric, diffuse, look_at, roughconductor  # noqa: F401

__all__ = ["Camera", "Scene", "conductor", "dielectric", "diffuse", "look_at", "roughconductor"]
