"""seevr_amd - MI355X (gfx950) implementation of seeVR's per-substep physics
(physics.comp) and LBVH build (bvh.comp) behind the C ABI of include/seevr.h.

The product is seevr_amd/libseevr_hip.so (build: python -m seevr_amd.build);
this package adds a thin Python view of that ABI for tests and benchmarks.
"""
from .api import Batch, Context, Scene, caps_for, default_caps, default_params  # noqa: F401
from ._lib import LIB_PATH, SeevrError  # noqa: F401

__all__ = ["Batch", "Context", "Scene", "caps_for", "default_caps", "default_params", "LIB_PATH", "SeevrError"]
