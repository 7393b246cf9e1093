"""MI355X (gfx950) dense RGB-D reconstruction: HIP kernels behind a C ABI
(include/fusion_amd.h) with a thin Python host layer (fusion_amd.api)."""
from .native import FusionError, lib, scene_lib  # noqa: F401

__all__ = ["FusionError", "lib", "scene_lib"]
