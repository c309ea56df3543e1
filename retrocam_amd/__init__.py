"""retrocam_amd — MI355X-native render hot path of RetroCam.

The product is libretrocam_hip.so (C ABI: include/retrocam_hip.h) with
hand-written gfx950 kernels; this package holds its sources (csrc/), build
recipe (build.py), ctypes bindings (_native.py), the host-side mirror of the
reference's Renderer/SceneUploader classes (renderer.py) and seeded workloads.
"""
from . import _native  # noqa: F401
from .scene import CameraSpec, Material, PtParams, SceneArrays, SceneBuilder, lookat_camera, orbit_camera  # noqa: F401

__all__ = ["CameraSpec", "Material", "PtParams", "SceneArrays", "SceneBuilder", "lookat_camera", "orbit_camera"]
