"""Batched views over the C ABI (rc_views_* in include/retrocam_hip.h).

Project-new: the reference renders one view per Renderer
(src/main/java/com/retrocam/core/Renderer.java:72-116). A ViewBatch traces many
requests — each with its own scene slot, camera/resolution, switches and
sample count — in one persistent launch; each view's images equal those of a
single-view render of the same request.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .renderer import Renderer
from .scene import CameraSpec, PtParams, SceneArrays


@dataclass
class ViewRequest:
    scene: int            # scene slot
    camera: CameraSpec
    params: PtParams
    spp: int


class ViewBatch:
    def __init__(self, ctx):
        self.ctx = ctx
        self._lib = ctx._lib
        self.sizes = []

    def upload_scene(self, slot: int, sc: SceneArrays):
        lights = sc.lights if sc.nLights > 0 else np.zeros((1, 20), np.float32)
        self.ctx.check(self._lib.rc_upload_scene_slot(
            self.ctx.h, int(slot), N.fptr(np.ascontiguousarray(sc.tris)), sc.nTris,
            N.fptr(np.ascontiguousarray(sc.nodes)), sc.nNodes,
            N.fptr(np.ascontiguousarray(sc.mats)), sc.mats.shape[0],
            N.fptr(np.ascontiguousarray(lights)), sc.nLights, float(sc.totalLightPower)))

    def render(self, reqs, reset: bool = True, fp64: bool = False):
        arr = (N.RcView * len(reqs))()
        for i, q in enumerate(reqs):
            arr[i].scene = int(q.scene)
            arr[i].nSamples = int(q.spp)
            arr[i].cam = q.camera.to_c()
            arr[i].params = q.params.to_c()
        flags = (N.RC_VIEWS_RESET if reset else 0) | (N.RC_VIEWS_FP64 if fp64 else 0)
        self.ctx.check(self._lib.rc_views_render(self.ctx.h, arr, len(reqs), flags))
        self.sizes = [(q.camera.width, q.camera.height) for q in reqs]

    def post(self, view: int, params: N.RcPostParams, temporal):
        t = temporal.to_c()
        self.ctx.check(self._lib.rc_views_post(self.ctx.h, int(view), C.byref(params), C.byref(t)))

    def total_samples(self, view: int) -> int:
        return int(self._lib.rc_views_total_samples(self.ctx.h, int(view)))

    def read(self, view: int, which: int) -> np.ndarray:
        dt, ch = Renderer._DT[which]
        nbytes = self._lib.rc_views_buffer_bytes(self.ctx.h, int(view), int(which))
        n = nbytes // (np.dtype(dt).itemsize * ch)
        out = np.empty((n, ch) if ch > 1 else n, dtype=dt)
        self.ctx.check(self._lib.rc_views_read(self.ctx.h, int(view), int(which), out.ctypes.data_as(C.c_void_p),
                                               nbytes))
        return out
