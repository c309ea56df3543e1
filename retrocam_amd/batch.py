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


# Units (one sample of one pixel) a launch needs before every lane of the GPU has
# work for the whole launch: 256 CUs x 4 SIMDs x 4 waves x 64 lanes, 32 units each.
FILL_UNITS = 256 * 4 * 4 * 64 * 32
MAX_VIEWS = 4096  # views per rc_views_render


def request_units(q: ViewRequest) -> int:
    """The work of a request in units: pixels x samples."""
    return int(q.camera.width) * int(q.camera.height) * int(q.spp)


def plan_requests(reqs, fill_units: int = FILL_UNITS, max_views: int = MAX_VIEWS) -> list:
    """Launch groups for serving requests that all arrive together: shortest job
    first (ascending units, ties in arrival order), consecutive requests packed
    into one launch while the group stays at or below fill_units, larger
    requests alone. Returns a list of groups, each a list of request indices.

    Shortest-first minimises the mean (and median) latency of a set of jobs on
    one machine; packing only requests too small to fill the GPU on their own
    keeps the throughput of one launch per request for the rest."""
    order = sorted(range(len(reqs)), key=lambda i: (request_units(reqs[i]), i))
    groups, cur, cur_units = [], [], 0
    for i in order:
        u = request_units(reqs[i])
        if cur and (cur_units + u > fill_units or len(cur) >= max_views):
            groups.append(cur)
            cur, cur_units = [], 0
        cur.append(i)
        cur_units += u
    if cur:
        groups.append(cur)
    return groups


def serve(vb: ViewBatch, reqs, post_params, temporal, on_done=None, fill_units: int = FILL_UNITS) -> list:
    """Serves requests in plan_requests order: per group one rc_views_render, then
    each view's post chain (rc_views_post). on_done(request_index, view) runs
    after a request's chain is queued, while its images are still readable
    through `vb` (the next group re-lays out the views). Returns the order in
    which requests completed."""
    done = []
    for group in plan_requests(reqs, fill_units):
        vb.render([reqs[i] for i in group], reset=True)
        for v, i in enumerate(group):
            vb.post(v, post_params, temporal)
            if on_done is not None:
                on_done(i, v)
            done.append(i)
    return done
