"""Host-side mirror of the reference's render classes over the C ABI.

  RenderContext   owns an rc_ctx (one HIP stream, all device buffers)
  SceneUploader   SceneUploader.upload/getLightCount/getTotalLightPower
                  (src/main/java/com/retrocam/scene/SceneUploader.java:54-122)
  Renderer        Renderer.render/reset/getTotalSamples + texture readback
                  (src/main/java/com/retrocam/core/Renderer.java:72-138)

Every call goes to libretrocam_hip.so; failures raise RenderError with the
library's message (the reference throws RuntimeException, ShaderProgram.java:98-119).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .scene import CameraSpec, PtParams, SceneArrays


class RenderError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[rc status {code}] {msg}")
        self.code = code


class RenderContext:
    def __init__(self, device: int = 0):
        self._lib = N.hip_lib()
        h = C.c_void_p()
        rc = self._lib.rc_create(device, C.byref(h))
        if rc != N.RC_OK:
            raise RenderError(rc, self._lib.rc_last_error(None).decode())
        self.h = h

    def check(self, rc: int):
        if rc != N.RC_OK:
            raise RenderError(rc, self._lib.rc_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self._lib.rc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def stream(self) -> int:
        return self._lib.rc_stream(self.h)

    def capture_begin(self):
        """Starts recording the device work of the following calls into a frame
        graph (rc_frame_capture_begin); it must begin with a reset."""
        self.check(self._lib.rc_frame_capture_begin(self.h))

    def capture_end(self):
        self.check(self._lib.rc_frame_capture_end(self.h))

    def replay(self, temporal=None):
        """Launches the recorded frame; temporal (TemporalInputs or RcTemporal)
        replaces the post chains' time inputs, None keeps the recorded ones."""
        t = None
        if temporal is not None:
            t = temporal.to_c() if hasattr(temporal, "to_c") else temporal
        self.check(self._lib.rc_frame_replay(self.h, C.byref(t) if t is not None else None))

    def set_frame_camera(self, camera=None):
        """rc_frame_set_camera: the camera (CameraSpec) later replays use; None
        restores the recorded one."""
        if camera is None:
            self.check(self._lib.rc_frame_set_camera(self.h, None))
        else:
            c = camera.to_c()
            self.check(self._lib.rc_frame_set_camera(self.h, C.byref(c)))

    def release_frame(self):
        self.check(self._lib.rc_frame_release(self.h))

    def sync(self):
        self.check(self._lib.rc_sync(self.h))


class SceneUploader:
    def __init__(self, ctx: RenderContext):
        self.ctx = ctx
        self.scene = None

    def upload(self, sc: SceneArrays) -> "SceneUploader":
        lights = sc.lights if sc.nLights > 0 else np.zeros((1, 20), np.float32)
        self.ctx.check(self.ctx._lib.rc_upload_scene(
            self.ctx.h, N.fptr(np.ascontiguousarray(sc.tris)), sc.nTris,
            N.fptr(np.ascontiguousarray(sc.nodes)), sc.nNodes,
            N.fptr(np.ascontiguousarray(sc.mats)), sc.mats.shape[0],
            N.fptr(np.ascontiguousarray(lights)), sc.nLights, float(sc.totalLightPower)))
        self.scene = sc
        return self

    def get_light_count(self) -> int:
        return self.scene.nLights if self.scene else 0

    def get_total_light_power(self) -> float:
        return self.scene.totalLightPower if self.scene else 0.0


class Renderer:
    def __init__(self, ctx: RenderContext):
        self.ctx = ctx
        self._lib = ctx._lib
        self.width = 0
        self.height = 0

    def render(self, camera: CameraSpec, params: PtParams, n_samples: int = 1):
        cam = camera.to_c()
        p = params.to_c()
        if n_samples == 1:
            self.ctx.check(self._lib.rc_render_pass(self.ctx.h, C.byref(cam), C.byref(p)))
        else:
            self.ctx.check(self._lib.rc_render_passes(self.ctx.h, C.byref(cam), C.byref(p), int(n_samples)))
        self.width, self.height = camera.width, camera.height

    def reset(self):
        self.ctx.check(self._lib.rc_reset(self.ctx.h))

    def get_total_samples(self) -> int:
        return self._lib.rc_total_samples(self.ctx.h)

    def set_accum_mode(self, mode: int):
        self.ctx.check(self._lib.rc_set_accum_mode(self.ctx.h, int(mode)))

    def set_pt_mode(self, mode: int):
        self.ctx.check(self._lib.rc_set_pt_mode(self.ctx.h, int(mode)))

    TUNING_FIELDS = ("laneThreshold", "laneVariant", "laneBufferMiB", "laneTileWidth", "ptMode")

    def get_tuning(self) -> dict:
        t = N.RcTuning()
        self.ctx.check(self._lib.rc_get_tuning(self.ctx.h, C.byref(t)))
        return {k: int(getattr(t, k)) for k in self.TUNING_FIELDS}

    def autotune(self, camera: CameraSpec, params: PtParams, samples: int = 8) -> dict:
        """rc_autotune: keeps the fastest lane-scheduler tuning for this scene and
        resolution; returns it with its milliseconds per sample. Resets the
        accumulation."""
        cam = camera.to_c()
        p = params.to_c()
        t = N.RcTuning()
        ms = C.c_float()
        self.ctx.check(self._lib.rc_autotune(self.ctx.h, C.byref(cam), C.byref(p), int(samples), C.byref(t),
                                             C.byref(ms)))
        self.width, self.height = camera.width, camera.height
        out = {k: int(getattr(t, k)) for k in self.TUNING_FIELDS}
        out["ms_per_sample"] = float(ms.value)
        return out

    def set_tuning(self, **kw):
        """Sets any of laneThreshold, laneVariant, laneBufferMiB (the rest stay)."""
        cur = self.get_tuning()
        unknown = set(kw) - set(cur)
        if unknown:
            raise ValueError(f"unknown tuning fields {sorted(unknown)}")
        cur.update({k: int(v) for k, v in kw.items()})
        t = N.RcTuning(*(cur[k] for k in self.TUNING_FIELDS))
        self.ctx.check(self._lib.rc_set_tuning(self.ctx.h, C.byref(t)))

    _DT = {N.RC_BUF_ACCUM: (np.float32, 4), N.RC_BUF_GBUFFER: (np.uint16, 4), N.RC_BUF_GALBEDO: (np.uint16, 4),
           N.RC_BUF_VARIANCE: (np.float32, 1), N.RC_BUF_PRIMARY_ID: (np.int32, 1),
           N.RC_BUF_GBUFFER_ID: (np.int32, 1), N.RC_BUF_ACCUM64: (np.float64, 4),
           N.RC_BUF_POST_HDR: (np.float32, 4), N.RC_BUF_DISPLAY: (np.uint8, 4)}

    def read(self, which: int) -> np.ndarray:
        dt, ch = self._DT[which]
        nbytes = self._lib.rc_buffer_bytes(self.ctx.h, which)
        n = nbytes // (np.dtype(dt).itemsize * ch)
        out = np.empty((n, ch) if ch > 1 else n, dtype=dt)
        self.ctx.check(self._lib.rc_read(self.ctx.h, which, out.ctypes.data_as(C.c_void_p), nbytes))
        return out

    def debug_counters(self, enable: bool, n: int = 12) -> np.ndarray:
        out = np.zeros(n, np.uint64)
        self.ctx.check(self._lib.rc_debug_counters(self.ctx.h, 1 if enable else 0,
                                                   out.ctypes.data_as(C.POINTER(C.c_uint64)), n))
        return out


def debug_pcg(ctx: RenderContext, xyf: np.ndarray, count: int) -> np.ndarray:
    """Device initSeed + `count` pcg outputs per (x, y, frame) row of xyf."""
    xyf = np.ascontiguousarray(xyf, np.int32).reshape(-1, 3)
    out = np.zeros((xyf.shape[0], count + 1), np.uint32)
    ctx.check(ctx._lib.rc_debug_pcg(ctx.h, xyf.ctypes.data_as(C.POINTER(C.c_int32)), xyf.shape[0], int(count),
                                    out.ctypes.data_as(C.POINTER(C.c_uint32))))
    return out


def debug_rays(ctx: RenderContext, ro: np.ndarray, rd: np.ndarray, max_dist: np.ndarray):
    """Device closest-hit (tri, t) and shadow (unoccluded) queries for n rays."""
    n = ro.shape[0]
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = ro
    rays[:, 3] = max_dist
    rays[:, 4:7] = rd
    tri = np.zeros(n, np.int32)
    t = np.zeros(n, np.float32)
    un = np.zeros(n, np.int32)
    ctx.check(ctx._lib.rc_debug_rays(ctx.h, N.fptr(rays), n, tri.ctypes.data_as(C.POINTER(C.c_int32)), N.fptr(t),
                                     un.ctypes.data_as(C.POINTER(C.c_int32))))
    return tri, t, un


COUNTER_FIELDS = ["rays", "shadowRays", "nodePops", "childTests", "triTests", "hitUpdates", "shadeHits",
                  "shadowPops", "shadowTriTests", "lightScans", "lightReads", "paths"]
