"""Post-chain settings and host-side mirror of PostProcessStack.

PostSettings mirrors the RenderSettings fields read by PostProcessStack.runChain
and AtrousDenoiser.denoise (src/main/java/com/retrocam/core/RenderSettings.java:26-191,
src/main/java/com/retrocam/post/PostProcessStack.java:140-280); its defaults come
from the library (rc_post_defaults). TemporalInputs is the TemporalState/frame
counter part (time, agcGain, frameIndex).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N

VHS_PASSES = [1, 2, 3, 4, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16]


def default_params() -> N.RcPostParams:
    p = N.RcPostParams()
    N.hip_lib().rc_post_defaults(C.byref(p))
    return p


def make_params(enable=(), atrous: bool = False, iterations: int = 4, **overrides) -> N.RcPostParams:
    """RenderSettings defaults, with the listed VHS passes enabled and fields overridden."""
    p = default_params()
    for k in enable:
        p.enable[int(k)] = 1
    p.atrousEnabled = 1 if atrous else 0
    p.atrousIterations = int(iterations)
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise AttributeError(f"unknown post setting {k}")
        setattr(p, k, v)
    return p


@dataclass
class TemporalInputs:
    time: float = 0.0
    agcGain: float = 1.0
    frameIndex: int = 0

    def to_c(self) -> N.RcTemporal:
        t = N.RcTemporal()
        t.time = float(self.time)
        t.agcGain = float(self.agcGain)
        t.frameIndex = int(self.frameIndex)
        return t


@dataclass
class PostResult:
    hdr: np.ndarray           # (H*W, 4) float32, post chain output
    rgba8: np.ndarray         # (H*W, 4) uint8, display encode
    extra: dict = field(default_factory=dict)


class PostProcessStack:
    """runOnAccum / runOnImage over the C ABI (PostProcessStack.java:109-136)."""

    def __init__(self, ctx):
        self.ctx = ctx
        self._lib = ctx._lib

    def _read(self) -> PostResult:
        from .renderer import Renderer
        r = Renderer(self.ctx)
        return PostResult(r.read(N.RC_BUF_POST_HDR), r.read(N.RC_BUF_DISPLAY))

    def run_on_accum(self, params: N.RcPostParams, temporal: TemporalInputs, read: bool = True):
        t = temporal.to_c()
        self.ctx.check(self._lib.rc_post(self.ctx.h, C.byref(params), C.byref(t)))
        return self._read() if read else None

    def export_frame(self, path: str | None = None) -> np.ndarray:
        """FrameExporter.exportFrame: the displayed frame, rows top-down (H, W, 4)
        uint8, also written as a PNG to `path` when given."""
        w, h = C.c_int32(), C.c_int32()
        self.ctx.check(self._lib.rc_display_size(self.ctx.h, C.byref(w), C.byref(h)))
        out = np.empty((h.value, w.value, 4), np.uint8)
        self.ctx.check(self._lib.rc_export_frame(self.ctx.h, path.encode() if path else None,
                                                 out.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out

    def run_on_image(self, rgba: np.ndarray, width: int, height: int, params: N.RcPostParams,
                     temporal: TemporalInputs, read: bool = True):
        img = np.ascontiguousarray(rgba, np.float32).reshape(height * width, 4)
        t = temporal.to_c()
        self.ctx.check(self._lib.rc_post_image(self.ctx.h, N.fptr(img), int(width), int(height), C.byref(params),
                                               C.byref(t)))
        return self._read() if read else None
