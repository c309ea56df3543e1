"""Per-video-frame inputs of the render path, over librc_host.so (rc_temporal.cpp).

TemporalState      TemporalState.java:13-111 (AGC and auto-focus IIR filters,
                   focus hunting; the values rc_temporal and the camera's
                   focalDist take per frame)
KeyframeTrack      KeyframeTrack.java:20-178 (STEP / LINEAR / Catmull-Rom)
KeyframeTimeline   KeyframeTimeline.java:25-115 (tracks applied to their targets)
LensSettings       the keyframeable lens fields of RenderSettings
                   (RenderSettings.java:222-246, with their setter clamps)
OrbitRig           the keyframeable OrbitCamera fields (OrbitCamera.java:200-225)
PostSettings       the keyframeable post fields of RenderSettings (:229-246)
frame_dt, total_frames, canonical_time   RenderJob.java:34-41, RenderPipeline.java:150

The filter arithmetic and the spline run in the C++ host library with Java's
binary32 semantics; these classes hold the state and name things as the Java
host does.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from enum import IntEnum

import numpy as np

from . import _native as N
from .post import TemporalInputs
from .scene import CameraSpec, orbit_camera

f32 = np.float32


def _f(x) -> float:
    return float(f32(x))


def _jclamp(lo, hi, v):
    """Math.max(lo, Math.min(hi, v)) on floats (NaN propagates)."""
    v = f32(v)
    if v != v:
        return float(v)
    return float(max(f32(lo), min(f32(hi), v)))


def _jmax(lo, v):
    v = f32(v)
    if v != v:
        return float(v)
    return float(max(f32(lo), v))


@dataclass
class TemporalSettings:
    """RenderSettings fields the filters read (RenderSettings.java:28, 32, 42, 43)."""
    focusDistM: float = 2.0
    afSpeed: float = 3.0
    agcMaxGain: float = 6.0
    agcSpeed: float = 1.5

    def to_c(self) -> N.RctSettings:
        s = N.RctSettings()
        s.focusDistM, s.afSpeed, s.agcMaxGain, s.agcSpeed = (
            float(self.focusDistM), float(self.afSpeed), float(self.agcMaxGain), float(self.agcSpeed))
        return s


class TemporalState:
    """TemporalState (TemporalState.java:13-111)."""

    def __init__(self):
        self._lib = N.host_lib()
        self._s = N.RctState()
        self._lib.rct_init(C.byref(self._s))

    # current values
    @property
    def agcGain(self) -> float:
        return self._s.agcGain

    @property
    def focalDist(self) -> float:
        return self._s.focalDist

    @property
    def time(self) -> float:
        return self._s.time

    @property
    def whiteBalance(self) -> list:
        return list(self._s.whiteBalance)

    # targets (TemporalState.java:68-69)
    def set_agc_gain_target(self, target: float):
        self._s.agcGainTarget = float(target)

    def set_focal_dist_target(self, target: float):
        self._s.focalDistTarget = float(target)

    def update(self, dt: float, settings: TemporalSettings):
        self._lib.rct_update(C.byref(self._s), float(dt), C.byref(settings.to_c()))

    def reset_for_video_rendering(self, settings: TemporalSettings):
        self._lib.rct_reset_for_video(C.byref(self._s), C.byref(settings.to_c()))

    def update_for_video_frame(self, canonical_time: float, frame_dt: float, settings: TemporalSettings):
        self._lib.rct_update_for_video_frame(C.byref(self._s), float(canonical_time), float(frame_dt),
                                             C.byref(settings.to_c()))

    def inputs(self, frame_index: int) -> TemporalInputs:
        """The post chain's temporal inputs (u_time, u_agcGain, u_frameIndex)."""
        return TemporalInputs(self.time, self.agcGain, int(frame_index))

    def run_video(self, settings: TemporalSettings, fps: float, n_frames: int, focus_targets=None):
        """n_frames calls of update_for_video_frame at t = i * dt (rct_run_video);
        returns (time, agcGain, focalDist) arrays, one entry per frame."""
        out = [np.zeros(n_frames, f32) for _ in range(3)]
        ft = None
        if focus_targets is not None:
            ft = np.ascontiguousarray(focus_targets, dtype=f32)
            if ft.shape != (n_frames,):
                raise ValueError("focus_targets must hold one value per frame")
        rc = self._lib.rct_run_video(C.byref(self._s), C.byref(settings.to_c()), N.fptr(ft) if ft is not None else None,
                                     float(fps), int(n_frames), *[N.fptr(o) for o in out])
        if rc != n_frames:
            raise ValueError("rct_run_video failed")
        return tuple(out)


def frame_dt(fps: float) -> float:
    return N.host_lib().rcj_frame_dt(float(fps))


def total_frames(duration_s: float, fps: float) -> int:
    return N.host_lib().rcj_total_frames(float(duration_s), float(fps))


def canonical_time(frame: int, dt: float) -> float:
    return N.host_lib().rcj_canonical_time(int(frame), float(dt))


# ---- keyframes ------------------------------------------------------------------------

class Interpolation(IntEnum):
    STEP = N.RCK_STEP
    LINEAR = N.RCK_LINEAR
    BEZIER = N.RCK_BEZIER  # Catmull-Rom with automatic tangents

    @staticmethod
    def from_index(i: int) -> "Interpolation":
        return Interpolation(max(0, min(2, int(i))))


class TargetType(IntEnum):
    """KeyframeTarget.Type (KeyframeTarget.java:13-20)."""
    CAMERA = 0
    SCENE_OBJECT = 1
    RENDER_SETTINGS = 2


class KeyframeTrack:
    """Keyframes of one property of one target (KeyframeTrack.java:20-178)."""

    def __init__(self, target, property_name: str, target_type: TargetType):
        self.target = target
        self.property_name = property_name
        self.target_type = TargetType(target_type)
        self._lib = N.host_lib()
        self._h = self._lib.rck_new()
        if not self._h:
            raise MemoryError("rck_new")

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._lib.rck_free(h)

    def add_keyframe(self, time: float, value: float, interp=Interpolation.LINEAR) -> int:
        return self._lib.rck_add(self._h, float(time), float(value), int(interp))

    def remove_keyframe(self, index: int):
        self._lib.rck_remove(self._h, int(index))

    def keyframe_count(self) -> int:
        return self._lib.rck_count(self._h)

    def keyframes(self) -> list:
        out = []
        t, v, i = C.c_float(), C.c_float(), C.c_int32()
        for k in range(self.keyframe_count()):
            self._lib.rck_get(self._h, k, C.byref(t), C.byref(v), C.byref(i))
            out.append((t.value, v.value, Interpolation(i.value)))
        return out

    def evaluate(self, time: float, default_value: float) -> float:
        return self._lib.rck_evaluate(self._h, float(time), float(default_value))

    def evaluate_many(self, times, default_value: float) -> np.ndarray:
        ts = np.ascontiguousarray(times, dtype=f32)
        out = np.zeros_like(ts)
        self._lib.rck_evaluate_many(self._h, N.fptr(ts), ts.size, float(default_value), N.fptr(out))
        return out

    def apply_to_target(self, time: float):
        """applyToTarget (KeyframeTrack.java:130-133)."""
        cur = self.target.get_keyframeable_property(self.property_name)
        self.target.set_keyframeable_property(self.property_name, self.evaluate(time, cur))


class KeyframeTimeline:
    """KeyframeTimeline (KeyframeTimeline.java:25-115)."""

    def __init__(self, duration_seconds: float):
        self.duration_seconds = float(duration_seconds)
        self._tracks: list[KeyframeTrack] = []

    def add_track(self, target, property_name: str, target_type: TargetType) -> KeyframeTrack:
        for t in self._tracks:
            if t.target is target and t.property_name == property_name:
                return t
        if property_name not in target.keyframeable_property_names():
            raise KeyError(f"{type(target).__name__} has no keyframeable property {property_name!r}")
        t = KeyframeTrack(target, property_name, target_type)
        self._tracks.append(t)
        return t

    def remove_track(self, track: KeyframeTrack):
        self._tracks.remove(track)

    def remove_track_at(self, index: int):
        if 0 <= index < len(self._tracks):
            del self._tracks[index]

    def tracks(self) -> tuple:
        return tuple(self._tracks)

    def track_count(self) -> int:
        return len(self._tracks)

    def apply(self, time: float):
        for t in self._tracks:
            t.apply_to_target(time)

    def has_scene_object_tracks(self) -> bool:
        return any(t.target_type == TargetType.SCENE_OBJECT for t in self._tracks)

    def has_tracks_on(self, target) -> bool:
        return any(t.target is target and t.keyframe_count() > 0 for t in self._tracks)

    def is_empty(self) -> bool:
        return all(t.keyframe_count() == 0 for t in self._tracks)

    def find_track(self, target, property_name: str):
        for t in self._tracks:
            if t.target is target and t.property_name == property_name:
                return t
        return None


class _Keyframeable:
    """Keyframeable (Keyframeable.java): named float properties with setter clamps."""
    _PROPS: dict = {}  # property name -> (attribute, setter clamp)

    def keyframeable_property_names(self):
        return list(self._PROPS)

    def get_keyframeable_property(self, name: str) -> float:
        if name not in self._PROPS:
            return 0.0
        return _f(getattr(self, self._PROPS[name][0]))

    def set_keyframeable_property(self, name: str, value: float):
        if name in self._PROPS:
            attr, clamp = self._PROPS[name]
            setattr(self, attr, clamp(value))


@dataclass
class LensSettings(_Keyframeable):
    """Lens fields of RenderSettings (RenderSettings.java:222-246): focal length (mm),
    aperture (f-stop), focus distance (m, the auto-focus target), spherical aberration."""
    focalLengthMm: float = 35.0
    apertureFStop: float = 2.0
    focusDistM: float = 2.0
    saStrength: float = 0.04

    _PROPS = {
        "lens.focalLength": ("focalLengthMm", lambda v: _jmax(1.0, v)),
        "lens.aperture": ("apertureFStop", lambda v: _jmax(0.7, v)),
        "lens.focusDist": ("focusDistM", lambda v: _jmax(0.1, v)),
        "lens.saStrength": ("saStrength", _f),
    }


@dataclass
class PostSettings(_Keyframeable):
    """Post fields of RenderSettings that keyframes drive (RenderSettings.java:229-246;
    defaults :74, :78, :144, :146)."""
    exposure: float = 1.0
    colorSaturation: float = 0.95
    colorTempBias: float = 0.0
    tapeAge: float = 0.3

    _PROPS = {
        "post.exposure": ("exposure", lambda v: _jmax(0.0, v)),
        "post.saturation": ("colorSaturation", lambda v: _jmax(0.0, v)),
        "post.colorTemp": ("colorTempBias", _f),
        "post.tapeAge": ("tapeAge", lambda v: _jclamp(0.0, 1.0, v)),
    }

    def write_to(self, pp: N.RcPostParams):
        """Copies the four fields into post-chain parameters (same field names)."""
        for _, (attr, _c) in self._PROPS.items():
            setattr(pp, attr, float(getattr(self, attr)))


_PHI_LO = float(f32(0.02))
_PHI_HI = float(f32(f32(math.pi) - f32(0.02)))


@dataclass
class OrbitRig(_Keyframeable):
    """OrbitCamera's keyframeable state (OrbitCamera.java:200-225; defaults :27-33)."""
    theta: float = float(f32(math.pi / 2.0))
    phi: float = 1.2
    radius: float = 10.0
    targetX: float = 0.0
    targetY: float = 5.0
    targetZ: float = 0.0
    fovY: float = 60.0

    _PROPS = {
        "cam.theta": ("theta", _f),
        "cam.phi": ("phi", lambda v: _jclamp(_PHI_LO, _PHI_HI, v)),
        "cam.radius": ("radius", lambda v: _jmax(0.5, v)),
        "cam.targetX": ("targetX", _f),
        "cam.targetY": ("targetY", _f),
        "cam.targetZ": ("targetZ", _f),
        "cam.fovY": ("fovY", lambda v: _jclamp(5.0, 150.0, v)),
    }

    def camera(self, width: int, height: int, lens: LensSettings, focal_dist: float, **kw) -> CameraSpec:
        """The path tracer's camera for this pose: the orbit basis, the lens fields,
        and the auto-focus distance (ThinLensCamera.uploadTo reads
        TemporalState.focalDist, ThinLensCamera.java:27-58)."""
        return orbit_camera(width, height, theta=self.theta, phi=self.phi, radius=self.radius,
                            target=(self.targetX, self.targetY, self.targetZ), focalDist=float(focal_dist),
                            focalLengthMm=lens.focalLengthMm, fStop=lens.apertureFStop,
                            saStrength=lens.saStrength, **kw)
