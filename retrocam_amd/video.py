"""Video rendering: RenderPipeline's video mode
(src/main/java/com/retrocam/export/RenderPipeline.java:147-172 per frame,
:209-239 accumulateFrame) over the C ABI.

Per frame i, at t = i * dt (dt = 1 / fps):
  1. keyframes are applied at t (KeyframeTimeline.apply);
  2. the auto-focus target is the lens focus distance, and TemporalState
     advances one frame (updateForVideoFrame);
  3. the frame is accumulated (reset, samplesPerFrame samples of path tracing,
     with SPPM photon passes when enabled, post chain, display) with the camera
     at the new pose and focal distance and the new temporal inputs;
  4. the displayed frame is exported as frame_%05d.png (RenderJob.getFramePath).

Step 3 is recorded once into a captured frame (rc_frame_capture_*) and then
replayed with each frame's camera (rc_frame_set_camera) and temporal inputs
(rc_frame_replay): one graph launch per video frame. Frame 0 runs eagerly (it
sizes every buffer the recording uses), frame 1 is recorded and replayed.
Keyframes on post settings or scene objects change what the frame computes
beyond the camera and the temporal inputs; with such tracks every frame runs
eagerly, and with scene-object tracks the scene editor's state is rebuilt and
uploaded per frame (RenderPipeline.applyKeyframes re-uploads the scene when
KeyframeTimeline.hasSceneObjectTracks, KeyframeTimeline.java:88-96).
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .post import PostProcessStack, TemporalInputs
from .renderer import RenderContext, RenderError, Renderer, SceneUploader
from .scene import CameraSpec, PtParams
from .temporal import (KeyframeTimeline, LensSettings, OrbitRig, PostSettings, TemporalSettings, TemporalState,
                       canonical_time, frame_dt, total_frames)


@dataclass
class VideoJob:
    """RenderJob fields of a video render (RenderJob.java:14-48)."""
    width: int
    height: int
    fps: float = 29.97
    duration_seconds: float = 1.0
    samples_per_frame: int = 8
    output_path: str | None = None   # frames go to <output_path>_frames/frame_%05d.png
    first_frame_index: int = 0       # settings.frameIndex at the first frame (post-chain noise seeds)

    def total_frames(self) -> int:
        return total_frames(self.duration_seconds, self.fps)

    def frame_dt(self) -> float:
        return frame_dt(self.fps)

    def frame_path(self, i: int) -> str | None:
        if self.output_path is None:
            return None
        return f"{self.output_path}_frames/frame_{i:05d}.png"


@dataclass
class SppmSettings:
    """The SPPM pass of accumulateFrame (RenderPipeline.java:211-226)."""
    manager: object                  # SPPMManager
    init_radius: float
    photons: int
    hash_size: int
    alpha: float


@dataclass
class FrameRecord:
    index: int
    time: float
    agc_gain: float
    focal_dist: float
    replayed: bool
    path: str | None = None
    image: np.ndarray | None = None  # (H, W, 4) uint8 top-down, when kept


@dataclass
class VideoScene:
    """What a video frame is rendered from: camera rig, lens, post settings and
    the keyframes that animate them."""
    rig: OrbitRig = field(default_factory=OrbitRig)
    lens: LensSettings = field(default_factory=LensSettings)
    post: PostSettings = field(default_factory=PostSettings)
    temporal: TemporalSettings = field(default_factory=TemporalSettings)
    timeline: KeyframeTimeline | None = None
    camera_kw: dict = field(default_factory=dict)  # further CameraSpec fields (blades, lcaDelta, pinhole)
    editor: object = None  # SceneEditor whose objects scene-object tracks animate


class VideoRenderer:
    def __init__(self, ctx: RenderContext, renderer: Renderer, post: PostProcessStack, pt_params: PtParams,
                 post_params: N.RcPostParams, sppm: SppmSettings | None = None, use_graph: bool = True):
        self.ctx = ctx
        self.renderer = renderer
        self.post = post
        self.pt_params = pt_params
        self.post_params = post_params
        self.sppm = sppm
        self.use_graph = use_graph

    def accumulate_frame(self, cam: CameraSpec, t: TemporalInputs, spp: int):
        """accumulateFrame (RenderPipeline.java:209-239)."""
        r = self.renderer
        r.reset()
        s = self.sppm
        if s is None:
            r.render(cam, self.pt_params, n_samples=spp)
        else:
            s.manager.reset(s.init_radius)
            for _ in range(spp):
                s.manager.trace_photons(s.photons, s.hash_size)
                r.render(cam, self.pt_params, n_samples=1)
                s.manager.gather_radiance(cam)
                s.manager.update_radius(s.alpha)
        self.post.run_on_accum(self.post_params, t, read=False)

    def _abandon_capture(self):
        for step in (self.ctx.capture_end, self.ctx.release_frame):
            try:
                step()
            except RenderError:
                pass

    def render(self, job: VideoJob, scene: VideoScene, keep_images: bool = False, on_frame=None) -> list:
        """Renders job.total_frames() frames; returns a FrameRecord per frame."""
        n = job.total_frames()
        dt = job.frame_dt()
        timeline = scene.timeline
        post_animated = timeline is not None and timeline.has_tracks_on(scene.post)
        objects_animated = timeline is not None and timeline.has_scene_object_tracks()
        if objects_animated and scene.editor is None:
            raise ValueError("scene-object keyframes need VideoScene.editor (the SceneEditor they animate)")
        graph_ok = self.use_graph and not post_animated and not objects_animated
        if job.output_path is not None:
            os.makedirs(os.path.dirname(job.frame_path(0)) or ".", exist_ok=True)
        temporal = TemporalState()
        temporal.reset_for_video_rendering(scene.temporal)
        records = []
        captured = False
        try:
            for i in range(n):
                t = canonical_time(i, dt)
                if timeline is not None:
                    timeline.apply(t)
                    if post_animated:
                        scene.post.write_to(self.post_params)
                    if objects_animated:
                        SceneUploader(self.ctx).upload(scene.editor.build())
                temporal.set_focal_dist_target(scene.lens.focusDistM)
                temporal.update_for_video_frame(t, dt, scene.temporal)
                cam = scene.rig.camera(job.width, job.height, scene.lens, temporal.focalDist, **scene.camera_kw)
                tin = temporal.inputs(job.first_frame_index + i)
                replayed = False
                if graph_ok and i >= 1:
                    if not captured:
                        self.ctx.capture_begin()
                        try:
                            self.accumulate_frame(cam, tin, job.samples_per_frame)
                        except BaseException:
                            self._abandon_capture()
                            raise
                        self.ctx.capture_end()
                        captured = True
                    self.ctx.set_frame_camera(cam)
                    self.ctx.replay(tin)
                    replayed = True
                else:
                    self.accumulate_frame(cam, tin, job.samples_per_frame)
                path = job.frame_path(i)
                img = None
                if path is not None or keep_images:
                    img = self.post.export_frame(path)
                rec = FrameRecord(i, t, temporal.agcGain, temporal.focalDist, replayed, path,
                                  img if keep_images else None)
                records.append(rec)
                if on_frame is not None:
                    on_frame(rec)
        finally:
            if captured:
                self.ctx.release_frame()
        return records
