"""Video frames through a captured frame (RenderPipeline.tickVideoFrame,
src/main/java/com/retrocam/export/RenderPipeline.java:147-172): a frame recorded
once and replayed with each video frame's camera (rc_frame_set_camera: pose,
auto-focus distance, lens) and temporal inputs must equal the frame computed
eagerly with that camera and those inputs, bit for bit, for the lane scheduler
and the megakernel, and over a keyframed clip rendered with and without the
recording."""
import numpy as np
import pytest

from retrocam_amd import _native as N
from retrocam_amd import temporal as T
from retrocam_amd.post import VHS_PASSES, PostProcessStack, TemporalInputs, make_params
from retrocam_amd.renderer import RenderError, Renderer, SceneUploader
from retrocam_amd.scene import CameraSpec
from retrocam_amd.video import VideoJob, VideoRenderer, VideoScene

pytestmark = pytest.mark.gpu

OUTS = (N.RC_BUF_ACCUM, N.RC_BUF_VARIANCE, N.RC_BUF_GBUFFER, N.RC_BUF_GALBEDO, N.RC_BUF_PRIMARY_ID,
        N.RC_BUF_POST_HDR, N.RC_BUF_DISPLAY)


def read_all(r):
    return {b: r.read(b) for b in OUTS}


def eager(r, post, cam, params, spp, pp, t):
    r.reset()
    r.render(cam, params, n_samples=spp)
    post.run_on_accum(pp, t, read=False)
    return read_all(r)


@pytest.mark.parametrize("mode", ["lanes", "megakernel"])
def test_replay_with_another_camera_equals_the_eager_frame(gpu_ctx, cornell, mode):
    w = cornell
    SceneUploader(gpu_ctx).upload(w.scene)
    r = Renderer(gpu_ctx)
    r.set_pt_mode(N.RC_PT_LANES if mode == "lanes" else N.RC_PT_MEGAKERNEL)
    post = PostProcessStack(gpu_ctx)
    pp = make_params(enable=VHS_PASSES, atrous=True, iterations=4)
    lens = T.LensSettings()
    cam_a = T.OrbitRig().camera(112, 64, lens, 2.0)
    cam_b = T.OrbitRig(theta=1.3, phi=1.05, radius=8.5, targetX=0.4).camera(
        112, 64, T.LensSettings(focalLengthMm=28.0, apertureFStop=1.4, saStrength=0.08), 4.75)
    ta, tb = TemporalInputs(0.5, 1.0, 3), TemporalInputs(1.25, 2.5, 4)
    ref_b = eager(r, post, cam_b, w.params, 6, pp, tb)
    ref_a = eager(r, post, cam_a, w.params, 6, pp, ta)
    for k in OUTS:
        if k != N.RC_BUF_PRIMARY_ID:
            assert not np.array_equal(ref_a[k], ref_b[k]), k  # the cameras see different images
    gpu_ctx.capture_begin()
    eager_calls = (r.reset, lambda: r.render(cam_a, w.params, n_samples=6),
                   lambda: post.run_on_accum(pp, ta, read=False))
    for c in eager_calls:
        c()
    gpu_ctx.capture_end()
    try:
        gpu_ctx.set_frame_camera(cam_b)
        gpu_ctx.replay(tb)
        got = read_all(r)
        for k in OUTS:
            assert np.array_equal(got[k], ref_b[k]), k
        gpu_ctx.set_frame_camera(None)  # back to the recorded camera
        gpu_ctx.replay(ta)
        got = read_all(r)
        for k in OUTS:
            assert np.array_equal(got[k], ref_a[k]), k
        with pytest.raises(RenderError) as e:
            gpu_ctx.set_frame_camera(CameraSpec(**{**cam_b.__dict__, "width": 96}))
        assert e.value.code == N.RC_ERR_INVALID_ARG
    finally:
        gpu_ctx.release_frame()
    with pytest.raises(RenderError) as e:
        gpu_ctx.set_frame_camera(cam_b)
    assert e.value.code == N.RC_ERR_STATE


def test_sppm_gather_follows_the_replay_camera(gpu_ctx):
    from retrocam_amd import workloads
    from retrocam_amd.sppm import SPPMManager
    w = workloads.sppm_caustics(width=128, height=72, spp=3, photons=100_000, hash_size=1 << 17)
    SceneUploader(gpu_ctx).upload(w.scene)
    r = Renderer(gpu_ctx)
    r.set_pt_mode(N.RC_PT_LANES)
    s = SPPMManager(gpu_ctx, init_radius=w.extra["init_radius"])
    cam_b = T.OrbitRig(theta=1.2, phi=1.15, radius=9.0).camera(128, 72, T.LensSettings(), 3.0, pinhole=True)

    def frame(cam):
        r.reset()
        s.reset(w.extra["init_radius"])
        for _ in range(3):
            s.trace_photons(w.extra["photons"], w.extra["hash_size"])
            r.render(cam, w.params, n_samples=1)
            s.gather_radiance(cam)
            s.update_radius(w.extra["alpha"])

    bufs = (N.RC_BUF_ACCUM, N.RC_BUF_VARIANCE, N.RC_BUF_GBUFFER, N.RC_BUF_GALBEDO, N.RC_BUF_PRIMARY_ID)
    frame(cam_b)
    ref = {b: r.read(b) for b in bufs}
    frame(w.camera)
    gpu_ctx.capture_begin()
    frame(w.camera)
    gpu_ctx.capture_end()
    try:
        gpu_ctx.set_frame_camera(cam_b)
        gpu_ctx.replay()
        got = {b: r.read(b) for b in bufs}
    finally:
        gpu_ctx.release_frame()
    for k in (N.RC_BUF_GBUFFER, N.RC_BUF_GALBEDO, N.RC_BUF_PRIMARY_ID):
        assert np.array_equal(got[k], ref[k]), k
    # the gather adds photons in atomic insertion order: compare by relMSE
    a, b = got[N.RC_BUF_ACCUM][:, :3].astype(np.float64), ref[N.RC_BUF_ACCUM][:, :3].astype(np.float64)
    assert float(np.mean((a - b) ** 2 / (b * b + 0.01))) <= 1e-6


def test_keyframed_clip_replayed_equals_eager(gpu_ctx, cornell, tmp_path):
    w = cornell
    SceneUploader(gpu_ctx).upload(w.scene)
    r = Renderer(gpu_ctx)
    r.set_pt_mode(N.RC_PT_LANES)
    post = PostProcessStack(gpu_ctx)

    def scene():
        s = VideoScene()
        tl = T.KeyframeTimeline(0.5)
        th = tl.add_track(s.rig, "cam.theta", T.TargetType.CAMERA)
        th.add_keyframe(0.0, 1.45)
        th.add_keyframe(0.5, 1.75)
        fd = tl.add_track(s.lens, "lens.focusDist", T.TargetType.RENDER_SETTINGS)
        fd.add_keyframe(0.0, 2.0, T.Interpolation.BEZIER)
        fd.add_keyframe(0.25, 6.0, T.Interpolation.BEZIER)
        fd.add_keyframe(0.5, 3.0, T.Interpolation.BEZIER)
        s.timeline = tl
        return s

    job = VideoJob(96, 54, fps=12.0, duration_seconds=0.5, samples_per_frame=4,
                   output_path=str(tmp_path / "clip"), first_frame_index=10)
    pp = make_params(enable=VHS_PASSES, atrous=True, iterations=3)
    graph = VideoRenderer(gpu_ctx, r, post, w.params, pp, use_graph=True).render(job, scene(), keep_images=True)
    plain = VideoRenderer(gpu_ctx, r, post, w.params, pp, use_graph=False).render(job, scene(), keep_images=True)
    assert len(graph) == len(plain) == 6
    assert [g.replayed for g in graph] == [False] + [True] * 5
    for g, p in zip(graph, plain):
        assert (g.time, g.agc_gain, g.focal_dist) == (p.time, p.agc_gain, p.focal_dist)
        assert np.array_equal(g.image, p.image), g.index
    # the focus pull and the pan change the frames; every frame was written
    assert len({g.focal_dist for g in graph}) == 6
    assert not np.array_equal(graph[0].image, graph[5].image)
    for g in graph:
        with open(g.path, "rb") as f:
            assert f.read(8) == b"\x89PNG\r\n\x1a\n"


def test_scene_object_keyframes_rebuild_the_scene(gpu_ctx):
    # RenderPipeline re-uploads the scene per frame when a track moves a scene
    # object; each frame must equal an eager render of the editor at that time.
    from retrocam_amd.scene import PtParams
    from retrocam_amd.scene_file import SceneEditor
    ed = SceneEditor.create_default()
    block = ed.objects[7]  # Metal Block
    tl = T.KeyframeTimeline(0.25)
    tr = tl.add_track(block, "pos.x", T.TargetType.SCENE_OBJECT)
    tr.add_keyframe(0.0, 1.5)
    tr.add_keyframe(0.25, -0.5)
    scene = VideoScene(editor=ed, timeline=tl, camera_kw={"pinhole": True})
    job = VideoJob(64, 36, fps=12.0, duration_seconds=0.25, samples_per_frame=2)
    r = Renderer(gpu_ctx)
    r.set_pt_mode(N.RC_PT_LANES)
    post = PostProcessStack(gpu_ctx)
    pp = make_params(enable=VHS_PASSES, atrous=True, iterations=2)
    params = PtParams(1, 1.0, 6)
    recs = VideoRenderer(gpu_ctx, r, post, params, pp).render(job, scene, keep_images=True)
    assert len(recs) == 3 and not any(g.replayed for g in recs)
    assert not np.array_equal(recs[0].image, recs[2].image)
    for rec in (recs[0], recs[2]):
        ref_ed = SceneEditor.create_default()
        ref_ed.objects[7].px = tr.evaluate(rec.time, 1.5)
        SceneUploader(gpu_ctx).upload(ref_ed.build())
        cam = T.OrbitRig().camera(64, 36, T.LensSettings(), rec.focal_dist, pinhole=True)
        r.reset()
        r.render(cam, params, n_samples=2)
        post.run_on_accum(pp, TemporalInputs(rec.time, rec.agc_gain, rec.index), read=False)
        assert np.array_equal(post.export_frame(), rec.image), rec.index
