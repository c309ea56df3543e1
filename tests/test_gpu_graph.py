"""Captured frames (rc_frame_* through the C ABI): the per-frame chain of
RenderPipeline.accumulateFrame (src/main/java/com/retrocam/export/RenderPipeline.java:209-239)
recorded once into a HIP graph and replayed. Every replay must reproduce the
eagerly computed frame bit for bit (path trace, SPPM gather order aside, post
chain with a-trous or the U-Net, VHS passes, display), also after unrelated
eager work in between, and with new temporal inputs. Calls that cannot be
recorded must be refused without disturbing the context."""
import numpy as np
import pytest

from retrocam_amd import _native as N
from retrocam_amd import workloads
from retrocam_amd.batch import ViewBatch, ViewRequest
from retrocam_amd.denoiser import NetDenoiser
from retrocam_amd.post import VHS_PASSES, PostProcessStack, TemporalInputs, make_params
from retrocam_amd.renderer import RenderError, Renderer, SceneUploader
from retrocam_amd.scene import CameraSpec
from retrocam_amd.sppm import SPPMManager

pytestmark = pytest.mark.gpu

OUTS = (N.RC_BUF_ACCUM, N.RC_BUF_VARIANCE, N.RC_BUF_GBUFFER, N.RC_BUF_GALBEDO, N.RC_BUF_PRIMARY_ID,
        N.RC_BUF_POST_HDR, N.RC_BUF_DISPLAY)


def read_all(r):
    return {b: r.read(b) for b in OUTS}


def assert_same(a, b, what):
    for k in a:
        assert np.array_equal(a[k], b[k]), (what, k)


def plain_frame(ctx, r, post, cam, params, spp, pp, t):
    r.reset()
    r.render(cam, params, n_samples=spp)
    post.run_on_accum(pp, t, read=False)


@pytest.mark.parametrize("denoiser", ["atrous", "net"])
def test_replays_equal_the_eager_frame(gpu_ctx, cornell, denoiser):
    w = cornell
    cam = CameraSpec(**{**w.camera.__dict__, "width": 128, "height": 80})
    SceneUploader(gpu_ctx).upload(w.scene)
    r = Renderer(gpu_ctx)
    r.set_pt_mode(N.RC_PT_LANES)
    post = PostProcessStack(gpu_ctx)
    net = NetDenoiser(gpu_ctx).load()
    net.enable(denoiser == "net")
    pp = make_params(enable=VHS_PASSES, atrous=True, iterations=5)
    t0, t1 = TemporalInputs(0.25, 1.0, 7), TemporalInputs(3.5, 1.3, 42)
    plain_frame(gpu_ctx, r, post, cam, w.params, 8, pp, t1)
    ref1 = read_all(r)
    plain_frame(gpu_ctx, r, post, cam, w.params, 8, pp, t0)
    ref0 = read_all(r)
    gpu_ctx.capture_begin()
    plain_frame(gpu_ctx, r, post, cam, w.params, 8, pp, t0)
    gpu_ctx.capture_end()
    for rep in range(3):
        # unrelated eager work between replays: another sample count and time
        plain_frame(gpu_ctx, r, post, cam, w.params, 3, pp, t1)
        gpu_ctx.sync()
        gpu_ctx.replay()
        assert r.get_total_samples() == 8
        assert_same(read_all(r), ref0, f"replay {rep}")
    gpu_ctx.replay(t1)
    assert_same(read_all(r), ref1, "replay with new temporal inputs")
    gpu_ctx.replay()
    assert_same(read_all(r), ref0, "replay with the recorded temporal inputs again")
    gpu_ctx.release_frame()
    net.enable(False)


def test_replayed_sppm_chain(gpu_ctx):
    # accumulateFrame with SPPM: per sample photon trace, path trace, gather,
    # radius update; the gather sums photons in atomic insertion order, so runs
    # are compared by relMSE (the oracle tests pin the arithmetic bit for bit).
    w = workloads.sppm_caustics(width=160, height=90, spp=4, photons=200_000, hash_size=1 << 18)
    SceneUploader(gpu_ctx).upload(w.scene)
    r = Renderer(gpu_ctx)
    s = SPPMManager(gpu_ctx, init_radius=w.extra["init_radius"], fp64_radius=True)
    post = PostProcessStack(gpu_ctx)
    pp = make_params(enable=VHS_PASSES, atrous=True, iterations=3)
    t = TemporalInputs(1.0, 1.0, 3)

    def frame():
        r.reset()
        s.reset(w.extra["init_radius"])
        for _ in range(4):
            s.trace_photons(w.extra["photons"], w.extra["hash_size"])
            r.render(w.camera, w.params, n_samples=1)
            s.gather_radiance(w.camera)
            s.update_radius(w.extra["alpha"])
        post.run_on_accum(pp, t, read=False)

    frame()
    ref = read_all(r)
    ref_state = s.state()
    gpu_ctx.capture_begin()
    frame()
    gpu_ctx.capture_end()
    for _ in range(2):
        r.reset()
        r.render(w.camera, w.params, n_samples=2)
        gpu_ctx.sync()
        gpu_ctx.replay()
        got = read_all(r)
        assert s.state() == ref_state
        assert r.get_total_samples() == 4
        for k in (N.RC_BUF_GBUFFER, N.RC_BUF_GALBEDO, N.RC_BUF_PRIMARY_ID):
            assert np.array_equal(got[k], ref[k]), k
        a, b = got[N.RC_BUF_ACCUM][:, :3].astype(np.float64), ref[N.RC_BUF_ACCUM][:, :3].astype(np.float64)
        assert float(np.mean((a - b) ** 2 / (b * b + 0.01))) <= 1e-6
    gpu_ctx.release_frame()


def test_replayed_views_chain(gpu_ctx):
    c = workloads.cornell_small()
    t = workloads.training_scene(3, size=64)
    reqs = [ViewRequest(0, CameraSpec(**{**c.camera.__dict__, "width": 96, "height": 54}), c.params, 6),
            ViewRequest(1, t.camera, t.params, 5)]
    vb = ViewBatch(gpu_ctx)
    vb.upload_scene(0, c.scene)
    vb.upload_scene(1, t.scene)
    Renderer(gpu_ctx).set_pt_mode(N.RC_PT_LANES)
    pp = make_params(enable=VHS_PASSES, atrous=True, iterations=3)
    tmp = TemporalInputs(0.5, 1.0, 3)

    def frame():
        vb.render(reqs, reset=True)
        for i in range(len(reqs)):
            vb.post(i, pp, tmp)

    frame()
    ref = [{b: vb.read(i, b) for b in (N.RC_BUF_ACCUM, N.RC_BUF_POST_HDR, N.RC_BUF_DISPLAY)}
           for i in range(len(reqs))]
    gpu_ctx.capture_begin()
    frame()
    gpu_ctx.capture_end()
    for _ in range(2):
        vb.render([ViewRequest(q.scene, q.camera, q.params, 2) for q in reqs], reset=True)  # same sizes: same arenas
        gpu_ctx.sync()
        with pytest.raises(RenderError) as e:
            gpu_ctx.replay()  # the tables of other views are on the device now
        assert e.value.code == N.RC_ERR_STATE
        frame()  # restore the recorded views' tables
        gpu_ctx.replay()
        for i in range(len(reqs)):
            assert vb.total_samples(i) == reqs[i].spp
            for b, v in ref[i].items():
                assert np.array_equal(vb.read(i, b), v), (i, b)
    gpu_ctx.release_frame()


def test_capture_refuses_what_it_cannot_record(gpu_ctx, cornell):
    w = cornell
    cam = CameraSpec(**{**w.camera.__dict__, "width": 64, "height": 40})
    SceneUploader(gpu_ctx).upload(w.scene)
    r = Renderer(gpu_ctx)
    r.set_pt_mode(N.RC_PT_LANES)
    r.reset()
    r.render(cam, w.params, n_samples=2)
    before = r.read(N.RC_BUF_ACCUM)
    gpu_ctx.capture_begin()
    with pytest.raises(RenderError) as e:
        r.render(cam, w.params, n_samples=1)  # no reset recorded yet
    assert e.value.code == N.RC_ERR_STATE
    for call in (lambda: r.read(N.RC_BUF_ACCUM), gpu_ctx.sync, lambda: SceneUploader(gpu_ctx).upload(w.scene),
                 lambda: r.set_accum_mode(N.RC_ACCUM_FP64), gpu_ctx.replay, gpu_ctx.capture_begin):
        with pytest.raises(RenderError) as e:
            call()
        assert e.value.code == N.RC_ERR_STATE
    r.reset()
    r.set_pt_mode(N.RC_PT_WAVEFRONT)
    with pytest.raises(RenderError) as e:
        r.render(cam, w.params, n_samples=1)
    assert e.value.code == N.RC_ERR_STATE
    r.set_pt_mode(N.RC_PT_LANES)
    with pytest.raises(RenderError):
        r.render(CameraSpec(**{**cam.__dict__, "width": 70}), w.params, n_samples=1)  # would reallocate
    gpu_ctx.capture_end()
    gpu_ctx.release_frame()
    with pytest.raises(RenderError) as e:
        gpu_ctx.replay()
    assert e.value.code == N.RC_ERR_STATE
    with pytest.raises(RenderError):
        gpu_ctx.capture_end()
    # the context still works, and the refused calls changed no image
    assert np.array_equal(r.read(N.RC_BUF_ACCUM), before)
    r.reset()
    r.render(cam, w.params, n_samples=2)
    assert np.array_equal(r.read(N.RC_BUF_ACCUM), before)


def test_replay_refused_after_reallocation(gpu_ctx, cornell):
    # An eager render at another resolution reallocates the images the recorded
    # frame writes: the replay must be refused rather than touch freed memory.
    w = cornell
    cam = CameraSpec(**{**w.camera.__dict__, "width": 64, "height": 40})
    SceneUploader(gpu_ctx).upload(w.scene)
    r = Renderer(gpu_ctx)
    r.set_pt_mode(N.RC_PT_LANES)
    r.reset()
    r.render(cam, w.params, n_samples=2)
    ref = r.read(N.RC_BUF_ACCUM)
    gpu_ctx.capture_begin()
    r.reset()
    r.render(cam, w.params, n_samples=2)
    gpu_ctx.capture_end()
    gpu_ctx.replay()
    assert np.array_equal(r.read(N.RC_BUF_ACCUM), ref)
    r.reset()
    r.render(CameraSpec(**{**cam.__dict__, "width": 48}), w.params, n_samples=1)
    with pytest.raises(RenderError) as e:
        gpu_ctx.replay()
    assert e.value.code == N.RC_ERR_STATE
    gpu_ctx.release_frame()
