"""Batched views (rc_views_* through the C ABI): every view of a persistent
multi-view launch must equal, bit for bit, a single-view render of the same
request (rc_render_passes), including continued accumulation, FP64 mode and
the post chain."""
import numpy as np
import pytest

from retrocam_amd import _native as N
from retrocam_amd import workloads
from retrocam_amd.batch import ViewBatch, ViewRequest, plan_requests, request_units, serve
from retrocam_amd.post import VHS_PASSES, PostProcessStack, TemporalInputs, make_params
from retrocam_amd.renderer import RenderContext, RenderError, Renderer, SceneUploader
from retrocam_amd.scene import CameraSpec

pytestmark = pytest.mark.gpu

BUFS = [N.RC_BUF_ACCUM, N.RC_BUF_GBUFFER, N.RC_BUF_GALBEDO, N.RC_BUF_VARIANCE, N.RC_BUF_PRIMARY_ID,
        N.RC_BUF_GBUFFER_ID]


def requests():
    c = workloads.cornell_small()
    t = workloads.training_scene(3, size=64)
    reqs = [
        ViewRequest(0, CameraSpec(**{**c.camera.__dict__, "width": 96, "height": 54}), c.params, 6),
        ViewRequest(1, t.camera, t.params, 5),
        ViewRequest(0, CameraSpec(**{**c.camera.__dict__, "width": 72, "height": 40, "pinhole": False,
                                     "fStop": 1.4, "focalDist": 9.0, "lcaDelta": (-0.2, 0.0, 0.3)}), c.params, 3),
    ]
    return [c.scene, t.scene], reqs


def single(scene, req, spp_list, fp64=False):
    # reference: single-view megakernel renders in a context of their own
    with RenderContext(0) as ctx:
        SceneUploader(ctx).upload(scene)
        r = Renderer(ctx)
        r.set_pt_mode(N.RC_PT_MEGAKERNEL)
        if fp64:
            r.set_accum_mode(N.RC_ACCUM_FP64)
        r.reset()
        for k in spp_list:
            r.render(req.camera, req.params, n_samples=k)
        out = {b: r.read(b) for b in BUFS}
        if fp64:
            out[N.RC_BUF_ACCUM64] = r.read(N.RC_BUF_ACCUM64)
        post = PostProcessStack(ctx).run_on_accum(make_params(enable=VHS_PASSES, atrous=True, iterations=3),
                                                  TemporalInputs(0.5, 1.0, 3))
        out["hdr"], out["rgba8"] = post.hdr, post.rgba8
    return out


VIEW_MODES = [pytest.param(N.RC_PT_LANES, id="lanes"), pytest.param(N.RC_PT_MEGAKERNEL, id="megakernel")]


@pytest.mark.parametrize("mode", VIEW_MODES)
def test_views_equal_single_renders(mode, gpu_ctx):
    scenes, reqs = requests()
    Renderer(gpu_ctx).set_pt_mode(mode)
    vb = ViewBatch(gpu_ctx)
    for k, sc in enumerate(scenes):
        vb.upload_scene(k, sc)
    vb.render(reqs, reset=True)
    vb.render([ViewRequest(q.scene, q.camera, q.params, 2) for q in reqs], reset=False)
    pp = make_params(enable=VHS_PASSES, atrous=True, iterations=3)
    for i, q in enumerate(reqs):
        assert vb.total_samples(i) == q.spp + 2
        ref = single(scenes[q.scene], q, [q.spp, 2])
        for b in BUFS:
            assert np.array_equal(vb.read(i, b), ref[b]), (i, b)
        vb.post(i, pp, TemporalInputs(0.5, 1.0, 3))
        assert np.array_equal(vb.read(i, N.RC_BUF_POST_HDR), ref["hdr"]), i
        assert np.array_equal(vb.read(i, N.RC_BUF_DISPLAY), ref["rgba8"]), i


@pytest.mark.parametrize("mode", VIEW_MODES)
def test_views_fp64(mode, gpu_ctx):
    scenes, reqs = requests()
    Renderer(gpu_ctx).set_pt_mode(mode)
    vb = ViewBatch(gpu_ctx)
    for k, sc in enumerate(scenes):
        vb.upload_scene(k, sc)
    vb.render(reqs, reset=True, fp64=True)
    for i, q in enumerate(reqs):
        ref = single(scenes[q.scene], q, [q.spp], fp64=True)
        assert np.array_equal(vb.read(i, N.RC_BUF_ACCUM64), ref[N.RC_BUF_ACCUM64]), i
        assert np.array_equal(vb.read(i, N.RC_BUF_ACCUM), ref[N.RC_BUF_ACCUM]), i


def test_views_reject_bad_requests(gpu_ctx):
    scenes, reqs = requests()
    Renderer(gpu_ctx).set_pt_mode(N.RC_PT_LANES)
    vb = ViewBatch(gpu_ctx)
    vb.upload_scene(0, scenes[0])
    with pytest.raises(RenderError):
        vb.render([ViewRequest(N.RC_MAX_SCENE_SLOTS, reqs[0].camera, reqs[0].params, 1)])
    with pytest.raises(RenderError):
        vb.render([ViewRequest(N.RC_MAX_SCENE_SLOTS - 1, reqs[0].camera, reqs[0].params, 1)])  # empty slot
    with pytest.raises(RenderError):
        vb.render([ViewRequest(0, reqs[0].camera, reqs[0].params, 0)])


def test_scheduled_serving_equals_single_renders(gpu_ctx):
    # serve() packs the small requests into shared launches and runs the rest
    # alone, shortest first: every request's display must still equal its own
    # single-view render.
    scenes, reqs = requests()
    reqs = reqs + [ViewRequest(1, reqs[1].camera, reqs[1].params, 9)]
    Renderer(gpu_ctx).set_pt_mode(N.RC_PT_LANES)
    vb = ViewBatch(gpu_ctx)
    for k, sc in enumerate(scenes):
        vb.upload_scene(k, sc)
    pp = make_params(enable=VHS_PASSES, atrous=True, iterations=3)
    fill = request_units(reqs[0]) + request_units(reqs[2])  # forces several launches
    groups = plan_requests(reqs, fill_units=fill)
    assert 1 < len(groups) < len(reqs)
    got = {}

    def keep(i, v):
        got[i] = (vb.read(v, N.RC_BUF_ACCUM), vb.read(v, N.RC_BUF_DISPLAY))

    order = serve(vb, reqs, pp, TemporalInputs(0.5, 1.0, 3), on_done=keep, fill_units=fill)
    assert order == [i for g in groups for i in g]
    for i, q in enumerate(reqs):
        ref = single(scenes[q.scene], q, [q.spp])
        assert np.array_equal(got[i][0], ref[N.RC_BUF_ACCUM]), i
        assert np.array_equal(got[i][1], ref["rgba8"]), i
