"""GPU parity of the post chain (p00, a-trous, p01..p16, display) through the C
ABI against the CPU oracle: outputs within 1e-5 absolute (the stated tolerance
for a-trous and VHS), RGBA8 display values equal up to one code on >= 99.9%."""
import numpy as np
import pytest

from oracle import rco
from retrocam_amd import _native as N
from retrocam_amd.post import VHS_PASSES, PostProcessStack, TemporalInputs, make_params
from retrocam_amd.renderer import Renderer, SceneUploader

pytestmark = pytest.mark.gpu

W, H = 96, 64


def make_image(seed=4):
    rng = np.random.default_rng(seed)
    img = rng.uniform(0.0, 1.0, (H, W, 3)).astype(np.float32)
    img[H // 3:H // 2, :, :] = rng.uniform(0.8, 1.5, 3).astype(np.float32)  # bright band: smear, edges
    img[:, W // 2:W // 2 + 3, 1] = 0.02                                    # vertical edge
    a = np.ones((H * W, 4), np.float32)
    a[:, :3] = img.reshape(-1, 3)
    return a


def check(got, exp, tol=1e-5):
    d = np.abs(got.astype(np.float64) - exp.astype(np.float64))
    assert np.nanmax(d) <= tol, f"max abs diff {np.nanmax(d)}"


def check_rgba8(got, exp):
    d = np.abs(got.astype(np.int32) - exp.astype(np.int32))
    assert d.max() <= 1 and np.mean(d == 0) >= 0.999


@pytest.mark.parametrize("k", VHS_PASSES)
def test_each_vhs_pass_matches_oracle(gpu_ctx, k):
    img = make_image()
    ps = PostProcessStack(gpu_ctx)
    over = {"trackingSeverity": 0.8, "dropoutBurstRate": 1.0, "tapeAge": 0.9, "ccdSmearLength": 40}
    p = make_params(enable=[k], **over)
    for f in (0, 31, 902):
        t = TemporalInputs(time=f / 29.97, agcGain=2.0, frameIndex=f)
        g = ps.run_on_image(img, W, H, p, t)
        hdr, rgba8 = rco.post_chain(img, W, H, 1, p, t.to_c())
        check(g.hdr, hdr)
        check_rgba8(g.rgba8, rgba8)


def test_all_passes_on_image(gpu_ctx):
    img = make_image(9)
    p = make_params(enable=VHS_PASSES, atrous=True, iterations=5)  # no G-buffer: a-trous skipped
    t = TemporalInputs(time=3.2, agcGain=1.3, frameIndex=97)
    g = PostProcessStack(gpu_ctx).run_on_image(img, W, H, p, t)
    hdr, rgba8 = rco.post_chain(img, W, H, 1, p, t.to_c())
    check(g.hdr, hdr)
    check_rgba8(g.rgba8, rgba8)


def test_full_chain_on_rendered_frame(gpu_ctx, cornell):
    w = cornell
    SceneUploader(gpu_ctx).upload(w.scene)
    r = Renderer(gpu_ctx)
    r.reset()
    r.render(w.camera, w.params, n_samples=w.spp)
    acc = r.read(N.RC_BUF_ACCUM)
    gb = r.read(N.RC_BUF_GBUFFER)
    ga = r.read(N.RC_BUF_GALBEDO)
    var = r.read(N.RC_BUF_VARIANCE)
    ps = PostProcessStack(gpu_ctx)
    for iters, passes in ((5, []), (5, VHS_PASSES), (3, [1, 2, 9, 15])):
        p = make_params(enable=passes, atrous=True, iterations=iters, exposure=1.25)
        t = TemporalInputs(time=0.5, agcGain=1.1, frameIndex=12)
        g = ps.run_on_accum(p, t)
        hdr, rgba8 = rco.post_chain(acc, w.camera.width, w.camera.height, w.spp, p, t.to_c(), gb, ga, var)
        check(g.hdr, hdr)
        check_rgba8(g.rgba8, rgba8)
    # atrousMaxSpp below the sample count bypasses the denoiser (PostProcessStack.java:142-147)
    p = make_params(atrous=True, atrousMaxSpp=4)
    g = ps.run_on_accum(p, TemporalInputs())
    hdr, _ = rco.post_chain(acc, w.camera.width, w.camera.height, w.spp, p, TemporalInputs().to_c(), None, None,
                            None)
    check(g.hdr, hdr)


def test_export_frame_matches_display(gpu_ctx, cornell, tmp_path):
    # FrameExporter.exportFrame: the display image, flipped to top-down rows, as a PNG.
    from retrocam_amd.post import VHS_PASSES, PostProcessStack, TemporalInputs, make_params
    from retrocam_amd.renderer import Renderer, SceneUploader
    from retrocam_amd.scene import CameraSpec
    from test_export_cpu import decode_png
    w = cornell
    cam = CameraSpec(**{**w.camera.__dict__, "width": 96, "height": 54})
    SceneUploader(gpu_ctx).upload(w.scene)
    r = Renderer(gpu_ctx)
    r.reset()
    r.render(cam, w.params, n_samples=4)
    post = PostProcessStack(gpu_ctx)
    out = post.run_on_accum(make_params(enable=VHS_PASSES, atrous=True, iterations=3), TemporalInputs(0.5, 1.0, 2))
    path = str(tmp_path / "frame.png")
    top_down = post.export_frame(path)
    disp = out.rgba8.reshape(54, 96, 4)
    assert np.array_equal(top_down, disp[::-1])
    assert np.array_equal(decode_png(open(path, "rb").read()), disp[::-1])
