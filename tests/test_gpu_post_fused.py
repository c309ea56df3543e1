"""The post chain as fused sweeps (RC_POST_FUSED, the default) against the same
chain launched one pass at a time (RC_POST_PER_PASS, PostProcessStack's FBO
chain, PostProcessStack.java:140-280): bit-identical HDR outputs and display
encodes on rendered frames with a-trous or the network, on images of awkward
sizes, and at settings that stretch every pass's reach (long smear, strong
tracking and dropouts, large timebase and head-switch jitter, interlace
offsets); the fused chain takes far fewer launches."""
import ctypes as C

import numpy as np
import pytest

from retrocam_amd import _native as N
from retrocam_amd import workloads
from retrocam_amd.denoiser import NetDenoiser
from retrocam_amd.post import VHS_PASSES, PostProcessStack, TemporalInputs, make_params
from retrocam_amd.renderer import Renderer, SceneUploader

pytestmark = pytest.mark.gpu

STRETCHED = dict(ccdSmearLength=150, ccdSmearThreshold=0.2, trackingSeverity=1.0, trackingMaxDisplacePx=40.0,
                 dropoutBurstRate=1.0, tapeAge=1.6, timebaseAmplitudePx=12.0, headSwitchJitter=15.0,
                 headSwitchLines=60, interlaceFieldOffsetPx=5.0, interlaceCombStrength=4.0,
                 interlaceCombEdgeThresh=0.005, chromaResSigma=30.0, lumaBlurSigma=4.0)


def set_mode(ctx, mode):
    ctx.check(N.hip_lib().rc_set_post_mode(ctx.h, mode))


def launches(ctx):
    n = C.c_int32()
    ctx.check(N.hip_lib().rc_post_launches(ctx.h, C.byref(n)))
    return n.value


def both_modes(ctx, run):
    out = {}
    for mode in (N.RC_POST_PER_PASS, N.RC_POST_FUSED):
        set_mode(ctx, mode)
        r = run()
        out[mode] = (r.hdr.copy(), r.rgba8.copy(), launches(ctx))
    set_mode(ctx, N.RC_POST_FUSED)
    (h0, d0, n0), (h1, d1, n1) = out[N.RC_POST_PER_PASS], out[N.RC_POST_FUSED]
    assert np.array_equal(h1.view(np.uint32), h0.view(np.uint32)), "HDR differs"
    assert np.array_equal(d1, d0), "display differs"
    return n0, n1


@pytest.fixture(scope="module")
def cornell_frame(gpu_ctx):
    w = workloads.cornell_small()
    SceneUploader(gpu_ctx).upload(w.scene)
    r = Renderer(gpu_ctx)
    r.reset()
    r.render(w.camera, w.params, n_samples=w.spp)
    return w


@pytest.mark.parametrize("case", ["default", "stretched", "no_vhs", "subset"])
def test_fused_chain_equals_per_pass_on_a_rendered_frame(gpu_ctx, cornell_frame, case):
    ps = PostProcessStack(gpu_ctx)
    if case == "default":
        p = make_params(enable=VHS_PASSES, atrous=True, iterations=5)
    elif case == "stretched":
        p = make_params(enable=VHS_PASSES, atrous=True, iterations=3, exposure=1.7, **STRETCHED)
    elif case == "no_vhs":
        p = make_params(atrous=True, iterations=5)
    else:
        p = make_params(enable=[2, 8, 12, 16], atrous=False)
    for f in (0, 31, 902):
        t = TemporalInputs(time=f / 29.97, agcGain=1.4, frameIndex=f)
        n0, n1 = both_modes(gpu_ctx, lambda: ps.run_on_accum(p, t))
        print(f"{case} frame {f}: per-pass {n0} launches, fused {n1}", flush=True)
        assert n1 < n0
        if case == "default":
            # per pass: p00 + 5 a-trous + 14 VHS + display = 21; fused: p00 + 5
            # a-trous + at most 5 sweeps for the VHS passes and the display
            assert n0 == 21 and n1 <= 11


@pytest.mark.parametrize("size", [(1, 1), (5, 3), (37, 29), (300, 7), (7, 300), (1000, 64), (854, 480)])
def test_fused_chain_equals_per_pass_on_images(gpu_ctx, size):
    w, h = size
    rng = np.random.default_rng(w * 1000 + h)
    img = np.ones((w * h, 4), np.float32)
    img[:, :3] = rng.gamma(1.0, 0.7, (w * h, 3))
    ps = PostProcessStack(gpu_ctx)
    for kw in ({}, STRETCHED):
        p = make_params(enable=VHS_PASSES, **kw)
        t = TemporalInputs(time=2.5, agcGain=1.9, frameIndex=333)
        both_modes(gpu_ctx, lambda: ps.run_on_image(img, w, h, p, t))


def test_fused_chain_equals_per_pass_with_the_network(gpu_ctx, cornell_frame):
    net = NetDenoiser(gpu_ctx).load()
    net.enable(True)
    try:
        ps = PostProcessStack(gpu_ctx)
        for flips in (1, 4):
            net.set_flips(flips)
            p = make_params(enable=VHS_PASSES, atrous=True, iterations=5)
            both_modes(gpu_ctx, lambda: ps.run_on_accum(p, TemporalInputs(1.0, 1.0, 7)))
    finally:
        net.set_flips(1)
        net.enable(False)
