"""What the trained network gains over the method without it, on frames it never
saw: the committed U-Net weights against the reference's a-trous denoiser
(AtrousDenoiser.java:60-125, 5 passes) on held-out full plain frames.

The frames are the benchmark generator (workloads.full_frame: 250k triangles,
854x480, thin lens) at seeds 901 and 902, which neither the training sets
(seeds 2-65, 300-379, 400-479) nor the selection of the weights (seed 1) used.
Inputs have 5 samples per pixel, the benchmark's operating point; the
reference is a 2048-sample FP64 accumulation of the same frame. relMSE =
mean over pixels and RGB of (x - r)^2 / (r^2 + 0.01) on linear HDR.

Stated margin: the network's relMSE is at most a quarter of a-trous' and
at most 0.03 (its error against a 4096-sample reference is about 0.01; the
2048-sample reference adds its own noise, about 0.005)."""
import numpy as np
import pytest

from retrocam_amd import _native as N
from retrocam_amd import workloads
from retrocam_amd.denoiser import NetDenoiser, rel_mse
from retrocam_amd.post import PostProcessStack, TemporalInputs, make_params
from retrocam_amd.renderer import Renderer, SceneUploader

pytestmark = pytest.mark.gpu

SPP_IN = 5
SPP_REF = 2048


@pytest.mark.parametrize("seed", [901, 902])
def test_net_beats_atrous_on_held_out_frames(gpu_ctx, seed):
    w = workloads.full_frame(seed=seed, spp=SPP_IN)
    SceneUploader(gpu_ctx).upload(w.scene)
    r = Renderer(gpu_ctx)
    r.set_pt_mode(N.RC_PT_LANES)
    r.set_accum_mode(N.RC_ACCUM_FP64)
    r.reset()
    done = 0
    while done < SPP_REF:
        r.render(w.camera, w.params, n_samples=256)
        done += 256
    acc64 = r.read(N.RC_BUF_ACCUM64)
    ref = acc64[:, :3] / acc64[:, 3:4]
    r.set_accum_mode(N.RC_ACCUM_FP32)
    r.reset()
    r.render(w.camera, w.params, n_samples=SPP_IN)
    net = NetDenoiser(gpu_ctx).load()
    post = PostProcessStack(gpu_ctx)
    p = make_params(atrous=True, iterations=5)
    t = TemporalInputs()
    err = {}
    for name, on, passes in (("atrous", False, 1), ("net1", True, 1), ("net4", True, 4)):
        net.enable(on)
        net.set_flips(passes)
        post.run_on_accum(p, t, read=False)
        err[name] = rel_mse(r.read(N.RC_BUF_POST_HDR)[:, :3], ref)
    net.enable(False)
    net.set_flips(1)
    print(f"held-out frame seed {seed}, {SPP_IN} spp vs {SPP_REF} spp: relMSE {err}", flush=True)
    assert err["net4"] <= 0.25 * err["atrous"] and err["net4"] <= 0.03, err
    assert err["net1"] <= 0.25 * err["atrous"], err
