"""GPU checks of the learned denoiser (rc_net_* through the C ABI) against the
FP64 oracle (oracle/rc_oracle_unet.py) with bfloat16 rounding emulated at the
points where the kernels store bfloat16.

Tolerances (bfloat16 activations, FP32 accumulation in a different order):
forward relMSE between GPU and oracle outputs <= 1e-4; loss within 1e-3
relative; per-layer gradient cosine >= 0.999 and relative L2 error <= 3e-2."""
import numpy as np
import pytest

from oracle import rc_oracle_unet as U
from retrocam_amd import _native as N
from retrocam_amd import workloads
from retrocam_amd.denoiser import NetDenoiser, frame_raw, make_dataset, rel_mse
from retrocam_amd.post import PostProcessStack, TemporalInputs, make_params
from retrocam_amd.renderer import Renderer, SceneUploader
from retrocam_amd.scene import CameraSpec

pytestmark = pytest.mark.gpu


def random_raw(seed, n, h, w):
    rng = np.random.default_rng(seed)
    raw = np.zeros((n, h, w, 12), np.float32)
    raw[..., 0:3] = rng.gamma(1.0, 0.5, (n, h, w, 3))
    raw[..., 3:6] = rng.uniform(0.05, 0.9, (n, h, w, 3))
    nrm = rng.standard_normal((n, h, w, 3))
    raw[..., 6:9] = nrm / np.linalg.norm(nrm, axis=-1, keepdims=True)
    raw[..., 9] = rng.uniform(0.5, 12.0, (n, h, w))
    luma = raw[..., 0] * 0.2126 + raw[..., 1] * 0.7152 + raw[..., 2] * 0.0722
    raw[..., 10] = luma * luma + rng.uniform(0.0, 2.0, (n, h, w))
    raw[..., 11] = rng.choice([4.0, 8.0, 64.0], (n, h, w))
    tgt = (raw[..., 0:3] * rng.uniform(0.7, 1.3, (n, h, w, 3))).astype(np.float32)
    return raw, tgt


def oracle_features(raw):
    return U.features(raw.astype(np.float64), bf=True)


def test_forward_matches_oracle(gpu_ctx):
    net = NetDenoiser(gpu_ctx).init(7)
    theta = net.get_params()
    raw, _ = random_raw(1, 2, 32, 48)
    out = net.forward(raw)
    ref, _ = U.forward(theta.astype(np.float64), oracle_features(raw), bf=True)
    ref = np.maximum(ref, 0.0)  # the denoised image is clamped at 0
    assert out.shape == ref.shape
    assert rel_mse(out, ref) <= 1e-4


def test_gradient_matches_oracle(gpu_ctx):
    net = NetDenoiser(gpu_ctx).init(11)
    theta = net.get_params()
    raw, tgt = random_raw(2, 2, 32, 32)
    loss, g = net.grad(raw, tgt)
    rloss, rg = U.loss_and_grad(theta.astype(np.float64), oracle_features(raw), tgt.astype(np.float64), bf=True)
    assert abs(loss - rloss) <= 1e-3 * abs(rloss)
    off = 0
    for name, ws, bs in U.layer_shapes():
        n = int(np.prod(ws)) + bs[0]
        a, b = g[off:off + n].astype(np.float64), rg[off:off + n]
        off += n
        cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30))
        rel = float(np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30))
        assert cos >= 0.999 and rel <= 3e-2, (name, cos, rel)


def test_training_reduces_loss(gpu_ctx):
    raw, tgt = make_dataset(gpu_ctx, range(6), size=64, spp_in=(8,), spp_ref=256)
    net = NetDenoiser(gpu_ctx).init(3)
    net.set_dataset(raw, tgt)
    losses = net.train(150, batch=4, lr=1e-3, seed=5)
    assert np.all(np.isfinite(losses))
    assert float(np.mean(losses[-20:])) < 0.5 * float(np.mean(losses[:5]))


def test_post_chain_runs_the_network(gpu_ctx):
    w = workloads.training_scene(3, size=64)
    cam = CameraSpec(**{**w.camera.__dict__, "width": 96, "height": 64})
    SceneUploader(gpu_ctx).upload(w.scene)
    r = Renderer(gpu_ctx)
    r.reset()
    r.render(cam, w.params, n_samples=8)
    net = NetDenoiser(gpu_ctx).init(5)
    net.enable(True)
    try:
        out = PostProcessStack(gpu_ctx).run_on_accum(make_params(atrous=True, iterations=5), TemporalInputs())
    finally:
        net.enable(False)
    raw = frame_raw(r, 96, 64)
    ref = net.forward(raw[None])[0].reshape(-1, 3)
    assert rel_mse(out.hdr[:, :3], ref) <= 1e-6
