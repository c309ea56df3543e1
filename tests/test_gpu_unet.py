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
from retrocam_amd.renderer import RenderError, Renderer, SceneUploader
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
        # the bias block on its own (column sums of the layer's output gradient)
        ab, rb = a[-bs[0]:], b[-bs[0]:]
        rel_b = float(np.linalg.norm(ab - rb) / (np.linalg.norm(rb) + 1e-30))
        assert rel_b <= 3e-2, (name, "bias", rel_b)


@pytest.mark.parametrize("shape", [(4, 480, 864), (8, 128, 128), (1, 48, 80), (3, 16, 16)])
def test_conv_kernels_are_bit_identical(gpu_ctx, shape):
    # The tiled kernels (input tiles in LDS, weight fragments prefetched, tiles
    # grouped per XCD), the direct kernels and the direct kernels with one
    # channel group per wave compute every output from the same operands in the
    # same K order: the network's outputs and gradients agree bit for bit.
    n, h, w = shape
    net = NetDenoiser(gpu_ctx).load()
    raw, tgt = random_raw(20 + n, n, h, w)
    outs, grads = [], []
    try:
        for mode in (N.RC_NET_KERNELS_TILED, N.RC_NET_KERNELS_DIRECT, N.RC_NET_KERNELS_SINGLE_GROUP):
            net.set_kernels(mode)
            outs.append(net.forward(raw))
            if n * h * w <= 8 * 128 * 128:
                grads.append(net.grad(raw, tgt)[1])
    finally:
        net.set_kernels(N.RC_NET_KERNELS_TILED)
    for o in outs[1:]:
        assert np.array_equal(o.view(np.uint32), outs[0].view(np.uint32))
    for g in grads[1:]:
        assert np.array_equal(g.view(np.uint32), grads[0].view(np.uint32))
    assert np.isfinite(outs[0]).all() and outs[0].max() > 0


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


@pytest.mark.parametrize("passes", [2, 4, 8])
def test_flip_ensemble_is_the_mean_of_mirrored_passes(gpu_ctx, passes):
    # rc_set_net_flips: pass k runs on the frame mirrored by k (bit 0 left-right,
    # bit 1 top-bottom) and, for k >= 4, moved 8 pixels right and down (edge
    # replication fills the top rows and left columns); its output is mapped
    # back and the passes averaged as (((p0 + p1) + p2) + ...) * (1 / passes).
    # The same composition of rc_net_forward calls on the host must give the
    # post chain's image bit for bit.
    w = workloads.training_scene(4, size=64)
    W, H = 90, 52  # not multiples of 16: the frame is padded by edge replication
    cam = CameraSpec(**{**w.camera.__dict__, "width": W, "height": H})
    SceneUploader(gpu_ctx).upload(w.scene)
    r = Renderer(gpu_ctx)
    r.reset()
    r.render(cam, w.params, n_samples=6)
    net = NetDenoiser(gpu_ctx).init(9)
    net.enable(True)
    net.set_flips(passes)
    try:
        out = PostProcessStack(gpu_ctx).run_on_accum(make_params(atrous=True, iterations=5), TemporalInputs())
    finally:
        net.set_flips(1)
        net.enable(False)
    raw = frame_raw(r, W, H)
    room = 8 if passes > 4 else 0  # the padded frame leaves room for the move
    Wp, Hp = (W + room + 15) // 16 * 16, (H + room + 15) // 16 * 16

    def fwd(x, s):
        xp = np.pad(x, ((s, Hp - H - s), (s, Wp - W - s), (0, 0)), mode="edge")
        return np.maximum(net.forward(xp[None])[0][s:s + H, s:s + W], np.float32(0))

    flips = [lambda a: a, lambda a: a[:, ::-1], lambda a: a[::-1], lambda a: a[::-1, ::-1]]
    acc = None
    for k in range(passes):
        f, s = flips[k & 3], 8 if k & 4 else 0
        y = f(fwd(f(raw), s))
        acc = y if acc is None else (acc + y).astype(np.float32)
    ref = (acc * np.float32(1.0 / passes)).astype(np.float32).reshape(-1, 3)
    assert np.array_equal(out.hdr[:, :3], ref)
    assert np.all(out.hdr[:, 3] == 1.0)
    with pytest.raises(RenderError):
        net.set_flips(3)
