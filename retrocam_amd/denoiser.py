"""Learned denoiser over the C ABI (rc_net_* in include/retrocam_hip.h).

Project-new: the reference denoises only with the a-trous filter
(src/main/java/com/retrocam/post/AtrousDenoiser.java:60-125), which stays the
fallback. The network's per-pixel inputs are the reference's render outputs
(pathtrace.comp:6-7, 644-654): the accumulated radiance mean, first-hit albedo,
shading normal, hit distance, mean luma^2 and the sample count ("raw", 12
floats per pixel).

  NetDenoiser     parameters, training, forward, and the switch that makes
                  rc_post run the network instead of a-trous
  frame_raw       raw inputs of the frame currently in a Renderer
  render_pair     (raw at low spp, radiance target at high spp) of one view
  make_dataset    pairs over seeded training scenes (workloads.training_scene)
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _native as N
from .renderer import Renderer, SceneUploader

RAW_CHANNELS = 12
WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "unet_denoiser.npy")


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def param_count() -> int:
    return int(N.hip_lib().rc_net_param_count())


class NetDenoiser:
    def __init__(self, ctx):
        self.ctx = ctx
        self._lib = ctx._lib
        self.n = int(self._lib.rc_net_param_count())

    def init(self, seed: int = 1) -> "NetDenoiser":
        self.ctx.check(self._lib.rc_net_init(self.ctx.h, int(seed) & 0xFFFFFFFF))
        return self

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.n, np.float32)
        self.ctx.check(self._lib.rc_net_get_params(self.ctx.h, _fp(out), self.n))
        return out

    def set_params(self, theta: np.ndarray) -> "NetDenoiser":
        t = np.ascontiguousarray(theta, np.float32)
        self.ctx.check(self._lib.rc_net_set_params(self.ctx.h, _fp(t), int(t.size)))
        return self

    def save(self, path: str = WEIGHTS_PATH):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.save(path, self.get_params())

    def load(self, path: str = WEIGHTS_PATH) -> "NetDenoiser":
        return self.set_params(np.load(path, allow_pickle=False))

    def set_dataset(self, raw: np.ndarray, target: np.ndarray):
        raw = np.ascontiguousarray(raw, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        n, h, w, c = raw.shape
        assert c == RAW_CHANNELS and target.shape == (n, h, w, 3)
        self.ctx.check(self._lib.rc_net_set_dataset(self.ctx.h, _fp(raw), _fp(target), n, h, w))

    def train(self, steps: int, batch: int = 8, lr: float = 1e-3, seed: int = 1) -> np.ndarray:
        losses = np.zeros(steps, np.float32)
        self.ctx.check(self._lib.rc_net_train(self.ctx.h, int(steps), int(batch), float(lr), int(seed) & 0xFFFFFFFF,
                                              _fp(losses)))
        return losses

    def grad(self, raw: np.ndarray, target: np.ndarray):
        raw = np.ascontiguousarray(raw, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        n, h, w, _ = raw.shape
        loss = np.zeros(1, np.float32)
        g = np.zeros(self.n, np.float32)
        self.ctx.check(self._lib.rc_net_grad(self.ctx.h, _fp(raw), _fp(target), n, h, w, _fp(loss), _fp(g)))
        return float(loss[0]), g

    def forward(self, raw: np.ndarray) -> np.ndarray:
        raw = np.ascontiguousarray(raw, np.float32)
        n, h, w, _ = raw.shape
        out = np.zeros((n, h, w, 3), np.float32)
        self.ctx.check(self._lib.rc_net_forward(self.ctx.h, _fp(raw), n, h, w, _fp(out)))
        return out

    def set_flips(self, passes: int = 1):
        """rc_set_net_flips: network passes per denoise (1, 2, 4 or 8), averaged
        over the frame, its mirror images and (8) those moved 8 pixels right and
        down."""
        self.ctx.check(self._lib.rc_set_net_flips(self.ctx.h, int(passes)))

    def set_kernels(self, mode: int = N.RC_NET_KERNELS_TILED):
        """rc_set_net_kernels: which convolution kernels the forward pass uses
        (tiled, direct or direct with one channel group); outputs are identical."""
        self.ctx.check(self._lib.rc_set_net_kernels(self.ctx.h, int(mode)))

    def enable(self, on: bool = True):
        """rc_post's denoise step runs the network (on) or a-trous (off)."""
        self.ctx.check(self._lib.rc_set_denoiser(self.ctx.h, N.RC_DENOISE_NET if on else N.RC_DENOISE_ATROUS))


def _half(a: np.ndarray) -> np.ndarray:
    return a.view(np.float16).astype(np.float32)


def frame_raw(r: Renderer, width: int, height: int) -> np.ndarray:
    """(height, width, 12) raw inputs of the frame in r (bottom-up rows, as the
    reference stores images)."""
    return _raw_from(r.read(N.RC_BUF_ACCUM), r.read(N.RC_BUF_GBUFFER), r.read(N.RC_BUF_GALBEDO),
                     r.read(N.RC_BUF_VARIANCE), width, height)


def render_pair(ctx, w, spp_in: int, spp_ref: int):
    """Raw inputs at spp_in and the FP64-accumulated spp_ref radiance target of workload w."""
    SceneUploader(ctx).upload(w.scene)
    r = Renderer(ctx)
    r.set_accum_mode(N.RC_ACCUM_FP32)
    r.reset()
    r.render(w.camera, w.params, n_samples=spp_in)
    raw = frame_raw(r, w.camera.width, w.camera.height)
    r.set_accum_mode(N.RC_ACCUM_FP64)
    r.reset()
    done = 0
    while done < spp_ref:
        k = min(512, spp_ref - done)
        r.render(w.camera, w.params, n_samples=k)
        done += k
    acc = r.read(N.RC_BUF_ACCUM64)
    tgt = (acc[:, 0:3] / acc[:, 3:4]).astype(np.float32).reshape(w.camera.height, w.camera.width, 3)
    r.set_accum_mode(N.RC_ACCUM_FP32)
    return raw, tgt


def _raw_from(acc, gbuf, galb, variance, w, h):
    """Raw inputs: radiance mean, albedo, normal, depth, mean luma^2, spp."""
    gb = _half(gbuf)
    ga = _half(galb)
    raw = np.zeros((h * w, RAW_CHANNELS), np.float32)
    raw[:, 0:3] = acc[:, 0:3] / acc[:, 3:4]
    raw[:, 3:6] = ga[:, 0:3]
    raw[:, 6:9] = gb[:, 0:3]
    raw[:, 9] = gb[:, 3]
    raw[:, 10] = variance / acc[:, 3]
    raw[:, 11] = acc[:, 3]
    return raw.reshape(h, w, RAW_CHANNELS)


def _raw_between(a0, v0, a1, v1, gbuf, galb, w, h):
    """Raw inputs of the samples added between two reads of one FP64
    accumulation (sums a0 -> a1 and luma^2 sums v0 -> v1): the mean of those
    samples alone, an input independent of the earlier ones."""
    n = a1[:, 3] - a0[:, 3]
    acc = np.empty((a1.shape[0], 4), np.float64)
    acc[:, 0:3] = a1[:, 0:3] - a0[:, 0:3]
    acc[:, 3] = n
    var = v1.astype(np.float64) - v0.astype(np.float64)
    return _raw_from(acc, gbuf, galb, var, w, h)


def make_dataset(ctx, seeds, size: int = 128, spp_in=(4, 8, 16, 32, 64, 128, 256), spp_ref: int = 1024, log=None,
                 group: int = N.RC_MAX_SCENE_SLOTS, levels: int = 1, seed: int = 1, independent: bool = False):
    """Pairs over workloads.training_scene(seed); spp_in cycles through the given
    counts. With levels > 1 each view yields `levels` inputs at increasing sample
    counts drawn from spp_in that share its high-spp target: reads of one
    accumulation at each count, or, with `independent`, the means of disjoint
    runs of samples (frames n1.., n1 + n2.., ...: independent noise at each
    level). Views are rendered in batches of `group` scenes, one persistent
    launch per batch and sample chunk (rc_views_render)."""
    from . import workloads
    from .batch import ViewBatch, ViewRequest
    seeds = list(seeds)
    rng = np.random.default_rng(seed)
    vb = ViewBatch(ctx)
    raws, tgts = [], []
    for g0 in range(0, len(seeds), group):
        ws = [workloads.training_scene(int(s), size=size) for s in seeds[g0:g0 + group]]
        for k, w in enumerate(ws):
            vb.upload_scene(k, w.scene)
        if levels <= 1:
            lv = [[int(spp_in[(g0 + k) % len(spp_in)])] for k in range(len(ws))]
        else:
            lv = [sorted(int(v) for v in rng.choice(spp_in, size=levels, replace=False)) for _ in ws]
        done = [0] * len(ws)
        per_view = [[] for _ in ws]
        prev = [None] * len(ws)
        for r in range(len(lv[0])):
            if independent:
                reqs = [ViewRequest(k, w.camera, w.params, lv[k][r]) for k, w in enumerate(ws)]
                vb.render(reqs, reset=r == 0, fp64=True)
                for k in range(len(ws)):
                    a1, v1 = vb.read(k, N.RC_BUF_ACCUM64), vb.read(k, N.RC_BUF_VARIANCE)
                    a0, v0 = prev[k] if prev[k] is not None else (np.zeros_like(a1), np.zeros_like(v1))
                    per_view[k].append(_raw_between(a0, v0, a1, v1, vb.read(k, N.RC_BUF_GBUFFER),
                                                    vb.read(k, N.RC_BUF_GALBEDO), size, size))
                    prev[k] = (a1, v1)
                continue
            reqs = [ViewRequest(k, w.camera, w.params, lv[k][r] - done[k]) for k, w in enumerate(ws)]
            vb.render(reqs, reset=r == 0)
            for k in range(len(ws)):
                done[k] = lv[k][r]
                per_view[k].append(_raw_from(vb.read(k, N.RC_BUF_ACCUM), vb.read(k, N.RC_BUF_GBUFFER),
                                             vb.read(k, N.RC_BUF_GALBEDO), vb.read(k, N.RC_BUF_VARIANCE), size,
                                             size))
        done_ref = 0
        while done_ref < spp_ref:
            kk = min(512, spp_ref - done_ref)
            vb.render([ViewRequest(k, w.camera, w.params, kk) for k, w in enumerate(ws)], reset=done_ref == 0,
                      fp64=True)
            done_ref += kk
            ctx.sync()
        for k in range(len(ws)):
            acc = vb.read(k, N.RC_BUF_ACCUM64)
            t = (acc[:, 0:3] / acc[:, 3:4]).astype(np.float32).reshape(size, size, 3)
            for raw in per_view[k]:
                raws.append(raw)
                tgts.append(t)
        if log:
            log(f"[dataset] {min(g0 + group, len(seeds))}/{len(seeds)} scenes")
    return np.stack(raws), np.stack(tgts)


def make_frame_crops(ctx, seeds, size: int = 128, spp_in=(4, 8, 16, 32, 64, 128, 256), spp_ref: int = 1024,
                     levels: int = 3, log=None, seed: int = 2, independent: bool = False):
    """Pairs from full frames of workloads.full_frame(seed) (854x480, 250k
    triangles, thin lens), cut into size x size crops: `levels` inputs per frame
    at increasing sample counts from spp_in (with `independent`, means of
    disjoint runs of samples, as in make_dataset), targets at spp_ref with FP64
    accumulation."""
    from . import workloads
    from .renderer import Renderer, SceneUploader
    rng = np.random.default_rng(seed)
    raws, tgts = [], []
    r = Renderer(ctx)
    for i, fs in enumerate(seeds):
        w = workloads.full_frame(seed=int(fs))
        W, H = w.camera.width, w.camera.height
        SceneUploader(ctx).upload(w.scene)
        lv = sorted(int(v) for v in rng.choice(spp_in, size=levels, replace=False))
        full = []
        if independent:
            r.set_accum_mode(N.RC_ACCUM_FP64)
            r.reset()
            a0, v0 = None, None
            for n in lv:
                r.render(w.camera, w.params, n_samples=n)
                a1, v1 = r.read(N.RC_BUF_ACCUM64), r.read(N.RC_BUF_VARIANCE)
                if a0 is None:
                    a0, v0 = np.zeros_like(a1), np.zeros_like(v1)
                full.append(_raw_between(a0, v0, a1, v1, r.read(N.RC_BUF_GBUFFER), r.read(N.RC_BUF_GALBEDO), W, H))
                a0, v0 = a1, v1
        else:
            r.set_accum_mode(N.RC_ACCUM_FP32)
            r.reset()
            done = 0
            for n in lv:
                r.render(w.camera, w.params, n_samples=n - done)
                done = n
                full.append(_raw_from(r.read(N.RC_BUF_ACCUM), r.read(N.RC_BUF_GBUFFER), r.read(N.RC_BUF_GALBEDO),
                                      r.read(N.RC_BUF_VARIANCE), W, H))
        r.set_accum_mode(N.RC_ACCUM_FP64)
        r.reset()
        done = 0
        while done < spp_ref:
            kk = min(256, spp_ref - done)
            r.render(w.camera, w.params, n_samples=kk)
            done += kk
            ctx.sync()
        acc = r.read(N.RC_BUF_ACCUM64)
        t = (acc[:, 0:3] / acc[:, 3:4]).astype(np.float32).reshape(H, W, 3)
        r.set_accum_mode(N.RC_ACCUM_FP32)
        for y0 in range(0, H - size + 1, size):
            for x0 in range(0, W - size + 1, size):
                for raw in full:
                    raws.append(raw[y0:y0 + size, x0:x0 + size])
                    tgts.append(t[y0:y0 + size, x0:x0 + size])
        if log:
            log(f"[dataset] full frame {i + 1}/{len(seeds)} (seed {fs})")
    return np.stack(raws), np.stack(tgts)


def rel_mse(x: np.ndarray, r: np.ndarray) -> float:
    """mean over pixels and RGB of (x - r)^2 / (r^2 + 0.01)."""
    x = np.asarray(x, np.float64)
    r = np.asarray(r, np.float64)
    return float(np.mean((x - r) ** 2 / (r * r + 0.01)))
