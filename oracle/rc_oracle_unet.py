"""FP64 CPU oracle of the learned denoiser (project-new: the reference has no
network; its inputs are the reference's render outputs, pathtrace.comp:6-7 and
:644-647 — accumulated radiance, first-hit albedo, shading normal and hit
distance).

TEST INFRASTRUCTURE ONLY: used by tests/ to check the HIP forward and backward
kernels of libretrocam_hip.so. Nothing in the product imports it.

Network (5-level U-Net, 3x3 convolutions, zero 'same' padding, ReLU):
  e1 = relu(conv(f))                      12 -> 32   level 0
  e2 = relu(conv(pool(e1)))               32 -> 64   level 1
  e3 = relu(conv(pool(e2)))               64 -> 64   level 2
  e4 = relu(conv(pool(e3)))               64 -> 96   level 3
  e5 = relu(conv(pool(e4)))               96 -> 128  level 4
  d4 = relu(conv([up(e5), e4]))          224 -> 64   level 3
  d3 = relu(conv([up(d4), e3]))          128 -> 64   level 2
  d2 = relu(conv([up(d3), e2]))          128 -> 32   level 1
  d1 = relu(conv([up(d2), e1]))           64 -> 32   level 0
  o  = conv1x1(d1)                         32 -> 3
  y  = o + f[..., 0:3]    (residual in the log1p domain)
  x  = m * expm1(min(y, 20)),  m = max(f[..., 3:6], 0.01)   (linear radiance;
       the loss takes x as is, the denoised image is max(x, 0))
pool = 2x2 mean, up = nearest. Parameters, flat, per layer in this order:
W[cout][ky][kx][cin] then b[cout].

Features f (12 channels) from a render: log1p(max(radiance, 0) / max(albedo, 0.01))
x3 (albedo-demodulated), albedo x3, normal x3, log1p(max(depth, 0)) / 4,
log1p(standard error of the mean luma), log2(spp) / 10. Raw inputs per pixel
(12 floats): radiance rgb, albedo rgb, normal xyz, depth, mean of luma^2 over
the samples, sample count.

Loss = mean over pixels and RGB of (x - r)^2 / (r^2 + 0.01)  (the relMSE that
the quality metric uses).

With bf16=True the forward rounds weights, features and every activation to
bfloat16 (round to nearest even) at the points where the HIP kernels store
bfloat16; the backward is the exact FP64 derivative of that forward with the
roundings treated as identity.

Parity of this oracle is pinned by central finite differences in FP64
(tests/test_oracle_unet.py).
"""
from __future__ import annotations

import numpy as np

# name, [(source, channels), ...], cout, kernel size, level
LAYERS = [
    ("e1", [("f", 12)], 32, 3, 0),
    ("e2", [("pool:e1", 32)], 64, 3, 1),
    ("e3", [("pool:e2", 64)], 64, 3, 2),
    ("e4", [("pool:e3", 64)], 96, 3, 3),
    ("e5", [("pool:e4", 96)], 128, 3, 4),
    ("d4", [("up:e5", 128), ("e4", 96)], 64, 3, 3),
    ("d3", [("up:d4", 64), ("e3", 64)], 64, 3, 2),
    ("d2", [("up:d3", 64), ("e2", 64)], 32, 3, 1),
    ("d1", [("up:d2", 32), ("e1", 32)], 32, 3, 0),
    ("out", [("d1", 32)], 3, 1, 0),
]


def layer_shapes():
    out = []
    for name, srcs, cout, k, _ in LAYERS:
        cin = sum(c for _, c in srcs)
        out.append((name, (cout, k, k, cin), (cout,)))
    return out


def param_count() -> int:
    return sum(int(np.prod(w)) + b[0] for _, w, b in layer_shapes())


def unpack(theta: np.ndarray) -> dict:
    p, off = {}, 0
    for name, ws, bs in layer_shapes():
        n = int(np.prod(ws))
        p[name] = (theta[off:off + n].reshape(ws), theta[off + n:off + n + bs[0]])
        off += n + bs[0]
    assert off == theta.size
    return p


def bf16(a: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def features(raw, bf=False) -> np.ndarray:
    """(..., 12) features of (..., 12) raw inputs. With bf=True the albedo is
    rounded to bfloat16 before it demodulates the radiance (as the kernels do)."""
    raw = np.asarray(raw, np.float64)
    rad, alb, nrm, dep, lsq, n = raw[..., 0:3], raw[..., 3:6], raw[..., 6:9], raw[..., 9], raw[..., 10], raw[..., 11]
    m = np.maximum(bf16(alb) if bf else alb, 0.01)
    n = np.maximum(n, 1.0)
    luma = rad[..., 0] * 0.2126 + rad[..., 1] * 0.7152 + rad[..., 2] * 0.0722
    se = np.sqrt(np.maximum(lsq - luma * luma, 0.0) / n)
    f = np.concatenate([np.log1p(np.maximum(rad, 0.0) / m), alb, nrm,
                        (np.log1p(np.maximum(dep, 0.0)) / 4.0)[..., None], np.log1p(se)[..., None],
                        (np.log2(n) * 0.1)[..., None]], axis=-1)
    return f


def conv(x, w, b):
    """'same' convolution, x (N,H,W,Cin), w (Cout,k,k,Cin)."""
    k = w.shape[1]
    r = k // 2
    n, h, wd, _ = x.shape
    xp = np.pad(x, ((0, 0), (r, r), (r, r), (0, 0)))
    out = np.zeros((n, h, wd, w.shape[0]))
    for dy in range(k):
        for dx in range(k):
            out += xp[:, dy:dy + h, dx:dx + wd, :] @ w[:, dy, dx, :].T
    return out + b


def conv_backward(x, w, dz):
    k = w.shape[1]
    r = k // 2
    n, h, wd, _ = x.shape
    xp = np.pad(x, ((0, 0), (r, r), (r, r), (0, 0)))
    dxp = np.zeros_like(xp)
    dw = np.zeros_like(w)
    dz2 = dz.reshape(-1, dz.shape[-1])
    for dy in range(k):
        for dx in range(k):
            xs = xp[:, dy:dy + h, dx:dx + wd, :]
            dw[:, dy, dx, :] = dz2.T @ xs.reshape(-1, xs.shape[-1])
            dxp[:, dy:dy + h, dx:dx + wd, :] += dz @ w[:, dy, dx, :]
    return dxp[:, r:r + h, r:r + wd, :], dw, dz2.sum(axis=0)


def pool(x):
    n, h, w, c = x.shape
    return x.reshape(n, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4))


def pool_backward(d):
    return np.repeat(np.repeat(d, 2, axis=1), 2, axis=2) / 4.0


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_backward(d):
    n, h, w, c = d.shape
    return d.reshape(n, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4))


def forward(theta, f, bf=False):
    """Returns (x, cache). f (N,H,W,10), H and W divisible by 16."""
    P = unpack(np.asarray(theta, np.float64))
    rnd = bf16 if bf else (lambda a: a)
    W = {k: (rnd(w), b) for k, (w, b) in P.items()}
    f = rnd(f)
    act = {"f": f}
    inputs = {}

    def src(s):
        if s.startswith("pool:"):
            return rnd(pool(act[s[5:]]))
        if s.startswith("up:"):
            return up(act[s[3:]])
        return act[s]

    for name, srcs, _, _, _ in LAYERS:
        xin = np.concatenate([src(s) for s, _ in srcs], axis=-1)
        inputs[name] = xin
        z = conv(xin, *W[name])
        if name == "out":
            act[name] = z
        else:
            act[name] = rnd(np.maximum(z, 0.0))
    y = act["out"] + f[..., 0:3]
    m = np.maximum(f[..., 3:6], 0.01)
    x = m * np.expm1(np.minimum(y, 20.0))
    return x, {"W": W, "act": act, "inputs": inputs, "y": y, "m": m}


def loss_and_grad(theta, f, target, bf=False):
    """relMSE loss and its gradient w.r.t. the flat parameter vector."""
    x, c = forward(theta, f, bf)
    r = np.asarray(target, np.float64)
    cnt = x.size
    den = r * r + 0.01
    loss = float(np.sum((x - r) ** 2 / den) / cnt)
    dx = 2.0 * (x - r) / den / cnt
    dy = dx * np.where(c["y"] < 20.0, x + c["m"], 0.0)
    W, act, inputs = c["W"], c["act"], c["inputs"]
    grads = {}
    dact = {name: None for name, *_ in LAYERS}
    dact["out"] = dy

    def add(name, g):
        dact[name] = g if dact[name] is None else dact[name] + g

    for name, srcs, _, _, _ in reversed(LAYERS):
        dz = dact[name]
        if name != "out":
            dz = dz * (act[name] > 0.0)
        dxin, dw, db = conv_backward(inputs[name], W[name][0], dz)
        grads[name] = (dw, db)
        off = 0
        for s, ch in srcs:
            g = dxin[..., off:off + ch]
            off += ch
            if s == "f":
                continue
            if s.startswith("pool:"):
                add(s[5:], pool_backward(g))
            elif s.startswith("up:"):
                add(s[3:], up_backward(g))
            else:
                add(s, g)
    flat = np.concatenate([np.concatenate([grads[n][0].ravel(), grads[n][1].ravel()]) for n, *_ in LAYERS])
    return loss, flat


def he_init(seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    parts = []
    for _, ws, bs in layer_shapes():
        fan_in = ws[1] * ws[2] * ws[3]
        parts.append(rng.standard_normal(int(np.prod(ws))) * np.sqrt(2.0 / fan_in))
        parts.append(np.zeros(bs[0]))
    return np.concatenate(parts)
