"""CPU checks that pin the post-chain oracle (rc_oracle_post.c): hash known
answers, closed forms (p00, display ACES/sRGB, the linear BT.601 round trips,
p09's geometric series), reductions to the textbook B3 a-trous, border modes
(colour clamps, G-buffer wraps) and identity settings of every VHS pass."""
import json
import os

import numpy as np
import pytest

from oracle import rco
from retrocam_amd.post import TemporalInputs, make_params

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
W, H = 40, 24


def rgba(img3):
    a = np.ones((img3.shape[0] * img3.shape[1], 4), np.float32)
    a[:, :3] = img3.reshape(-1, 3)
    return a


def rand_img(seed, lo=0.0, hi=1.0):
    return np.random.default_rng(seed).uniform(lo, hi, (H, W, 3)).astype(np.float32)


def t0(**kw):
    return TemporalInputs(**kw).to_c()


def f16(a):
    return np.asarray(a, np.float32).astype(np.float16).view(np.uint16)


def test_hash_known_answers():
    d = json.load(open(os.path.join(GOLDEN, "hash_vectors.json")))
    assert all(rco.hash_u32(0, a) == h for a, h in d["pcg"])
    assert all(rco.hash_u32(1, a) == h for a, h in d["uhash1"])
    assert all(rco.hash_u32(2, a) == h for a, h in d["uhash2"])
    assert all(rco.hash_u32(3, a, b) == h for a, b, h in d["fhash1"])
    assert all(rco.hash_u32(4, a, b) == h for a, b, h in d["fhash2"])


def test_p00_normalize_closed_form():
    acc = rgba(rand_img(1, 0, 50))
    out = rco.post_pass(0, acc, W, H, make_params(exposure=1.7), t0(), total=37)
    exp = (acc[:, :3] / np.float32(37)) * np.float32(1.7)
    assert np.array_equal(out[:, :3], exp.astype(np.float32))
    assert np.all(out[:, 3] == 1.0)


def test_display_encode_closed_form():
    x = np.linspace(0, 8, W * H).astype(np.float64)
    img = np.repeat(x[:, None], 3, axis=1).reshape(H, W, 3).astype(np.float32)
    out = rco.post_pass(200, rgba(img), W, H, make_params(), t0())[:, 0].astype(np.float64)
    a = np.clip((x * (2.51 * x + 0.03)) / (x * (2.43 * x + 0.59) + 0.14), 0, 1)
    s = np.where(a < 0.0031308, 12.92 * a, 1.055 * a ** (1 / 2.4) - 0.055)
    assert np.max(np.abs(out - s)) < 2e-6


def _m601(coefs):
    return np.array(coefs, np.float64)


M5 = _m601([[0.299, 0.587, 0.114], [-0.16874, -0.33126, 0.5], [0.5, -0.41869, -0.08131]])
I5 = _m601([[1, 0, 1.402], [1, -0.34414, -0.71414], [1, 1.772, 0]])
M3 = _m601([[0.299, 0.587, 0.114], [-0.169, -0.331, 0.5], [0.5, -0.419, -0.081]])
I3 = _m601([[1, 0, 1.402], [1, -0.344, -0.714], [1, 1.772, 0]])


def roundtrip(img, M, I, clamp=False):
    r = img.reshape(-1, 3).astype(np.float64) @ (I @ M).T
    return np.clip(r, 0, 1) if clamp else r


@pytest.mark.parametrize("pass_id,kw,matrix,clamp", [
    (3, dict(colorSaturation=1.0, colorHueRot=0.0, colorTempBias=0.0, colorDriftSpeed=0.0), "3", True),
    (4, dict(drShadowCrush=0.0, drBlackLift=0.0, drHighlightKnee=1.0, drHighlightClip=1.0), "3", True),
    (7, dict(), "5", False),  # agcGain 0 below: zero noise amplitude
    (15, dict(), "5", False),  # constant image: no detail
    (2, dict(), "5", False),   # constant image: chroma blur is the identity
    (9, dict(), "5", False),
])
def test_identity_settings_reduce_to_bt601_round_trip(pass_id, kw, matrix, clamp):
    const = pass_id in (15, 2, 9)
    img = np.full((H, W, 3), (0.3, 0.55, 0.8), np.float32) if const else rand_img(2, 0.05, 0.95)
    out = rco.post_pass(pass_id, rgba(img), W, H, make_params(**kw), t0(agcGain=0.0))[:, :3]
    M, I = (M3, I3) if matrix == "3" else (M5, I5)
    assert np.max(np.abs(out - roundtrip(img, M, I, clamp))) < 2e-6


@pytest.mark.parametrize("pass_id,kw", [
    (1, dict()),                       # constant image: Gaussian weights normalise
    (8, dict(ccdSmearThreshold=1.5)),  # no luma above threshold: clamp only
    (10, dict(timebaseAmplitudePx=0.0)),
    (11, dict(dotCrawlIntensity=0.0)),
    (12, dict(dropoutBurstRate=0.0)),
    (13, dict(headSwitchLines=0)),
    (14, dict(trackingSeverity=0.0)),
    (16, dict(interlaceFieldOffsetPx=0.0, interlaceCombStrength=0.0, interlaceLineWeighting=0.0)),
])
def test_identity_settings(pass_id, kw):
    img = np.full((H, W, 3), (0.3, 0.55, 0.8), np.float32) if pass_id == 1 else rand_img(3, 0.0, 1.0)
    for f in (0, 5, 77):
        out = rco.post_pass(pass_id, rgba(img), W, H, make_params(**kw), t0(frameIndex=f, time=f / 29.97))[:, :3]
        assert np.max(np.abs(out - img.reshape(-1, 3))) < 2e-6, (pass_id, f)


def test_p09_geometric_series_at_a_step_edge():
    k = 0.35
    A = np.array([0.9, 0.2, 0.1], np.float32)
    B = np.array([0.1, 0.3, 0.9], np.float32)
    x0 = 10
    img = np.zeros((H, W, 3), np.float32)
    img[:, :x0] = A
    img[:, x0:] = B
    out = rco.post_pass(9, rgba(img), W, H, make_params(chromaBleedFactor=k), t0())[:, :3].reshape(H, W, 3)
    ca, cb = (M5 @ A.astype(np.float64))[1:], (M5 @ B.astype(np.float64))[1:]
    for m in range(0, 20):
        x = x0 + m
        wb = (1 - k ** (m + 1)) / (1 - k ** 24)  # weight of the B side (taps 0..m)
        chroma = wb * cb + (1 - wb) * ca
        y = (M5 @ B.astype(np.float64))[0]
        exp = I5 @ np.array([y, chroma[0], chroma[1]])
        assert np.max(np.abs(out[5, x] - exp)) < 2e-6


def _flat_gbuf(depth=5.0, normal=(0.0, 1.0, 0.0), albedo=(1.0, 1.0, 1.0)):
    gb = np.zeros((H * W, 4), np.float32)
    gb[:, 0:3] = normal
    gb[:, 3] = depth
    ga = np.ones((H * W, 4), np.float32)
    ga[:, 0:3] = albedo
    return gb, ga


def test_atrous_reduces_to_b3_atrous_when_edge_stopping_is_off():
    # sigmaNormal 0 (wn = 1), flat G-buffer (wd = 1), huge variance and sigmaColor
    # (wc = 1): every pass is the textbook B3-spline a-trous (Dammertz 2010) with
    # clamped colour taps at offsets (i, j) * 2^pass.
    img = rand_img(5, 0.05, 1.0)
    gb, ga = _flat_gbuf()
    var = np.full(H * W, 1e6, np.float32)
    p = make_params(atrousSigmaNormal=0.0, atrousSigmaColor=1e10, atrousSigmaDepth=1.0)
    h = np.array([1, 4, 6, 4, 1], np.float64) / 16
    cur = img.astype(np.float64)
    src = rgba(img)
    for step in (1, 2, 4):
        out = rco.post_pass(100 + int(np.log2(step)), src, W, H, p, t0(), f16(gb), f16(ga), var, total=1)
        ref = np.zeros_like(cur)
        for j in range(-2, 3):
            for i in range(-2, 3):
                yy = np.clip(np.arange(H) + j * step, 0, H - 1)
                xx = np.clip(np.arange(W) + i * step, 0, W - 1)
                ref += h[i + 2] * h[j + 2] * cur[yy][:, xx]
        assert np.max(np.abs(out[:, :3].reshape(H, W, 3) - ref)) < 2e-6, step
        cur = ref
        src = out


def test_atrous_border_modes_colour_clamps_gbuffer_wraps():
    # Colour is 1 in the last column, 0 elsewhere; columns 0-1 of the G-buffer
    # are at depth 50 (rest 5). For x = W-1 the taps x+1, x+2 clamp to column W-1
    # for colour (value 1) but wrap to columns 0-1 for the G-buffer, whose depth
    # weight is ~0, so only taps -2..0 count: result = h2 / (h0 + h1 + h2).
    img = np.zeros((H, W, 3), np.float32)
    img[:, W - 1] = 1.0
    gb, ga = _flat_gbuf()
    gb.reshape(H, W, 4)[:, 0:2, 3] = 50.0
    var = np.full(H * W, 1e6, np.float32)
    p = make_params(atrousSigmaNormal=0.0, atrousSigmaColor=1e10, atrousSigmaDepth=0.15)
    out = rco.post_pass(100, rgba(img), W, H, p, t0(), f16(gb), f16(ga), var, total=1)
    v = out.reshape(H, W, 4)[H // 2, W - 1, 0]
    assert abs(v - (6 / 16) / (1 / 16 + 4 / 16 + 6 / 16)) < 1e-6
    # with a clamped (non-wrapping) G-buffer the answer would be (6+4+1)/16 = 11/16


def test_atrous_sky_pixels_pass_through():
    img = rand_img(6)
    gb, ga = _flat_gbuf(depth=1e4)
    var = np.full(H * W, 0.5, np.float32)
    out = rco.post_pass(101, rgba(img), W, H, make_params(), t0(), f16(gb), f16(ga), var, total=4)
    assert np.array_equal(out[:, :3], img.reshape(-1, 3))


def test_post_chain_order_and_rgba8():
    img = rand_img(7, 0.0, 2.0)
    gb, ga = _flat_gbuf(albedo=(0.5, 0.6, 0.7))
    var = np.random.default_rng(8).uniform(0, 4, H * W).astype(np.float32)
    p = make_params(enable=[1, 2, 3, 4, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16], atrous=True, iterations=5)
    t = t0(time=1.3, agcGain=1.5, frameIndex=45)
    hdr, rgba8 = rco.post_chain(rgba(img) * 3.0, W, H, 3, p, t, f16(gb), f16(ga), var)
    # same result pass by pass
    cur = rco.post_pass(0, rgba(img) * 3.0, W, H, p, t, total=3)
    for i in range(5):
        cur = rco.post_pass(100 + i, cur, W, H, p, t, f16(gb), f16(ga), var, total=3)
    for k in (1, 2, 3, 4, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16):
        cur = rco.post_pass(k, cur, W, H, p, t, total=3)
    assert np.array_equal(hdr, cur)
    disp = rco.post_pass(200, cur, W, H, p, t)
    assert np.array_equal(rgba8[:, :3], np.floor(np.clip(disp[:, :3], 0, 1) * 255 + 0.5).astype(np.uint8))
