"""CPU pins of the VHS passes of the post oracle (rc_oracle_post.c) in their
active regimes, restated independently from the shaders' mathematics (float64
numpy, Python-integer hashes that reproduce tests/golden/hash_vectors.json):

- p03 (post/p03_color_matrix.frag:56-90): saturation, hue rotation, temperature
  and drift as a chroma-plane map between the 3-decimal BT.601 matrices;
  saturation 0 gives grey;
- p04 (post/p04_dynamic_range.frag:52-98): the crush / lift / knee transfer
  curve sampled against its piecewise formula, chroma scaled with luma;
- p07 (post/p07_ccd_noise.frag:45-115): the exact PCG noise field and row
  banding;
- p08 (post/p08_ccd_smear.frag:63-112): a single bright row against the closed
  form primary (below) and 0.7^i secondary (above) sums, with the edge counts;
- p10 (post/p10_timebase.frag:44-79): a horizontal ramp comes back shifted by
  the closed-form row offset;
- p12 (post/p12_dropout.frag:59-200): burst windows and event boxes recomputed
  from the hashes; every pixel outside them unchanged, every pixel inside equal
  to its displaced, stretched, brightened sample;
- p13 (post/p13_headswitch.frag:78-191): rows at or above the switch zone
  unchanged; ripple, chroma rotation and jitter in the zone;
- p14 (post/p14_tracking.frag:83-269): tracking bands recomputed from the epoch
  hashes; rows outside unchanged; band interiors displaced by the closed-form
  warp, which stays within the severity-scaled maximum;
- p16 (post/p16_interlace.frag:81-141): field line weights, the half-line
  vertical offset and the comb shift on even lines;
- p01 / p02 (post/p01_downsample.frag:36-53, p02_chroma_res.frag:55-85): the
  9-tap RGB and the radius-clamp(2.5 sigma) chroma Gaussians as edge-clamped
  convolutions; p11 (post/p11_dot_crawl.frag:48-105): the gated I/Q subcarrier;
  p15 (post/p15_edge_enhance.frag:80-120): binomial unsharp mask with coring.

Each pin also runs against oracle builds with one planted mistake."""
import importlib.util
import json
import os

import numpy as np
import pytest

from oracle import rco
from retrocam_amd.post import TemporalInputs, make_params
from oracle_variants import pin_fails

HERE = os.path.dirname(os.path.abspath(__file__))
_spec = importlib.util.spec_from_file_location("gen_hash_vectors", os.path.join(HERE, "golden", "gen_hash_vectors.py"))
HV = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(HV)
M32 = 0xFFFFFFFF
W, H = 64, 48

M3 = np.array([[0.299, 0.587, 0.114], [-0.169, -0.331, 0.5], [0.5, -0.419, -0.081]])
M5 = np.array([[0.299, 0.587, 0.114], [-0.16874, -0.33126, 0.5], [0.5, -0.41869, -0.08131]])


def rgb3(y, cb, cr):
    return np.stack([y + 1.402 * cr, y - 0.344 * cb - 0.714 * cr, y + 1.772 * cb], -1)


def rgb5(y, cb, cr):
    return np.stack([y + 1.402 * cr, y - 0.34414 * cb - 0.71414 * cr, y + 1.772 * cb], -1)


def fhash(kind, a, b):
    """fHash of p12 (kind 1, one-round uHash) or p13/p14 (kind 2) as float32."""
    u = HV.uhash1 if kind == 1 else HV.uhash2
    return float(np.float32(u((a & M32) ^ ((u(b & M32) * 2654435761) & M32))) / np.float32(4294967295.0))


def run(pass_id, img, t=TemporalInputs(), **kw):
    a = np.ones((img.shape[0] * img.shape[1], 4), np.float32)
    a[:, :3] = img.reshape(-1, 3)
    out = rco.post_pass(pass_id, a, img.shape[1], img.shape[0], make_params(**kw), t.to_c())
    return out[:, :3].reshape(img.shape).astype(np.float64)


def rand_img(seed, lo=0.05, hi=0.95):
    return np.random.default_rng(seed).uniform(lo, hi, (H, W, 3)).astype(np.float32)


def ramp_x(slope=0.004, base=0.2):
    x = np.arange(W, dtype=np.float64)
    return np.repeat(np.repeat((base + slope * x)[None, :, None], H, 0), 3, 2).astype(np.float32)


def smooth(e0, e1, x):
    t = np.clip((x - e0) / (e1 - e0), 0.0, 1.0)
    return t * t * (3 - 2 * t)


def test_python_hashes_reproduce_the_golden_vectors():
    d = json.load(open(os.path.join(HERE, "golden", "hash_vectors.json")))
    assert all(HV.pcg_hash(a) == h for a, h in d["pcg"])
    for k, key in ((1, "fhash1"), (2, "fhash2")):
        for a, b, bits in d[key]:
            assert np.array([fhash(k, a, b)], np.float32).view(np.uint32)[0] == bits


# ---- p03 ---------------------------------------------------------------------
def pin_p03_colour_matrix():
    img = rand_img(40).astype(np.float64)
    for sat, hue, temp, time, drift in ((0.0, 0.0, 0.0, 0.0, 0.0), (0.0, 30.0, 0.0, 2.0, 0.0),
                                        (1.3, 25.0, 0.0, 0.0, 0.0), (0.7, -140.0, 1.0, 3.0, 0.4),
                                        (1.0, 0.0, -2.0, 11.0, 1.3)):
        out = run(3, img.astype(np.float32), TemporalInputs(time=time), colorSaturation=sat, colorHueRot=hue,
                  colorTempBias=temp, colorDriftSpeed=drift)
        q = img @ M3.T
        cb, cr = q[..., 1] * sat, q[..., 2] * sat            # chroma centred at 0 (the +0.5 cancels)
        th = np.radians(hue)
        cb, cr = np.cos(th) * cb - np.sin(th) * cr, np.sin(th) * cb + np.cos(th) * cr
        shift = temp * 0.05 + np.sin(np.float32(time) * np.float32(drift)) * 0.015
        exp = np.clip(rgb3(q[..., 0], cb - 0.75 * shift, cr + shift), 0, 1)
        assert np.max(np.abs(out - exp)) < 3e-6, (sat, hue, temp, np.max(np.abs(out - exp)))
        if sat == 0.0 and shift == 0.0:
            assert np.max(np.abs(out - out[..., :1])) < 2e-6, "saturation 0 gives grey"


# ---- p04 ---------------------------------------------------------------------
def transfer(y, crush, lift, knee, clip):
    t = y / max(crush, 1e-4)
    y1 = np.where(y < crush, t * t * (3 - 2 * t) * crush, y)
    y2 = lift + (1 - lift) * y1
    tk = (y2 - knee) / max(1 - knee, 1e-4)
    return np.where(y2 > knee, np.minimum(knee + (clip - knee) * (1 - np.exp(-3.5 * tk)), clip), y2)


def pin_p04_transfer_curve():
    g = np.linspace(0.0, 1.2, W * H).reshape(H, W)
    grey = np.repeat(g[..., None], 3, 2)
    col = rand_img(41, 0.0, 1.1).astype(np.float64)
    for crush, lift, knee, clip in ((0.03, 0.01, 0.78, 0.95), (0.2, 0.1, 0.5, 0.8), (0.0, 0.0, 0.3, 1.0)):
        kw = dict(drShadowCrush=crush, drBlackLift=lift, drHighlightKnee=knee, drHighlightClip=clip)
        for img in (grey, col):
            out = run(4, img.astype(np.float32), **kw)
            q = img @ M3.T
            yin = q[..., 0]
            y = transfer(yin, crush, lift, knee, clip)
            sc = np.where(yin > 0.001, y / np.where(yin > 0.001, yin, 1), 1.0)
            exp = np.clip(rgb3(y, q[..., 1] * sc, q[..., 2] * sc), 0, 1)
            assert np.max(np.abs(out - exp)) < 3e-6, (crush, knee, np.max(np.abs(out - exp)))


# ---- p07 ---------------------------------------------------------------------
def p07_rand(px, py, seed):
    h = HV.pcg_hash(((px & M32) ^ ((py * 1619) & M32) ^ ((seed * 2654435761) & M32)) & M32)
    return float(np.float32(h) / np.float32(4294967295.0)) * 2.0 - 1.0


def pin_p07_noise_field():
    img = rand_img(42).astype(np.float64)
    for f, agc, base, cs in ((0, 1.0, 0.015, 1.6), (37, 2.5, 0.03, 0.8), (1000, 0.7, 0.0, 2.0)):
        out = run(7, img.astype(np.float32), TemporalInputs(frameIndex=f, agcGain=agc), ccdNoiseLumaBase=base,
                  ccdNoiseChromaScale=cs)
        q = img @ M5.T
        lum = np.clip(q[..., 0], 0, 1)
        la = (base + 0.04 * (1 - lum)) * agc
        ca = la * cs
        nY = np.array([[p07_rand(x, y, f) for x in range(W)] for y in range(H)])
        nB = np.array([[p07_rand(x, y, f + 137) for x in range(W)] for y in range(H)])
        nR = np.array([[p07_rand(x, y, f + 271) for x in range(W)] for y in range(H)])
        band = np.array([p07_rand(0, y * 7 + f // 8, 0) * 0.005 * agc for y in range(H)])[:, None]
        exp = rgb5(q[..., 0] + nY * la + band, q[..., 1] + nB * ca, q[..., 2] + nR * ca)
        assert np.max(np.abs(out - exp)) < 3e-6, (f, np.max(np.abs(out - exp)))


# ---- p08 ---------------------------------------------------------------------
def pin_p08_single_bright_row():
    thr, inten, v = 0.3, 0.5, 0.9
    for L, y0 in ((20, 30), (7, 45), (150, 12), (3, 1)):
        img = np.zeros((H, W, 3), np.float32)
        img[y0] = v
        out = run(8, img, ccdSmearLength=L, ccdSmearThreshold=thr, ccdSmearIntensity=inten)
        slen = max(1, int(L * 0.3))
        for y in range(H):
            prim = sec = 0.0
            i = y0 - y
            if 1 <= i <= L:
                prim = (v - thr) / min(L, H - 1 - y) * inten          # taps up to the image top
            j = y - y0
            if 1 <= j <= slen:
                sec = (v - thr) * 0.7 ** j / min(slen, y) * inten * 0.25  # taps down to row 0
            exp = min(max((v if y == y0 else 0.0) + prim + sec, 0.0), 1.0)
            assert np.max(np.abs(out[y] - exp)) < 2e-6, (L, y0, y, out[y, 0], exp)


# ---- p10 ---------------------------------------------------------------------
def p10_hash(x):
    s = np.float32(np.sin(np.float64(np.float32(np.float32(x) * np.float32(127.1)) + np.float32(311.7))))
    v = np.float32(s * np.float32(43758.5453))
    return float(v - np.floor(v))


def pin_p10_ramp_shift():
    img = ramp_x()
    for A, freq, speed, time, f in ((1.5, 0.02, 0.7, 0.0, 0), (2.5, 0.3, 1.1, 4.2, 13), (0.8, 0.05, 0.0, 0.0, 7)):
        out = run(10, img, TemporalInputs(time=time, frameIndex=f), timebaseAmplitudePx=A, timebaseFreq=freq,
                  timebaseSpeed=speed)
        fb = np.float32(f // 4) * np.float32(0.137)
        for y in range(H):
            sy = np.float32((y + 0.5) / H) * np.float32(H)
            sp = np.sin(np.float64(np.float32(sy * np.float32(freq) + np.float32(time) * np.float32(speed))))
            nx = np.float32(sy * np.float32(0.05) + fb)
            i, fr = np.floor(nx), nx - np.floor(nx)
            u = fr * fr * (3 - 2 * fr)
            nz = (p10_hash(i) * (1 - u) + p10_hash(i + 1) * u) * 2 - 1
            off = A * (sp * 0.5 + nz * 0.5)
            xs = np.arange(4, W - 4)
            exp = 0.2 + 0.004 * (xs + off)
            assert np.max(np.abs(out[y, 4:W - 4, 0] - exp)) < 3e-5, (A, f, y, off)


# ---- p12 ---------------------------------------------------------------------
def p12_events(frame, rate, age):
    """Event boxes of one frame in evaluation order: (y0, h, x0, w, dx, dy, stretch)."""
    vf = frame // 2
    win = vf // 30
    thr = 1.0 - rate
    this_a = 0.0 if fhash(1, win, 0xDEADBEEF) < thr else 1.0
    prev_a = 0.0 if fhash(1, (win - 1) & M32, 0xDEADBEEF) < thr else 1.0
    act = max(this_a, prev_a * float(smooth(0.7, 1.0, 1.0 - (vf % 30) / 30.0)))
    if act < 0.001:
        return act, []
    n = 1 + int(np.round(act * (0.05 + age * 0.95) * 13.0))
    ev = []
    for i in range(n):
        y0 = np.floor(fhash(1, vf * 31 + i, i * 13 + 1) * H)
        h = 2 if fhash(1, vf + i * 3, i + 77) > 0.88 else 1
        wr = fhash(1, vf * 7 + i, i * 7 + 2)
        w = 80 + wr * 140 if fhash(1, vf * 2 + i, i + 55) > 0.88 else 8 + wr * wr * 32
        x0 = fhash(1, vf * 5 + i, i * 7 + 3) * max(0.0, W - w)
        dx = (fhash(1, vf * 11 + i, i * 7 + 4) * 2 - 1) * (4 + age * 16)
        dyr = fhash(1, vf * 13 + i, i * 7 + 5)
        dy = (-1.0 if dyr > 0.8 else 1.0) if dyr > 0.6 else 0.0
        st = fhash(1, vf * 17 + i, i * 7 + 6)
        stretch = 1 + st * age * 0.5 if st > 0.7 else 1.0
        ev.append((y0, h, x0, w, dx, dy, stretch))
    return act, ev


def pin_p12_dropout_events():
    img = ramp_x(slope=0.005, base=0.1)
    rate, age, bright = 0.6, 0.9, 0.92
    active = checked = 0
    for frame in range(0, 400, 7):
        out = run(12, img, TemporalInputs(frameIndex=frame), dropoutBurstRate=rate, tapeAge=age, dropoutBright=bright)
        act, ev = p12_events(frame, rate, age)
        active += bool(ev)
        exp = img.astype(np.float64).copy()
        skip = np.zeros((H, W), bool)
        for y in range(H):
            for x in range(W):
                fx = x + 0.5
                for (y0, h, x0, w, dx, dy, stretch) in ev:
                    if not (y0 <= y < y0 + h):
                        continue
                    if not (x0 <= fx < x0 + w):
                        continue
                    dL, dR = fx - x0, x0 + w - fx
                    wt = float(smooth(0, 1.5, dL) * smooth(0, 5, dR))
                    if wt < 0.001:
                        break
                    cx = x0 + w * 0.5
                    sx = cx + (fx - cx) * stretch + dx
                    if not (1.0 <= sx <= W - 1.0):
                        skip[y, x] = True          # clamped sample: not a closed form
                        break
                    rep = 0.1 + 0.005 * (sx - 0.5)  # the ramp at texel position sx - 0.5
                    bb = float(smooth(3, 0, dL)) * bright
                    rep *= 1 + bb * 0.3                # grey: the desaturation changes nothing
                    exp[y, x] = exp[y, x] * (1 - wt) + rep * wt
                    break                             # the first event on the pixel decides
        assert np.max(np.abs(out - exp)[~skip]) < 2e-5, (frame, np.max(np.abs(out - exp)[~skip]))
        untouched = np.all(exp == img, axis=2) & ~skip
        assert np.array_equal(out[untouched], img[untouched].astype(np.float64)), frame
        checked += int((~untouched).sum())
    assert active >= 5 and checked > 200, (active, checked)


# ---- p13 ---------------------------------------------------------------------
def p13_noise(x, seed):
    i = np.floor(x)
    f = x - i
    u = f * f * (3 - 2 * f)
    return (fhash(2, int(i), seed) * (1 - u) + fhash(2, int(i) + 1, seed) * u) * 2 - 1


def pin_p13_head_switch():
    const = np.full((H, W, 3), (0.35, 0.5, 0.65), np.float32)
    ramp = ramp_x()
    rnd = rand_img(43)
    for f, zone, R, C, J in ((0, 12, 0.1, 0.35, 4.0), (77, 20, 0.25, -0.8, 2.5), (5, 1, 0.1, 0.35, 4.0)):
        kw = dict(headSwitchLines=zone, headSwitchRipple=R, headSwitchChroma=C, headSwitchJitter=J)
        t = TemporalInputs(frameIndex=f)
        out = run(13, rnd, t, **kw)
        assert np.array_equal(out[zone:], rnd[zone:].astype(np.float64)), "rows above the switch zone unchanged"
        # constant image: ripple on luma, chroma rotated by the settle-scaled angle
        out = run(13, const, t, **kw)
        q = const[0, 0].astype(np.float64) @ M5.T
        pr = fhash(2, f, 0xCAFEBABE) * 6.283185
        sgn = 1.0 if fhash(2, f, 0xBEEF) > 0.5 else -1.0
        for y in range(zone):
            settle = (1 - y / zone) ** 2
            rip = R * settle * np.sin(y * 1.570796 + pr)
            a = C * settle * sgn
            cb = np.cos(a) * q[1] - np.sin(a) * q[2]
            cr = np.sin(a) * q[1] + np.cos(a) * q[2]
            exp = rgb5(q[0] + rip, cb, cr)
            assert np.max(np.abs(out[y] - exp)) < 3e-6, (f, y)
        # ramp, no ripple or chroma: the row comes back shifted by the jitter
        out = run(13, ramp, t, **{**kw, "headSwitchRipple": 0.0, "headSwitchChroma": 0.0})
        for y in range(zone):
            settle = (1 - y / zone) ** 2
            jit = J * settle * (p13_noise(np.float32(y * 0.8), f) * 0.7 + p13_noise(np.float32(y * 3.0), f + 7) * 0.3)
            xs = np.arange(6, W - 6)
            assert np.max(np.abs(out[y, 6:W - 6, 0] - (0.2 + 0.004 * (xs + jit)))) < 3e-6, (f, y, jit)


# ---- p14 ---------------------------------------------------------------------
def p14_noise(x, seed):
    i = np.floor(x)
    f = x - i
    u = f * f * (3 - 2 * f)
    return fhash(2, int(i), seed) * (1 - u) + fhash(2, int(i) + 1, seed) * u


def p14_row(frame, y, sev, maxd, fringe):
    """(weight, displacement px, fringe) of the event that decides row y."""
    ce = frame // 30
    best = (0.0, 0.0, 0.0)
    for e in range(2):
        for ep in range(3):
            epoch = (ce - ep) & M32
            if fhash(2, epoch * 7919 + e * 3, 0xDEADBEEF) > 0.08 + 0.62 * sev:
                continue
            dur = int(15 + fhash(2, epoch * 1009 + e, 0xCAFE) * 105)
            st = (epoch * 30 + int(fhash(2, epoch * 503 + e, 0xBEEF) * 30)) & M32
            if not (st <= frame < st + dur):
                continue
            eh = 20 + fhash(2, epoch * 3001 + e, 0x2222) * 60 * sev
            y0 = fhash(2, epoch * 2017 + e, 0x1111) * (H - eh)
            y1 = y0 + eh
            if not (y0 <= y < y1):
                continue
            bt = (y - y0) / max(eh - 1, 1)
            ds = 1.0 if fhash(2, epoch * 4003 + e, 0x3333) > 0.5 else -1.0
            base = ds * maxd * sev * (bt - 0.5) * 2
            wob = np.sin(bt * 6.28318 + (frame * 0.37 + e * 1.57)) * maxd * 0.2 * sev
            mj = (p14_noise(np.float32(y * 0.25), frame + e * 17) * 2 - 1) * 2.5 * sev
            dT, dB = y - y0, y1 - y
            ew = float(smooth(0, 1.5, dT) * smooth(0, 4, dB))
            fr = (1 - float(smooth(0, fringe, min(dT, dB)))) * ew
            if ew > best[0]:
                best = (ew, base + wob + mj, fr)
    return best


def pin_p14_tracking_bands():
    sev, maxd, fringe = 0.8, 6.0, 4.0
    img = ramp_x(slope=0.003, base=0.3)
    rnd = rand_img(44)
    banded = 0
    for frame in range(0, 900, 31):
        kw = dict(trackingSeverity=sev, trackingMaxDisplacePx=maxd, trackingFringeWidthPx=fringe)
        t = TemporalInputs(frameIndex=frame)
        out_r = run(14, rnd, t, **kw)
        out = run(14, img, t, **kw)
        for y in range(H):
            w, disp, fr = p14_row(frame, y, sev, maxd, fringe)
            if w < 0.001:
                assert np.array_equal(out_r[y], rnd[y].astype(np.float64)), (frame, y)
                continue
            assert abs(disp) <= maxd * sev * 1.2 + 2.5 * sev + 1e-4
            if w == 1.0 and fr == 0.0:             # band interior: a pure horizontal shift
                xs = np.arange(12, W - 12)
                exp = 0.3 + 0.003 * (xs + disp)
                assert np.max(np.abs(out[y, 12:W - 12, 0] - exp)) < 2e-5, (frame, y, disp)
                banded += 1
    assert banded >= 40, banded


# ---- p16 ---------------------------------------------------------------------
def pin_p16_interlace():
    rnd = rand_img(45, 0.2, 0.8)
    for wgt in (0.05, 0.3):
        out = run(16, rnd, interlaceFieldOffsetPx=0.0, interlaceCombStrength=0.0, interlaceLineWeighting=wgt)
        m = np.where(np.arange(H) % 2 == 0, 1 - wgt, 1 + wgt / 2)[:, None, None]
        assert np.max(np.abs(out - rnd * m)) < 2e-6, wgt
    # I = 0.3 + 0.002 x + 0.01 y: even lines comb by the strength (|ya - yb| = 0.02 is
    # above 3x the edge threshold, yHere > yLeft); every line moves by +-offset/2 rows
    xs, ys = np.meshgrid(np.arange(W), np.arange(H))
    g = (0.3 + 0.002 * xs + 0.01 * ys).astype(np.float32)
    img = np.repeat(g[..., None], 3, 2)
    for S, F, wgt in ((1.5, 0.0, 0.05), (2.0, 0.5, 0.1), (-1.0, 1.2, 0.0)):
        out = run(16, img, interlaceFieldOffsetPx=F, interlaceCombStrength=S, interlaceCombEdgeThresh=0.005,
                  interlaceLineWeighting=wgt)
        for y in range(3, H - 3):
            even = y % 2 == 0
            xv = np.arange(4, W - 4)
            sx = xv + (S if even else 0.0)
            sy = y + (F / 2 if even else -F / 2)
            exp = (0.3 + 0.002 * sx + 0.01 * sy) * ((1 - wgt) if even else (1 + wgt / 2))
            assert np.max(np.abs(out[y, 4:W - 4, 0] - exp)) < 3e-6, (S, F, y)


# ---- p01 / p02 / p11 / p15 ------------------------------------------------------
def hconv(a, taps, weights):
    """sum_i w_i a[:, clamp(x + t_i)] / sum_i w_i along rows (edge-clamped)."""
    out = np.zeros_like(a)
    for t, w in zip(taps, weights):
        out += w * a[:, np.clip(np.arange(a.shape[1]) + t, 0, a.shape[1] - 1)]
    return out / np.sum(weights)


def pin_p01_gaussian():
    img = rand_img(46).astype(np.float64)
    for sigma in (0.6, 1.5, 3.0):
        out = run(1, img.astype(np.float32), lumaBlurSigma=sigma)
        taps = np.arange(-4, 5)
        exp = hconv(img, taps, np.exp(-taps ** 2 / (2 * sigma * sigma)))
        assert np.max(np.abs(out - exp)) < 2e-6, sigma


def pin_p02_chroma_gaussian():
    img = rand_img(47).astype(np.float64)
    for sigma in (0.3, 2.0, 7.0, 20.0):
        out = run(2, img.astype(np.float32), chromaResSigma=sigma)
        sg = max(sigma, 0.5)
        r = min(max(int(sg * 2.5 + 0.5), 1), 32)
        taps = np.arange(-r, r + 1)
        q = img @ M5.T
        wts = np.exp(-taps ** 2 / (2 * sg * sg))
        exp = rgb5(q[..., 0], hconv(q[..., 1], taps, wts), hconv(q[..., 2], taps, wts))
        assert np.max(np.abs(out - exp)) < 3e-6, sigma


def pin_p11_dot_crawl():
    img = rand_img(48, 0.1, 0.9)
    img[:, W // 2:] += 0.05
    img = np.clip(img, 0, 1).astype(np.float32)
    a = img.astype(np.float64)
    l709 = a @ np.array([0.2126, 0.7152, 0.0722])
    for f, inten, sub, thr in ((0, 0.25, 0.14, 0.04), (5, 0.6, 0.3, 0.01), (1001, 0.25, 0.14, 0.2)):
        out = run(11, img, TemporalInputs(frameIndex=f), dotCrawlIntensity=inten, dotCrawlSubcarrier=sub,
                  dotCrawlEdgeThresh=thr)
        xs = np.arange(W)
        ll = l709[:, np.clip(xs - 1, 0, W - 1)]
        lr = l709[:, np.clip(xs + 1, 0, W - 1)]
        grad = np.maximum(np.abs(l709 - ll), np.abs(lr - l709))
        gate = smooth(thr, 3 * thr, grad)
        pi = np.float32(3.14159265)
        hp = ((np.float32(xs + 0.5) * np.float32(sub)) * np.float32(2.0)) * pi
        lp = np.float32(np.arange(H))[:, None] * pi
        ph = ((hp[None, :] + lp) + np.float32(f) * (pi * np.float32(0.5))).astype(np.float64)
        I = np.sin(ph) * inten * gate
        Q = np.cos(ph) * inten * gate * 0.5
        exp = np.clip(a + np.stack([I + 0.9562 * Q, -0.2721 * I - 0.6474 * Q, -1.1069 * I + 1.7046 * Q], -1), 0, 1)
        assert np.max(np.abs(out - exp)) < 2e-5, (f, np.max(np.abs(out - exp)))


def pin_p15_unsharp():
    img = rand_img(49, 0.1, 0.9).astype(np.float64)
    k = np.array([1, 4, 6, 4, 1]) / 16.0
    for amount, core in ((1.5, 0.025), (0.7, 0.2), (2.0, 0.0001)):
        out = run(15, img.astype(np.float32), edgeEnhanceAmount=amount, edgeEnhanceCoreThreshold=core)
        q = img @ M5.T
        Y = q[..., 0]
        blur = np.zeros_like(Y)
        for j in range(-2, 3):
            for i in range(-2, 3):
                blur += k[i + 2] * k[j + 2] * Y[np.clip(np.arange(H) + j, 0, H - 1)][:, np.clip(np.arange(W) + i, 0, W - 1)]
        det = Y - blur
        det = det * smooth(core * 0.5, core, np.abs(det))
        exp = rgb5(Y + amount * det, q[..., 1], q[..., 2])
        assert np.max(np.abs(out - exp)) < 3e-6, (amount, core)


PINS = {f.__name__: f for f in (pin_p03_colour_matrix, pin_p04_transfer_curve, pin_p07_noise_field,
                                pin_p08_single_bright_row, pin_p10_ramp_shift, pin_p12_dropout_events,
                                pin_p13_head_switch, pin_p14_tracking_bands, pin_p16_interlace, pin_p01_gaussian,
                                pin_p02_chroma_gaussian, pin_p11_dot_crawl, pin_p15_unsharp)}


@pytest.mark.parametrize("name", list(PINS))
def test_pin(name):
    PINS[name]()


MISTAKES = [
    ("pin_p03_colour_matrix", "nx = nx - shift * 0.75f;", "nx = nx + shift * 0.75f;"),
    ("pin_p03_colour_matrix", "float nx = ch * cx - sh * cy, ny = sh * cx + ch * cy;",
     "float nx = ch * cx + sh * cy, ny = ch * cy - sh * cx;"),
    ("pin_p03_colour_matrix", "float drift = rc_sinf(t->time * p->colorDriftSpeed) * 0.015f;",
     "float drift = rc_sinf(t->time * p->colorDriftSpeed) * 0.15f;"),
    ("pin_p04_transfer_curve", "rc_expf(-3.5f * t)", "rc_expf(-3.0f * t)"),
    ("pin_p04_transfer_curve", "float Y = crush(yin, p->drShadowCrush);\n    Y = p->drBlackLift + (1.0f - p->drBlackLift) * Y;",
     "float Y = p->drBlackLift + (1.0f - p->drBlackLift) * yin;\n    Y = crush(Y, p->drShadowCrush);"),
    ("pin_p04_transfer_curve", "float sc = (yin > 0.001f) ? (Y / yin) : 1.0f;", "float sc = 1.0f;"),
    ("pin_p07_noise_field", "int rowSeed = y * 7 + f / 8;", "int rowSeed = y * 7 + f / 4;"),
    ("pin_p07_noise_field", "float nCb = p07_rand(x, y, f + 137) * ca;", "float nCb = p07_rand(x, y, f + 138) * ca;"),
    ("pin_p07_noise_field", "0.04f * (1.0f - lum)", "0.04f * lum"),
    ("pin_p08_single_bright_row", "rc_powf(0.70f, (float)i)", "rc_powf(0.75f, (float)i)"),
    ("pin_p08_single_bright_row", "int slen = (int)((float)p->ccdSmearLength * 0.30f);",
     "int slen = (int)((float)p->ccdSmearLength * 0.25f);"),
    ("pin_p08_single_bright_row", "float prim = pt > 0 ? (pa / (float)pt) * p->ccdSmearIntensity : 0.0f;",
     "float prim = pt > 0 ? (pa / (float)p->ccdSmearLength) * p->ccdSmearIntensity : 0.0f;"),
    ("pin_p10_ramp_shift", "float fb = (float)(t->frameIndex / 4) * 0.137f;",
     "float fb = (float)(t->frameIndex / 2) * 0.137f;"),
    ("pin_p10_ramp_shift", "float sp = rc_sinf(sy * p->timebaseFreq + t->time * p->timebaseSpeed);",
     "float sp = rc_sinf(sy * p->timebaseFreq);"),
    ("pin_p10_ramp_shift", "float off = p->timebaseAmplitudePx * (sp * 0.5f + np * 0.5f);",
     "float off = p->timebaseAmplitudePx * (sp * 0.5f - np * 0.5f);"),
    ("pin_p12_dropout_events", "uint32_t vf = (uint32_t)(t->frameIndex / 2);", "uint32_t vf = (uint32_t)t->frameIndex;"),
    ("pin_p12_dropout_events", "int evH = (fhash1(vf + si * 3u, si + 77u) > 0.88f) ? 2 : 1;",
     "int evH = (fhash1(vf + si * 3u, si + 77u) > 0.5f) ? 2 : 1;"),
    ("pin_p12_dropout_events", "float x0 = fhash1(vf * 5u + si, si * 7u + 3u) * maxStart;",
     "float x0 = fhash1(vf * 5u + si, si * 7u + 4u) * maxStart;"),
    ("pin_p12_dropout_events", "float g = 1.0f + bb * 0.3f;", "float g = 1.0f + bb * 0.5f;"),
    ("pin_p13_head_switch", "if (dist >= zone) return fetch(s, x, y);", "if (dist > zone) return fetch(s, x, y);"),
    ("pin_p13_head_switch", "float ripple = (p->headSwitchRipple * settle) * rs;",
     "float ripple = (p->headSwitchRipple * prox) * rs;"),
    ("pin_p13_head_switch", "float sgn = (fhash2(fs, 0xBEEFu) > 0.5f) ? 1.0f : -1.0f;",
     "float sgn = (fhash2(fs, 0xBEEFu) < 0.5f) ? 1.0f : -1.0f;"),
    ("pin_p13_head_switch", "(sn * 0.7f + fn * 0.3f)", "(sn * 0.3f + fn * 0.7f)"),
    ("pin_p14_tracking_bands", "float eh = 20.0f + (hr * 60.0f) * sev;", "float eh = 30.0f + (hr * 60.0f) * sev;"),
    ("pin_p14_tracking_bands", "uint32_t st = epoch * 30u + (uint32_t)(sr * 30.0f);", "uint32_t st = epoch * 30u;"),
    ("pin_p14_tracking_bands", "* sev) * (bt - 0.5f)) * 2.0f;", "* sev) * bt) * 2.0f;"),
    ("pin_p14_tracking_bands", "float fph = (float)frame * 0.37f + (float)ei * 1.57f;",
     "float fph = (float)frame * 0.37f;"),
    ("pin_p16_interlace", "float m = even ? (1.0f - p->interlaceLineWeighting) : (1.0f + p->interlaceLineWeighting * 0.5f);",
     "float m = even ? (1.0f + p->interlaceLineWeighting * 0.5f) : (1.0f - p->interlaceLineWeighting);"),
    ("pin_p16_interlace", "    if (even) {\n        float ew", "    if (!even) {\n        float ew"),
    ("pin_p16_interlace", "float hd = signf(yh - yl);", "float hd = signf(yl - yh);"),
    ("pin_p01_gaussian", "for (int i = -4; i <= 4; ++i) {\n        float w = gaussf((float)i, sigma);\n        c3 c",
     "for (int i = -3; i <= 3; ++i) {\n        float w = gaussf((float)i, sigma);\n        c3 c"),
    ("pin_p02_chroma_gaussian", "int radius = clampi((int)(sigma * 2.5f + 0.5f), 1, 32);",
     "int radius = clampi((int)(sigma * 2.0f + 0.5f), 1, 32);"),
    ("pin_p02_chroma_gaussian", "float sigma = fmaxf(p->chromaResSigma, 0.5f);", "float sigma = fmaxf(p->chromaResSigma, 0.3f);"),
    ("pin_p11_dot_crawl", "float lp = lineY * 3.14159265f;", "float lp = lineY * 1.5707963f;"),
    ("pin_p11_dot_crawl", "c3 c = C3(1.0000f * I + 0.9562f * Q, -0.2721f * I - 0.6474f * Q,",
     "c3 c = C3(1.0000f * I + 0.9562f * Q, -0.2721f * I + 0.6474f * Q,"),
    ("pin_p15_unsharp", "float cw = smoothstepf(p->edgeEnhanceCoreThreshold * 0.5f, p->edgeEnhanceCoreThreshold, fabsf(det));",
     "float cw = smoothstepf(0.0f, p->edgeEnhanceCoreThreshold, fabsf(det));"),
    ("pin_p15_unsharp", "static const float w[5] = {0.0625f, 0.25f, 0.375f, 0.25f, 0.0625f};",
     "static const float w[5] = {0.0f, 0.25f, 0.5f, 0.25f, 0.0f};"),
]


@pytest.mark.parametrize("pin,old,new", MISTAKES, ids=[f"{m[0]}-{i}" for i, m in enumerate(MISTAKES)])
def test_planted_mistakes_fail(pin, old, new):
    assert pin_fails(PINS[pin], "rc_oracle_post.c", old, new), f"{pin} passes with the planted mistake"
