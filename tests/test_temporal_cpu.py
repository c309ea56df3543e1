"""Per-video-frame inputs: TemporalState (TemporalState.java:44-111), keyframe
tracks (KeyframeTrack.java:71-178) and video timing (RenderJob.java:34-41).

The oracle (oracle/rc_oracle_temporal.py) is pinned first by what the
mathematics fixes: the IIR filters' closed form (at t = 0 the hunting term is
exactly 0), the hunting waveform itself (with alpha = 1 the focal distance is
target + hunt(t)), the clamps, Catmull-Rom's exactness on linear and (interior,
uniformly spaced) quadratic data, pass-through at keys, and Math.round's
tie rule. librc_host.so's C++ restatement is then compared with the oracle bit
for bit on seeded random sequences."""
import math

import numpy as np
import pytest

from oracle import rc_oracle_temporal as O
from retrocam_amd import temporal as T
from retrocam_amd.scene import orbit_camera

F = np.float32


def bits(x):
    return np.float32(x).view(np.uint32)


# ---- the oracle against closed forms -----------------------------------------------------

@pytest.mark.parametrize("target", [1.0, 2.5, 5.9])
def test_oracle_agc_iir_closed_form(target):
    s = O.Settings(agcSpeed=1.5, agcMaxGain=6.0)
    st = O.TemporalState()
    st.agcGainTarget = F(target)
    dt = 1.0 / 30.0
    alpha = 1.0 - math.exp(-1.5 * float(F(dt)))
    for n in range(1, 120):
        st.update_for_video_frame(0.0, dt, s)
        want = target + (1.0 - target) * (1.0 - alpha) ** n
        assert abs(float(st.agcGain) - want) <= 2e-6 * max(1.0, target) * n ** 0.5 + 1e-7


def test_oracle_agc_clamps():
    s = O.Settings(agcSpeed=50.0, agcMaxGain=4.0)
    st = O.TemporalState()
    st.agcGainTarget = F(100.0)
    for _ in range(50):
        st.update(F(0.1), s)
        assert 1.0 <= st.agcGain <= 4.0
    assert st.agcGain == F(4.0)
    st.agcGainTarget = F(0.01)
    for _ in range(50):
        st.update(F(0.1), s)
        assert st.agcGain >= 1.0
    assert st.agcGain == F(1.0)


@pytest.mark.parametrize("f0,target", [(2.0, 5.0), (8.0, 0.5)])
def test_oracle_focus_iir_closed_form(f0, target):
    # canonical time 0 each frame: sin(0) = 0, so no hunting; a pure IIR
    s = O.Settings(focusDistM=f0, afSpeed=3.0)
    st = O.TemporalState()
    st.reset_for_video(s)
    st.focalDistTarget = F(target)
    dt = F(1.0 / 24.0)
    alpha = 1.0 - math.exp(-3.0 * float(dt))
    for n in range(1, 100):
        st.update_for_video_frame(0.0, dt, s)
        want = target + (f0 - target) * (1.0 - alpha) ** n
        assert abs(float(st.focalDist) - want) <= 4e-6 * max(f0, target) * n ** 0.5


def test_oracle_focus_hunting_waveform():
    # afSpeed * dt = 40: (float) exp(-40) < 2^-25, so afAlpha rounds to 1 and
    # focalDist = max(0.1, target + (0.006 sin(7.31 t) + 0.004 sin(13.77 t)))
    s = O.Settings(afSpeed=1000.0)
    st = O.TemporalState()
    st.focalDistTarget = F(1.5)
    for i in range(200):
        t = F(i * 0.0137)
        st.update_for_video_frame(t, 0.04, s)
        hunt = 0.006 * math.sin(float(t) * 7.31) + 0.004 * math.sin(float(t) * 13.77)
        assert abs(float(st.focalDist) - (1.5 + hunt)) <= 2e-7
        assert st.time == t
    # the waveform stays within +-0.01 m and reaches past 0.009 somewhere
    ts = np.arange(0, 20, 0.001)
    h = 0.006 * np.sin(ts * 7.31) + 0.004 * np.sin(ts * 13.77)
    assert np.max(np.abs(h)) <= 0.01 and np.max(np.abs(h)) > 0.009


def test_oracle_focus_floor_and_time_accumulation():
    s = O.Settings(afSpeed=10.0)
    st = O.TemporalState()
    st.focalDistTarget = F(0.0)
    dt = F(0.05)
    for n in range(1, 400):
        st.update(dt, s)
        assert st.focalDist >= F(0.1)
        assert abs(float(st.time) - n * float(dt)) <= 1e-6 * n
    assert abs(float(st.focalDist) - 0.1) < 0.011


def test_oracle_reset_for_video():
    s = O.Settings(focusDistM=3.25)
    st = O.TemporalState()
    st.agcGain, st.time, st.agcGainTarget = F(3.0), F(9.0), F(2.0)
    st.reset_for_video(s)
    assert (st.agcGain, st.focalDist, st.time, st.agcGainTarget, st.focalDistTarget) == (1.0, 3.25, 0.0, 1.0, 3.25)


def _track(keys, interp):
    k = O.KeyframeTrack()
    for t, v in keys:
        k.add(t, v, interp)
    return k


def test_oracle_catmull_rom_reproduces_lines_and_passes_through_keys():
    rng = np.random.default_rng(3)
    ts = np.sort(rng.uniform(0, 10, 7)).astype(F)
    a, b = 0.75, -1.25
    k = _track([(t, a + b * float(t)) for t in ts], O.BEZIER)
    for t in np.linspace(ts[0], ts[-1], 301, dtype=F):
        assert abs(float(k.evaluate(t, 0.0)) - (a + b * float(t))) <= 5e-5
    for t, v, _ in k.keys:
        assert k.evaluate(t, 0.0) == v


def test_oracle_catmull_rom_exact_for_quadratics_on_interior_segments():
    # Uniform spacing: the centred tangents are the quadratic's derivative, and
    # cubic Hermite interpolation with exact derivatives reproduces degree <= 3.
    q = lambda t: 0.5 * t * t - 2.0 * t + 1.0
    ts = [0.0, 1.0, 2.0, 3.0, 4.0, 5.0]
    k = _track([(t, q(t)) for t in ts], O.BEZIER)
    for t in np.linspace(1.0, 4.0, 61, dtype=F):
        assert abs(float(k.evaluate(t, 0.0)) - q(float(t))) <= 1e-5
    # the end segments use one-sided tangents: exact only on lines
    assert abs(float(k.evaluate(0.5, 0.0)) - q(0.5)) > 1e-3


def test_oracle_step_linear_clamp_and_edit_rules():
    k = _track([(1.0, 10.0), (3.0, 30.0), (2.0, 20.0)], O.LINEAR)
    assert [kk[0] for kk in k.keys] == [1.0, 2.0, 3.0]                       # sorted insert
    assert k.evaluate(1.5, 0.0) == F(15.0) and k.evaluate(2.75, 0.0) == F(27.5)
    assert k.evaluate(-5.0, 0.0) == F(10.0) and k.evaluate(9.0, 0.0) == F(30.0)  # clamped
    assert k.add(2.0004, 99.0, O.STEP) == 1 and len(k.keys) == 3             # within 0.001 s: overwrite
    assert k.evaluate(2.5, 0.0) == F(99.0)                                   # STEP holds the earlier key
    assert k.add(2.0011, 5.0, 7) == 2 and k.keys[2][2] == O.BEZIER           # new key; interp clamped
    assert k.remove(17) == 4 and k.remove(0) == 3
    e = O.KeyframeTrack()
    assert e.evaluate(1.0, 4.5) == F(4.5)
    e.add(2.0, 7.0, O.LINEAR)
    assert e.evaluate(-1.0, 4.5) == F(7.0) and e.evaluate(9.0, 4.5) == F(7.0)


def test_oracle_job_timing():
    assert O.total_frames(1.0, 29.97) == 30
    assert O.total_frames(2.5, 30.0) == 75
    assert O.total_frames(0.0, 30.0) == 1
    assert O.total_frames(2.5, 1.0) == 3          # ties round up
    assert O.total_frames(-2.5, 1.0) == 1         # Math.round(-2.5f) = -2, then max(1, .)
    assert O.total_frames(float(np.nextafter(F(0.5), F(0))), 1.0) == 1
    assert O.frame_dt(0.0) == F(1.0) / F(29.97) and O.frame_dt(25.0) == F(0.04)
    assert O.canonical_time(3, O.frame_dt(25.0)) == F(3) * F(0.04)


# ---- librc_host.so against the oracle, bit for bit ----------------------------------

def test_host_temporal_state_matches_oracle():
    rng = np.random.default_rng(11)
    for trial in range(20):
        kw = dict(focusDistM=float(rng.uniform(0.2, 8)), afSpeed=float(rng.uniform(0.5, 20)),
                  agcMaxGain=float(rng.uniform(1, 8)), agcSpeed=float(rng.uniform(0.2, 10)))
        so, sh = O.Settings(**kw), T.TemporalSettings(**kw)
        o, h = O.TemporalState(), T.TemporalState()
        o.reset_for_video(so)
        h.reset_for_video_rendering(sh)
        fps = float(rng.choice([24.0, 25.0, 29.97, 30.0, 60.0]))
        dt = O.frame_dt(fps)
        for i in range(120):
            tgt = float(rng.uniform(0.05, 10)) if i % 17 == 0 else None
            gt = float(rng.uniform(0.5, 9)) if i % 23 == 0 else None
            if tgt is not None:
                o.focalDistTarget = F(tgt)
                h.set_focal_dist_target(tgt)
            if gt is not None:
                o.agcGainTarget = F(gt)
                h.set_agc_gain_target(gt)
            if trial % 2:
                o.update(dt, so)
                h.update(float(dt), sh)
            else:
                t = O.canonical_time(i, dt)
                o.update_for_video_frame(t, dt, so)
                h.update_for_video_frame(float(t), float(dt), sh)
            for name in ("agcGain", "focalDist", "time"):
                assert bits(getattr(o, name)) == bits(getattr(h, name)), (trial, i, name)


def test_host_run_video_matches_per_frame_updates():
    s = T.TemporalSettings(focusDistM=1.0, afSpeed=4.0)
    targets = np.linspace(1.0, 6.0, 90, dtype=F)
    a = T.TemporalState()
    a.reset_for_video_rendering(s)
    a.set_agc_gain_target(3.0)
    tm, gain, fd = a.run_video(s, 30.0, 90, focus_targets=targets)
    o = O.TemporalState()
    o.reset_for_video(O.Settings(focusDistM=1.0, afSpeed=4.0))
    o.agcGainTarget = F(3.0)
    dt = O.frame_dt(30.0)
    for i in range(90):
        o.focalDistTarget = targets[i]
        o.update_for_video_frame(O.canonical_time(i, dt), dt, O.Settings(focusDistM=1.0, afSpeed=4.0))
        assert (bits(tm[i]), bits(gain[i]), bits(fd[i])) == (bits(o.time), bits(o.agcGain), bits(o.focalDist))


def test_host_job_timing_matches_oracle():
    rng = np.random.default_rng(5)
    cases = [(1.0, 29.97), (2.5, 30.0), (0.0, 30.0), (2.5, 1.0), (-2.5, 1.0), (1e30, 1e30), (float("nan"), 1.0),
             (float(np.nextafter(F(0.5), F(0))), 1.0)]
    cases += [(float(d), float(f)) for d, f in zip(rng.uniform(0, 100, 200), rng.uniform(1, 120, 200))]
    cases += [(float(k) + 0.5, 1.0) for k in range(-4, 40)]
    for d, f in cases:
        assert T.total_frames(d, f) == O.total_frames(d, f), (d, f)
    for f in (0.0, -1.0, 24.0, 29.97, 59.94, 1000.0):
        assert bits(T.frame_dt(f)) == bits(O.frame_dt(f))
        for i in (0, 1, 7, 1234, 99999):
            assert bits(T.canonical_time(i, O.frame_dt(f))) == bits(O.canonical_time(i, O.frame_dt(f)))


@pytest.mark.parametrize("interp", [O.STEP, O.LINEAR, O.BEZIER, "mixed"])
def test_host_keyframe_tracks_match_oracle(interp):
    rng = np.random.default_rng(7 + (3 if interp == "mixed" else int(interp)))
    for trial in range(25):
        o, h = O.KeyframeTrack(), T.KeyframeTrack(None, "x", T.TargetType.RENDER_SETTINGS)
        n = int(rng.integers(0, 9))
        for _ in range(n):
            t = float(rng.choice([rng.uniform(-2, 12), float(rng.integers(0, 10))]))
            v = float(rng.normal(0, 5))
            it = int(rng.integers(-1, 4)) if interp == "mixed" else int(interp)
            assert o.add(t, v, it) == h.add_keyframe(t, v, it)
        if n > 2 and trial % 3 == 0:
            idx = int(rng.integers(-1, n + 1))
            o.remove(idx)
            h.remove_keyframe(idx)
        assert [(bits(a), bits(b), c) for a, b, c in o.keys] == \
               [(bits(a), bits(b), int(c)) for a, b, c in h.keyframes()]
        times = np.concatenate([rng.uniform(-3, 13, 200), [k[0] for k in o.keys]]).astype(F)
        got = h.evaluate_many(times, 0.625)
        want = np.array([o.evaluate(t, 0.625) for t in times], F)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), trial
        for t in times[:5]:
            assert bits(h.evaluate(float(t), 0.625)) == bits(o.evaluate(t, 0.625))


# ---- the Python host layer -------------------------------------------------------------

def test_timeline_applies_tracks_with_setter_clamps():
    lens, rig, post = T.LensSettings(), T.OrbitRig(), T.PostSettings()
    tl = T.KeyframeTimeline(4.0)
    tr = tl.add_track(lens, "lens.focusDist", T.TargetType.RENDER_SETTINGS)
    assert tl.add_track(lens, "lens.focusDist", T.TargetType.RENDER_SETTINGS) is tr   # one track per property
    tr.add_keyframe(0.0, 2.0)
    tr.add_keyframe(2.0, -1.0)                 # below the 0.1 m floor of the setter
    ph = tl.add_track(rig, "cam.phi", T.TargetType.CAMERA)
    ph.add_keyframe(0.0, 1.0, T.Interpolation.STEP)
    ph.add_keyframe(1.0, 5.0, T.Interpolation.STEP)
    ta = tl.add_track(post, "post.tapeAge", T.TargetType.RENDER_SETTINGS)
    ta.add_keyframe(0.0, 0.5)
    ta.add_keyframe(1.0, 1.5)
    with pytest.raises(KeyError):
        tl.add_track(rig, "cam.roll", T.TargetType.CAMERA)
    tl.apply(0.5)
    assert lens.focusDistM == pytest.approx(1.25) and rig.phi == pytest.approx(1.0) and post.tapeAge == 1.0
    tl.apply(3.0)
    assert lens.focusDistM == float(F(0.1))
    assert rig.phi == float(F(F(math.pi) - F(0.02)))  # OrbitCamera clamps phi below pi - 0.02
    assert tl.has_tracks_on(post) and not tl.has_scene_object_tracks() and not tl.is_empty()
    assert tl.find_track(rig, "cam.phi") is ph
    tl.remove_track(ph)
    assert tl.find_track(rig, "cam.phi") is None and tl.track_count() == 2


def test_rig_camera_and_post_settings():
    rig = T.OrbitRig(theta=0.3, phi=1.1, radius=7.0, targetX=0.5, targetY=4.0, targetZ=-1.0)
    lens = T.LensSettings(focalLengthMm=50.0, apertureFStop=2.8, saStrength=0.1)
    cam = rig.camera(64, 48, lens, 3.5)
    ref = orbit_camera(64, 48, theta=0.3, phi=1.1, radius=7.0, target=(0.5, 4.0, -1.0), focalDist=3.5,
                       focalLengthMm=50.0, fStop=2.8, saStrength=0.1)
    assert cam == ref
    from retrocam_amd import _native as N
    pp = N.RcPostParams()
    T.PostSettings(exposure=1.5, colorSaturation=0.8, colorTempBias=-0.25, tapeAge=0.6).write_to(pp)
    assert (pp.exposure, pp.colorSaturation, pp.colorTempBias, pp.tapeAge) == (1.5, F(0.8), -0.25, F(0.6))


def test_video_job_frame_paths_and_count():
    from retrocam_amd.video import VideoJob
    j = VideoJob(64, 48, fps=25.0, duration_seconds=0.5, output_path="/tmp/clip")
    assert j.total_frames() == 13  # 12.5 frames: Math.round ties up
    assert j.frame_path(7) == "/tmp/clip_frames/frame_00007.png"
    assert VideoJob(8, 8).frame_path(0) is None
