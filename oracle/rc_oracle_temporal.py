"""CPU oracle of the per-video-frame inputs: the camcorder's temporal state
(TemporalState, src/main/java/com/retrocam/core/TemporalState.java:13-111),
keyframe tracks (KeyframeTrack, src/main/java/com/retrocam/keyframe/KeyframeTrack.java:20-178)
and the video job's frame timing (RenderJob.java:34-41, RenderPipeline.java:150-151).

TEST INFRASTRUCTURE ONLY: used by tests/ to check librc_host.so's rct_*, rck_*
and rcj_* functions. Nothing in the product imports it.

Java semantics, written out one operation at a time: float arithmetic is
binary32 (every intermediate is rounded with numpy.float32), Math.exp and
Math.sin take and return binary64 (Python floats), a (float) cast rounds once,
and Math.min / Math.max return NaN when an operand is NaN.
"""
import math

import numpy as np

F = np.float32
STEP, LINEAR, BEZIER = 0, 1, 2


def _jmin(a, b):
    if a != a:
        return a
    if b != b:
        return b
    return a if a <= b else b


def _jmax(a, b):
    if a != a:
        return a
    if b != b:
        return b
    return a if a >= b else b


class Settings:
    """RenderSettings fields read by the filters (RenderSettings.java:28,32,42,43)."""

    def __init__(self, focusDistM=2.0, afSpeed=3.0, agcMaxGain=6.0, agcSpeed=1.5):
        self.focusDistM = F(focusDistM)
        self.afSpeed = F(afSpeed)
        self.agcMaxGain = F(agcMaxGain)
        self.agcSpeed = F(agcSpeed)


class TemporalState:
    def __init__(self):  # field initialisers, TemporalState.java:16-29
        self.agcGain = F(1.0)
        self.focalDist = F(2.0)
        self.time = F(0.0)
        self.agcGainTarget = F(1.0)
        self.focalDistTarget = F(2.0)

    def _filters(self, dt, s):
        # TemporalState.java:48-50 / :99-101
        agc_alpha = F(F(1.0) - F(math.exp(float(F(-s.agcSpeed) * F(dt)))))
        self.agcGain = F(self.agcGain + F(F(self.agcGainTarget - self.agcGain) * agc_alpha))
        self.agcGain = _jmax(F(1.0), _jmin(s.agcMaxGain, self.agcGain))
        # :53-54 / :103-104
        af_alpha = F(F(1.0) - F(math.exp(float(F(-s.afSpeed) * F(dt)))))
        self.focalDist = F(self.focalDist + F(F(self.focalDistTarget - self.focalDist) * af_alpha))
        # :59-63 / :106-110
        t = float(self.time)
        hunt = F(F(math.sin(t * 7.31) * 0.006 + math.sin(t * 13.77) * 0.004) * af_alpha)
        self.focalDist = _jmax(F(0.1), F(self.focalDist + hunt))

    def update(self, dt, s):  # :44-64
        self.time = F(self.time + F(dt))
        self._filters(dt, s)

    def reset_for_video(self, s):  # :77-84
        self.__init__()
        self.focalDist = F(s.focusDistM)
        self.focalDistTarget = F(s.focusDistM)

    def update_for_video_frame(self, canonical_time, frame_dt, s):  # :96-111
        self.time = F(canonical_time)
        self._filters(frame_dt, s)


def frame_dt(fps):  # RenderJob.getFrameDt, RenderJob.java:39-41
    fps = F(fps)
    return F(F(1.0) / fps) if fps > 0 else F(F(1.0) / F(29.97))


def total_frames(duration, fps):  # RenderJob.getTotalFrames, :34-37 (video formats)
    with np.errstate(over="ignore"):
        x = F(F(duration) * F(fps))  # binary32 product; may overflow to inf
    if x != x:
        return 1  # Math.round(NaN) = 0
    if math.isinf(x):
        return 2 ** 31 - 1 if x > 0 else 1
    # Math.round(float): closest int, ties up (x + 0.5 is exact in binary64 below 2^52; above, x is an integer)
    r = max(-2 ** 31, min(2 ** 31 - 1, math.floor(float(x) + 0.5)))
    return max(1, int(r))


def canonical_time(frame, dt):  # RenderPipeline.java:150
    return F(F(frame) * F(dt))


class KeyframeTrack:
    def __init__(self):
        self.keys = []  # [time, value, interp], ascending time

    def add(self, time, value, interp):  # addKeyframe, KeyframeTrack.java:71-83
        time, value = F(time), F(value)
        interp = max(STEP, min(BEZIER, int(interp)))
        for i, k in enumerate(self.keys):
            if abs(F(k[0] - time)) < F(0.001):
                k[1], k[2] = value, interp
                return i
        i = 0
        while i < len(self.keys) and self.keys[i][0] < time:
            i += 1
        self.keys.insert(i, [time, value, interp])
        return i

    def remove(self, index):  # removeKeyframe, :85-88
        if 0 <= index < len(self.keys):
            del self.keys[index]
        return len(self.keys)

    def evaluate(self, time, default):  # evaluate, :102-125
        time = F(time)
        k = self.keys
        n = len(k)
        if n == 0:
            return F(default)
        if n == 1:
            return k[0][1]
        if time <= k[0][0]:
            return k[0][1]
        if time >= k[n - 1][0]:
            return k[n - 1][1]
        hi = 1
        while hi < n and k[hi][0] < time:
            hi += 1
        lo = hi - 1
        t0, v0, interp = k[lo]
        t1, v1, _ = k[hi]
        lt = F(F(time - t0) / F(t1 - t0))
        if interp == STEP:
            return v0
        if interp == LINEAR:
            return F(v0 + F(F(v1 - v0) * lt))
        return self._catmull_rom(lo, hi, lt)

    def _catmull_rom(self, lo, hi, t):  # :146-178
        k = self.keys
        t0, v0 = k[lo][0], k[lo][1]
        t1, v1 = k[hi][0], k[hi][1]
        dt = F(t1 - t0)
        if lo > 0:
            m0 = F(F(v1 - k[lo - 1][1]) / F(t1 - k[lo - 1][0]))
        else:
            m0 = F(F(v1 - v0) / dt)
        if hi < len(k) - 1:
            m1 = F(F(k[hi + 1][1] - v0) / F(k[hi + 1][0] - t0))
        else:
            m1 = F(F(v1 - v0) / dt)
        t2 = F(t * t)
        t3 = F(t2 * t)
        h00 = F(F(F(F(2) * t3) - F(F(3) * t2)) + F(1))
        h10 = F(F(t3 - F(F(2) * t2)) + t)
        h01 = F(F(F(-2) * t3) + F(F(3) * t2))
        h11 = F(t3 - t2)
        a = F(h00 * v0)
        b = F(F(h10 * dt) * m0)
        c = F(h01 * v1)
        d = F(F(h11 * dt) * m1)
        return F(F(F(a + b) + c) + d)
