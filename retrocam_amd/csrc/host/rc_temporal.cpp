// rc_temporal.cpp — per-video-frame inputs of the render path, with Java's
// binary32 semantics: the camcorder's temporal state (auto-gain and auto-focus
// IIR filters, focus hunting), keyframe tracks, and the video job's frame
// timing. Each function cites the Java method it restates.
//
// Arithmetic follows the Java source operation by operation: float operations
// stay in binary32 (the build uses -ffp-contract=off, so no fused multiply-add
// is formed), Math.exp / Math.sin are evaluated in binary64 and rounded to
// float where Java casts, and Math.min / Math.max keep Java's NaN rule.
#include <cmath>
#include <cstdint>
#include <new>
#include <vector>

#include "rc_host.h"

namespace {

// java.lang.Math.min(float, float) / max(float, float) for the operands these
// filters see: NaN if either operand is NaN (signed zeros never reach them).
float java_minf(float a, float b) {
    if (a != a) return a;
    if (b != b) return b;
    return a <= b ? a : b;
}

float java_maxf(float a, float b) {
    if (a != a) return a;
    if (b != b) return b;
    return a >= b ? a : b;
}

// One step of both IIR filters (TemporalState.java:47-63 and :99-110, which
// differ only in how `time` advances; the caller sets s->time first).
void advance_filters(rct_state *s, float dt, const rct_settings *st) {
    // agcAlpha = 1 - (float) exp(-agcSpeed * dt)
    const float agcAlpha = 1.0f - (float)std::exp((double)(-st->agcSpeed * dt));
    s->agcGain += (s->agcGainTarget - s->agcGain) * agcAlpha;
    s->agcGain = java_maxf(1.0f, java_minf(st->agcMaxGain, s->agcGain));

    const float afAlpha = 1.0f - (float)std::exp((double)(-st->afSpeed * dt));
    s->focalDist += (s->focalDistTarget - s->focalDist) * afAlpha;

    // hunting: two incommensurate sinusoids in binary64, rounded once, scaled by afAlpha
    const double t = (double)s->time;
    const float hunt = (float)(std::sin(t * 7.31) * 0.006 + std::sin(t * 13.77) * 0.004) * afAlpha;
    s->focalDist = java_maxf(0.1f, s->focalDist + hunt);
}

struct Keyframe {
    float time;
    float value;
    int32_t interp;
};

int32_t clamp_interp(int32_t i) {  // Interpolation.fromIndex (KeyframeTrack.java:29-31)
    return i < RCK_STEP ? RCK_STEP : (i > RCK_BEZIER ? RCK_BEZIER : i);
}

}  // namespace

struct rck_track {
    std::vector<Keyframe> keys;  // ascending time
};

namespace {

// KeyframeTrack.catmullRom (KeyframeTrack.java:146-178): cubic Hermite between
// keys lo and hi with tangents from the neighbours (one-sided at the ends).
float catmull_rom(const std::vector<Keyframe> &k, int lo, int hi, float t) {
    const Keyframe &k0 = k[lo], &k1 = k[hi];
    const float dt = k1.time - k0.time;
    float m0;
    if (lo > 0) {
        const Keyframe &kp = k[lo - 1];
        m0 = (k1.value - kp.value) / (k1.time - kp.time);
    } else {
        m0 = (k1.value - k0.value) / dt;
    }
    float m1;
    if (hi < (int)k.size() - 1) {
        const Keyframe &kn = k[hi + 1];
        m1 = (kn.value - k0.value) / (kn.time - k0.time);
    } else {
        m1 = (k1.value - k0.value) / dt;
    }
    const float t2 = t * t, t3 = t2 * t;
    const float h00 = 2.0f * t3 - 3.0f * t2 + 1.0f;
    const float h10 = t3 - 2.0f * t2 + t;
    const float h01 = -2.0f * t3 + 3.0f * t2;
    const float h11 = t3 - t2;
    return h00 * k0.value + h10 * dt * m0 + h01 * k1.value + h11 * dt * m1;
}

}  // namespace

extern "C" {

// ---- TemporalState ----------------------------------------------------------

void rct_init(rct_state *s) {  // field initialisers, TemporalState.java:16-29
    if (!s) return;
    s->agcGain = 1.0f;
    s->focalDist = 2.0f;
    s->time = 0.0f;
    s->agcGainTarget = 1.0f;
    s->focalDistTarget = 2.0f;
    for (int i = 0; i < 9; ++i) s->whiteBalance[i] = (i % 4 == 0) ? 1.0f : 0.0f;
}

void rct_update(rct_state *s, float dt, const rct_settings *st) {  // TemporalState.update, :44-64
    if (!s || !st) return;
    s->time += dt;
    advance_filters(s, dt, st);
}

void rct_reset_for_video(rct_state *s, const rct_settings *st) {  // resetForVideoRendering, :77-84
    if (!s || !st) return;
    rct_init(s);
    s->focalDist = st->focusDistM;
    s->focalDistTarget = st->focusDistM;
}

void rct_update_for_video_frame(rct_state *s, float canonicalTime, float frameDt,
                                const rct_settings *st) {  // updateForVideoFrame, :96-111
    if (!s || !st) return;
    s->time = canonicalTime;  // set, not accumulated
    advance_filters(s, frameDt, st);
}

int32_t rct_run_video(rct_state *s, const rct_settings *st, const float *focusTargets, float fps, int32_t nFrames,
                      float *outTime, float *outAgcGain, float *outFocalDist) {
    if (!s || !st || nFrames < 0) return -1;
    const float dt = rcj_frame_dt(fps);
    for (int32_t i = 0; i < nFrames; ++i) {
        if (focusTargets) s->focalDistTarget = focusTargets[i];
        rct_update_for_video_frame(s, rcj_canonical_time(i, dt), dt, st);
        if (outTime) outTime[i] = s->time;
        if (outAgcGain) outAgcGain[i] = s->agcGain;
        if (outFocalDist) outFocalDist[i] = s->focalDist;
    }
    return nFrames;
}

// ---- RenderJob timing ---------------------------------------------------------

float rcj_frame_dt(float fps) {  // RenderJob.getFrameDt, RenderJob.java:39-41
    return fps > 0.0f ? 1.0f / fps : 1.0f / 29.97f;
}

int32_t rcj_total_frames(float durationSeconds, float fps) {  // RenderJob.getTotalFrames, :34-37 (video)
    // Math.round(float): the closest int, ties towards +infinity (exactly, in
    // binary64, with no rounding of x + 0.5); NaN -> 0; saturating to the int range
    const float x = durationSeconds * fps;
    if (x != x) return 1;
    const double xd = (double)x, f = std::floor(xd);
    const double r = (xd - f >= 0.5) ? f + 1.0 : f;
    const long long n = r >= 2147483647.0 ? 2147483647LL : (r <= -2147483648.0 ? -2147483648LL : (long long)r);
    return n < 1 ? 1 : (int32_t)n;
}

float rcj_canonical_time(int32_t frame, float frameDt) {  // RenderPipeline.java:150
    return (float)frame * frameDt;
}

// ---- KeyframeTrack ------------------------------------------------------------

rck_track *rck_new(void) { return new (std::nothrow) rck_track(); }

void rck_free(rck_track *t) { delete t; }

int32_t rck_add(rck_track *t, float time, float value, int32_t interp) {  // addKeyframe, :71-83
    if (!t) return -1;
    interp = clamp_interp(interp);
    for (size_t i = 0; i < t->keys.size(); ++i) {
        Keyframe &k = t->keys[i];
        if (std::fabs(k.time - time) < 0.001f) {
            k.value = value;
            k.interp = interp;
            return (int32_t)i;
        }
    }
    size_t i = 0;
    while (i < t->keys.size() && t->keys[i].time < time) ++i;
    t->keys.insert(t->keys.begin() + (long)i, Keyframe{time, value, interp});
    return (int32_t)i;
}

int32_t rck_remove(rck_track *t, int32_t index) {  // removeKeyframe, :85-88
    if (!t) return -1;
    if (index >= 0 && index < (int32_t)t->keys.size()) t->keys.erase(t->keys.begin() + index);
    return (int32_t)t->keys.size();
}

int32_t rck_count(const rck_track *t) { return t ? (int32_t)t->keys.size() : -1; }

int32_t rck_get(const rck_track *t, int32_t index, float *time, float *value, int32_t *interp) {
    if (!t || index < 0 || index >= (int32_t)t->keys.size()) return -1;
    const Keyframe &k = t->keys[index];
    if (time) *time = k.time;
    if (value) *value = k.value;
    if (interp) *interp = k.interp;
    return 0;
}

float rck_evaluate(const rck_track *t, float time, float defaultValue) {  // evaluate, :102-125
    if (!t) return defaultValue;
    const std::vector<Keyframe> &k = t->keys;
    const int n = (int)k.size();
    if (n == 0) return defaultValue;
    if (n == 1) return k[0].value;
    if (time <= k[0].time) return k[0].value;
    if (time >= k[n - 1].time) return k[n - 1].value;
    int hi = 1;
    while (hi < n && k[hi].time < time) ++hi;
    const int lo = hi - 1;
    const Keyframe &k0 = k[lo], &k1 = k[hi];
    const float localT = (time - k0.time) / (k1.time - k0.time);
    switch (k0.interp) {
    case RCK_STEP: return k0.value;
    case RCK_LINEAR: return k0.value + (k1.value - k0.value) * localT;
    default: return catmull_rom(k, lo, hi, localT);
    }
}

int32_t rck_evaluate_many(const rck_track *t, const float *times, int32_t n, float defaultValue, float *out) {
    if (!t || n < 0 || (n > 0 && (!times || !out))) return -1;
    for (int32_t i = 0; i < n; ++i) out[i] = rck_evaluate(t, times[i], defaultValue);
    return n;
}

}  // extern "C"
