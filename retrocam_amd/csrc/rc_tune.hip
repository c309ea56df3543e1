// rc_tune.hip — the autotuner of the lane-scheduler path tracer: times launch
// variants on the uploaded scene at a camera's resolution and keeps the
// fastest. Every candidate computes the same image (the variants differ in
// occupancy, LDS stack depth, triangles per walk step, the wait threshold that
// starts a shading round, and the shape of the 64-pixel unit tiles), so
// tuning changes speed only. The search starts from the context's tuning and
// moves only for a gain above kMinGain (2%): candidates within timing noise of
// each other do not change the choice from run to run. Project-new: the reference
// has no counterpart.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <vector>

#include "rc_internal.h"
#include "rc_shade.h"
#include "retrocam_hip.h"

namespace rc {
namespace {

struct Candidate {
    int variant, threshold, tileShift;
    int32_t mode;  // RC_PT_LANES, or RC_PT_MEGAKERNEL
    bool operator==(const Candidate &o) const {
        return variant == o.variant && threshold == o.threshold && tileShift == o.tileShift && mode == o.mode;
    }
};

// A candidate replaces the best so far only when it is this much faster
// (launch-to-launch spread of one candidate is about 1%).
constexpr float kMinGain = 0.02f;

// Milliseconds per sample of `samples` samples from a reset accumulation,
// between events on the context's stream (the reset is outside the window).
int time_candidate(rc_ctx *ctx, const rc_camera &cam, const rc_pt_params &p, int samples, const Candidate &c,
                   hipEvent_t a, hipEvent_t b, float &msPerSample) {
    ctx->laneVariant = c.variant;
    ctx->laneThresh = c.threshold;
    ctx->laneTileShift = c.tileShift;
    ctx->ptMode = c.mode;
    int rc = rc_reset(ctx);
    if (rc) return rc;
    hipError_t e = hipEventRecord(a, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipEventRecord");
    if ((rc = rc_render_passes(ctx, &cam, &p, samples))) return rc;
    e = hipEventRecord(b, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    if (e != hipSuccess) return hip_fail(ctx, e, "autotune timing");
    msPerSample = ms / (float)samples;
    return RC_OK;
}

}  // namespace
}  // namespace rc

using namespace rc;

extern "C" int rc_autotune(rc_ctx *ctx, const rc_camera *cam, const rc_pt_params *p, int32_t samplesPerTrial,
                           rc_tuning *out, float *msPerSample) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (int rc = capture_block(ctx, "rc_autotune")) return rc;
    if (!cam || !p) return set_error(ctx, RC_ERR_INVALID_ARG, "camera or params is NULL");
    if (samplesPerTrial < 1 || samplesPerTrial > 4096)
        return set_error(ctx, RC_ERR_INVALID_ARG, "samplesPerTrial must be in 1..4096");
    if (!ctx->scene.valid) return set_error(ctx, RC_ERR_STATE, "no scene uploaded");
    if (!lanes_supported(ctx->scene))
        return set_error(ctx, RC_ERR_STATE, "the scene's BVH is 30 or more levels deep: the lane scheduler (and so "
                                             "its tuning) does not apply; the megakernel renders it");
    (void)hipSetDevice(ctx->device);
    const Candidate keep{ctx->laneVariant, ctx->laneThresh, ctx->laneTileShift, ctx->ptMode};
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e != hipSuccess) {
        if (a) (void)hipEventDestroy(a);
        return hip_fail(ctx, e, "hipEventCreate");
    }
    int rc = RC_OK;
    const Candidate start{keep.variant, keep.threshold, keep.tileShift, RC_PT_LANES};
    Candidate best = start;
    float bestMs = FLT_MAX;
    auto trial = [&](const Candidate &c) {
        if (!(bestMs == FLT_MAX) && c == best) return;  // already timed
        float ms = 0.0f;
        if (rc == RC_OK) rc = time_candidate(ctx, *cam, *p, 1, c, a, b, ms);  // first launch of a variant: warm-up
        if (rc == RC_OK) rc = time_candidate(ctx, *cam, *p, samplesPerTrial, c, a, b, ms);
        if (rc == RC_OK && (bestMs == FLT_MAX || ms < bestMs * (1.0f - kMinGain))) {
            bestMs = ms;
            best = c;
        }
    };
    // an untimed run at the trial size first: it grows the sample buffer and
    // warms the caches, which would otherwise be charged to the first candidate
    float warm = 0.0f;
    rc = time_candidate(ctx, *cam, *p, samplesPerTrial, start, a, b, warm);
    // coordinate descent from the current tuning: kernel variant, then wait
    // threshold, then tile shape; last, the megakernel, which wins when every
    // lane has a single short unit (e.g. one-sample passes over a small scene)
    trial(start);
    for (int v = 0; v < RC_LANE_VARIANTS; ++v) trial(Candidate{v, best.threshold, best.tileShift, RC_PT_LANES});
    for (int th : {16, 24, 32, 40, 48, 56}) trial(Candidate{best.variant, th, best.tileShift, RC_PT_LANES});
    for (int sh = 2; sh <= 5; ++sh) trial(Candidate{best.variant, best.threshold, sh, RC_PT_LANES});
    trial(Candidate{best.variant, best.threshold, best.tileShift, RC_PT_MEGAKERNEL});
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    const Candidate use = rc == RC_OK ? best : keep;
    ctx->laneVariant = use.variant;
    ctx->laneThresh = use.threshold;
    ctx->laneTileShift = use.tileShift;
    ctx->ptMode = use.mode;
    if (rc) return rc;
    if ((rc = rc_reset(ctx))) return rc;
    if (out && (rc = rc_get_tuning(ctx, out))) return rc;
    if (msPerSample) *msPerSample = bestMs;
    return RC_OK;
}
