// rc_noalloc.hip — rc_post_image_noalloc: the runOnImage post chain on device
// buffers with no device allocation, for callers that cannot allocate (code
// under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc). The chain's two RGBA32F ping-pong images live in a file-scope
// device array, one copy per GPU, whose address is looked up once per device
// when a context is created there.
#include <mutex>

#include "rc_internal.h"

namespace rc {

// Largest image the entry point takes: 1280 x 720 (the largest serving
// resolution); the array holds two RGBA32F images of that size (29.5 MB per GPU,
// held until the process ends).
constexpr size_t kNoAllocMaxPixels = (size_t)1280 * 720;
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) float4 scratch[2 * kNoAllocMaxPixels];

// Device address of `scratch` on each GPU (the host-side name is only a handle).
static float4 *g_scratch[kNoAllocMaxDevices];
static std::mutex g_scratch_mu;

// Called by rc_create with `device` current: looks the array's address up once
// per device. Returns an RC_* code.
int noalloc_setup(rc_ctx *ctx, int device) {
    if (device < 0 || device >= kNoAllocMaxDevices)
        return set_error(ctx, RC_ERR_LIMIT, "device index beyond the no-allocation scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return RC_OK;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return hip_fail(ctx, e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    g_scratch[device] = (float4 *)ptr;
    return RC_OK;
}

}  // namespace rc

using namespace rc;

extern "C" int rc_post_image_noalloc(rc_ctx *ctx, const float *rgba, int32_t width, int32_t height,
                                     const rc_post_params *p, const rc_temporal *t, float *hdrOut,
                                     uint8_t *rgba8Out) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (!p || !t) return set_error(ctx, RC_ERR_INVALID_ARG, "post params or temporal state is NULL");
    if (p->ccdSmearLength < 0 || p->headSwitchLines < 0)
        return set_error(ctx, RC_ERR_INVALID_ARG, "ccdSmearLength and headSwitchLines must be >= 0");
    if (!rgba || !rgba8Out || width <= 0 || height <= 0)
        return set_error(ctx, RC_ERR_INVALID_ARG, "image pointers or size invalid");
    if ((size_t)width * (size_t)height > kNoAllocMaxPixels)
        return set_error(ctx, RC_ERR_LIMIT,
                         "rc_post_image_noalloc: image larger than its scratch (1280 x 720 = 921600 pixels)");
    if (int rc = capture_block(ctx, "rc_post_image_noalloc")) return rc;
    float4 *s = (ctx->device >= 0 && ctx->device < kNoAllocMaxDevices) ? g_scratch[ctx->device] : nullptr;
    if (!s) return set_error(ctx, RC_ERR_STATE, "no-allocation scratch not set up on this device");
    (void)hipSetDevice(ctx->device);
    const size_t n = (size_t)width * (size_t)height;
    ChainIO io;
    io.src = (const float4 *)rgba;
    io.w = width;
    io.h = height;
    io.total = 1;
    io.a = s;
    io.b = s + n;
    io.display = (uchar4 *)rgba8Out;
    // PostProcessStack.runOnImage (PostProcessStack.java:130-136): p00 with N = 1
    // and exposure 1, no a-trous (no G-buffer).
    rc_post_params q = *p;
    q.exposure = 1.0f;
    int rc = run_chain(ctx, io, q, *t, false);
    if (rc) return rc;
    if (hdrOut) {
        hipError_t e = hipMemcpyAsync(hdrOut, io.out, n * sizeof(float4), hipMemcpyDeviceToDevice, ctx->stream);
        if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpyAsync(post output)");
    }
    return RC_OK;
}
