// rc_debug.hip — instrumentation entry points (no counterpart in the reference):
// device-side PCG streams for the bit-exactness tests of the RNG.
#include "rc_device.h"
#include "rc_internal.h"

namespace rc {

__global__ void pcg_stream_kernel(const int *__restrict__ xyf, int n, int count, uint32_t *__restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t s = init_seed(xyf[3 * i], xyf[3 * i + 1], xyf[3 * i + 2]);
    uint32_t *o = out + (size_t)i * (size_t)(count + 1);
    o[0] = s;
    for (int k = 0; k < count; ++k) o[k + 1] = pcg(s);
}

// rays: 8 floats each {ro.xyz, maxDist, rd.xyz, -}
__global__ void ray_query_kernel(SceneView S, const float4 *__restrict__ rays, int n, int *__restrict__ tri,
                                 float *__restrict__ t, int *__restrict__ unocc) {
    __shared__ int2 stk[kShortStack][kBlock];
    StackRef st{stk, (int)threadIdx.x};
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 a = rays[2 * i], b = rays[2 * i + 1];
    v3 ro = mk(a.x, a.y, a.z), rd = mk(b.x, b.y, b.z);
    Counters C;
    Hit h = closest_hit<false>(S, ro, rd, st, C);
    tri[i] = h.tri;
    t[i] = h.t;
    unocc[i] = unoccluded<false>(S, ro, rd, a.w, st, C) ? 1 : 0;
}

// Every binary32 x with 1e-8 <= |x| <= 2^126: recip_rn(x) against 1.0f / x.
__global__ void recip_check_kernel(uint32_t base, unsigned long long *mismatches) {
    const uint32_t i = base + (uint32_t)(blockIdx.x * 256 + threadIdx.x);
    const float x = __uint_as_float(i);
    if (!(fabsf(x) >= 1e-8f) || fabsf(x) > 0x1p126f) return;
    if (__float_as_uint(recip_rn(x)) != __float_as_uint(1.0f / x)) atomicAdd(mismatches, 1ull);
}

}  // namespace rc

extern "C" int rc_debug_recip_check(rc_ctx *ctx, uint64_t *mismatches) {
    using namespace rc;
    if (!ctx || !mismatches) return RC_ERR_INVALID_ARG;
    if (int rc_ = capture_block(ctx, "rc_debug_recip_check")) return rc_;
    (void)hipSetDevice(ctx->device);
    unsigned long long *d = nullptr;
    hipError_t e = hipMalloc((void **)&d, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemsetAsync(d, 0, sizeof(unsigned long long), ctx->stream);
    const uint32_t chunk = 1u << 28;  // 16 launches of 2^28 bit patterns cover all 2^32
    for (uint32_t k = 0; k < 16 && e == hipSuccess; ++k) {
        hipLaunchKernelGGL(recip_check_kernel, dim3(chunk / 256), dim3(256), 0, ctx->stream, k * chunk, d);
        e = hipGetLastError();
    }
    unsigned long long h = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&h, d, sizeof h, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (d) (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(ctx, e, "rc_debug_recip_check");
    *mismatches = h;
    return RC_OK;
}

extern "C" int rc_debug_rays(rc_ctx *ctx, const float *rays, int32_t n, int32_t *triOut, float *tOut,
                             int32_t *unoccludedOut) {
    using namespace rc;
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (int rc_ = capture_block(ctx, "rc_debug_rays")) return rc_;
    if (!rays || !triOut || !tOut || !unoccludedOut || n <= 0 || n > (1 << 22))
        return set_error(ctx, RC_ERR_INVALID_ARG, "rc_debug_rays: bad arguments");
    if (!ctx->scene.valid) return set_error(ctx, RC_ERR_STATE, "no scene uploaded");
    (void)hipSetDevice(ctx->device);
    float4 *dr = nullptr;
    int *dtri = nullptr, *dun = nullptr;
    float *dt = nullptr;
    hipError_t e = hipMalloc((void **)&dr, (size_t)n * 32);
    if (e == hipSuccess) e = hipMalloc((void **)&dtri, (size_t)n * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&dt, (size_t)n * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&dun, (size_t)n * 4);
    if (e == hipSuccess) e = hipMemcpyAsync(dr, rays, (size_t)n * 32, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(ray_query_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream,
                           scene_view(ctx->scene), (const float4 *)dr, n, dtri, dt, dun);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(triOut, dtri, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(tOut, dt, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(unoccludedOut, dun, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (dr) (void)hipFree(dr);
    if (dtri) (void)hipFree(dtri);
    if (dt) (void)hipFree(dt);
    if (dun) (void)hipFree(dun);
    if (e != hipSuccess) return hip_fail(ctx, e, "rc_debug_rays");
    return RC_OK;
}

extern "C" int rc_debug_pcg(rc_ctx *ctx, const int32_t *xyf, int32_t n, int32_t count, uint32_t *out) {
    using namespace rc;
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (int rc_ = capture_block(ctx, "rc_debug_pcg")) return rc_;
    if (!xyf || !out || n <= 0 || count < 0 || n > (1 << 20) || count > 1024)
        return set_error(ctx, RC_ERR_INVALID_ARG, "rc_debug_pcg: bad arguments");
    (void)hipSetDevice(ctx->device);
    int *dx = nullptr;
    uint32_t *dout = nullptr;
    size_t outBytes = (size_t)n * (size_t)(count + 1) * sizeof(uint32_t);
    hipError_t e = hipMalloc((void **)&dx, (size_t)n * 3 * sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void **)&dout, outBytes);
    if (e == hipSuccess) e = hipMemcpyAsync(dx, xyf, (size_t)n * 3 * sizeof(int), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(pcg_stream_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, dx, n, count, dout);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, outBytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (dx) (void)hipFree(dx);
    if (dout) (void)hipFree(dout);
    if (e != hipSuccess) return hip_fail(ctx, e, "rc_debug_pcg");
    return RC_OK;
}
