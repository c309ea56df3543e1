// rc_capi.hip — the extern "C" entry points of libretrocam_hip.so: context,
// scene upload (wire layout -> device layout), progressive passes and readback.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "rc_internal.h"
#include "rc_shade.h"
#include "retrocam_hip.h"

namespace rc {

static thread_local std::string g_create_error;

int set_error(rc_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    else g_create_error = msg;
    return code;
}

int hip_fail(rc_ctx *ctx, hipError_t e, const char *what) {
    char buf[256];
    std::snprintf(buf, sizeof(buf), "%s failed: %s (%d)", what, hipGetErrorString(e), (int)e);
    return set_error(ctx, e == hipErrorOutOfMemory ? RC_ERR_OOM : RC_ERR_HIP, buf);
}

int capture_block(rc_ctx *ctx, const char *what) {
    if (!ctx->capturing) return RC_OK;
    return set_error(ctx, RC_ERR_STATE,
                     std::string(what) + " is not allowed while a frame is being captured (it would allocate, wait "
                                         "for the device or copy to the host); run the frame once before capturing it");
}

__global__ void fill_cam_kernel(CamArgs *p, int n, CamArgs c) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = c;
}

int camera_slot(rc_ctx *ctx, const CamArgs &c, const CamArgs **out) {
    *out = nullptr;
    if (ctx->capturing) {
        if (ctx->nCams >= kMaxCameraSlots)
            return set_error(ctx, RC_ERR_LIMIT, "a captured frame records at most 4096 path-trace launches and SPPM "
                                                "gathers (camera slots)");
        ctx->hCams[ctx->nCams] = c;
        *out = ctx->dCams + ctx->nCams++;
        return RC_OK;
    }
    CamArgs *slot = ctx->dCams + kMaxCameraSlots;
    hipLaunchKernelGGL(fill_cam_kernel, dim3(1), dim3(64), 0, ctx->stream, slot, 1, c);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(ctx, e, "camera upload");
    *out = slot;
    return RC_OK;
}

const rc_temporal *temporal_slot(rc_ctx *ctx, const rc_temporal &t) {
    if (!ctx->capturing) return nullptr;
    if (ctx->nTemporal >= kMaxTemporalSlots) {
        ctx->slotsExhausted = 1;  // later post chains keep their recorded inputs
        return nullptr;
    }
    ctx->hTemporal[ctx->nTemporal] = t;
    return ctx->dTemporal + ctx->nTemporal++;
}

// Recorded single-view work must follow an rc_reset of the same capture, so that
// every replay recomputes the frame instead of adding to the previous one.
static int capture_needs_reset(rc_ctx *ctx) {
    if (!ctx->capturing || ctx->captureResetSeen) return RC_OK;
    return set_error(ctx, RC_ERR_STATE, "a captured frame must begin with rc_reset before rendering into it");
}

__global__ void fill_temporal_kernel(rc_temporal *p, int n, rc_temporal t) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = t;
}

// Zero fill of up to 8 buffers in one launch (a graph-capturable kernel node;
// no memset nodes in captured frames). end[k] = words in buffers 0..k.
struct ZeroSegs {
    uint32_t *p[8];
    size_t end[8];
    int count;
};

__global__ void zero_segs_kernel(ZeroSegs z) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= z.end[z.count - 1]) return;
    int k = 0;
    while (i >= z.end[k]) ++k;
    z.p[k][i - (k ? z.end[k - 1] : 0)] = 0u;
}

template <typename T>
static void dfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

static void free_scene(DeviceScene &d) {
    dfree(d.inner); dfree(d.triPos); dfree(d.triShade); dfree(d.mats);
    dfree(d.lights); dfree(d.cdf); dfree(d.leafBigCount);
    d = DeviceScene{};
}

static void free_images(Images &im) {
    dfree(im.accum); dfree(im.accum64); dfree(im.gbuf); dfree(im.galbedo);
    dfree(im.variance); dfree(im.primId); dfree(im.gbufId);
    dfree(im.postA); dfree(im.postB); dfree(im.postIn); dfree(im.display);
    im = Images{};
}

template <typename T>
static int dalloc(rc_ctx *ctx, T *&p, size_t count, const char *what);

// Post buffers for a w x h chain (kept until the resolution changes).
static int ensure_post(rc_ctx *ctx, int w, int h) {
    Images &im = ctx->img;
    if (im.postW == w && im.postH == h && im.postA) return RC_OK;
    if (int rc = capture_block(ctx, "allocating the post-chain images")) return rc;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
    ++ctx->freeGen;
    dfree(im.postA); dfree(im.postB); dfree(im.postIn); dfree(im.display);
    im.postOut = nullptr;
    im.postW = im.postH = 0;
    size_t n = (size_t)w * (size_t)h;
    int rc;
    if ((rc = dalloc(ctx, im.postA, n, "hipMalloc(postA)"))) return rc;
    if ((rc = dalloc(ctx, im.postB, n, "hipMalloc(postB)"))) return rc;
    if ((rc = dalloc(ctx, im.postIn, n, "hipMalloc(postIn)"))) return rc;
    if ((rc = dalloc(ctx, im.display, n, "hipMalloc(display)"))) return rc;
    im.postW = w;
    im.postH = h;
    return RC_OK;
}

// PostProcessStack.runChain (PostProcessStack.java:140-280) from the p00 input
// io.src (accumulation sum, or an image with total = 1), then the display pass.
static int run_chain_per_pass(rc_ctx *ctx, ChainIO &io, const rc_post_params &p, const rc_temporal &t,
                              bool withGbuf, const rc_temporal *dt) {
    const int w = io.w, h = io.h;
    float4 *a = io.a, *b = io.b;
    int nl = 1;
    hipError_t e = launch_post_pass(ctx->stream, 0, io.src, a, w, h, p, t, io.gbuf, io.galbedo, io.variance,
                                    io.total, 0, nullptr, dt);
    if (e != hipSuccess) return hip_fail(ctx, e, "post p00 launch");
    bool exceeded = p.atrousMaxSpp > 0 && io.total > p.atrousMaxSpp;
    if (p.atrousEnabled && withGbuf && !exceeded && ctx->denoiser == RC_DENOISE_NET && net_ready(ctx)) {
        int rc = run_net_denoise(ctx, a, io.gbuf, io.galbedo, io.variance, io.total, b, w, h);
        if (rc) return rc;
        float4 *tmp = a; a = b; b = tmp;
    } else if (p.atrousEnabled && withGbuf && !exceeded) {
        int np = p.atrousIterations < 1 ? 1 : (p.atrousIterations > 5 ? 5 : p.atrousIterations);
        for (int i = 0; i < np; ++i) {  // AtrousDenoiser.denoise (AtrousDenoiser.java:60-125)
            e = launch_post_pass(ctx->stream, 100, a, b, w, h, p, t, io.gbuf, io.galbedo, io.variance, io.total,
                                 1 << i, nullptr, dt);
            if (e != hipSuccess) return hip_fail(ctx, e, "post a-trous launch");
            float4 *tmp = a; a = b; b = tmp;
            ++nl;
        }
    }
    for (int k = 1; k <= 16; ++k) {
        if (k == 5 || k == 6 || !p.enable[k]) continue;  // p05/p06 pass through (PostProcessStack.java:184)
        e = launch_post_pass(ctx->stream, k, a, b, w, h, p, t, nullptr, nullptr, nullptr, io.total, 0, nullptr, dt);
        if (e != hipSuccess) return hip_fail(ctx, e, "post pass launch");
        float4 *tmp = a; a = b; b = tmp;
        ++nl;
    }
    e = launch_post_pass(ctx->stream, 200, a, nullptr, w, h, p, t, nullptr, nullptr, nullptr, 1, 0, io.display, dt);
    if (e != hipSuccess) return hip_fail(ctx, e, "display launch");
    io.out = a;
    ctx->postLaunches = nl + 1;
    return RC_OK;
}

// The same chain as fused sweeps (rc_post.hip launch_post_sweeps): p00 is read
// on the fly by the first sweep; an a-trous pass starts a sweep of its own, the
// last one carrying the row-local passes that follow it; with the network, p00
// is written once for the network's input and the VHS passes start from its
// output.
int run_chain(rc_ctx *ctx, ChainIO &io, const rc_post_params &p, const rc_temporal &t, bool withGbuf) {
    const rc_temporal *dt = temporal_slot(ctx, t);
    if (ctx->capturing && !dt) return set_error(ctx, RC_ERR_LIMIT, "captured frame: too many post chains");
    if (ctx->postMode == RC_POST_PER_PASS) return run_chain_per_pass(ctx, io, p, t, withGbuf, dt);
    const int w = io.w, h = io.h;
    int ops[24], steps[24], n = 0, nl = 0, l = 0;
    SweepSource src;
    src.p = io.src;
    src.p00 = true;
    hipError_t e;
    const bool exceeded = p.atrousMaxSpp > 0 && io.total > p.atrousMaxSpp;
    if (p.atrousEnabled && withGbuf && !exceeded && ctx->denoiser == RC_DENOISE_NET && net_ready(ctx)) {
        const float4 *p00out = nullptr;
        e = launch_post_sweeps(ctx->stream, src, nullptr, nullptr, 0, w, h, io.a, io.b, nullptr, p, t, dt, io.gbuf,
                               io.galbedo, io.variance, io.total, &p00out, &l);
        if (e != hipSuccess) return hip_fail(ctx, e, "post p00 sweep");
        nl += l;
        float4 *netOut = (p00out == io.a) ? io.b : io.a;
        int rc = run_net_denoise(ctx, p00out, io.gbuf, io.galbedo, io.variance, io.total, netOut, w, h);
        if (rc) return rc;
        src.p = netOut;
        src.p00 = false;
    } else if (p.atrousEnabled && withGbuf && !exceeded) {
        int np = p.atrousIterations < 1 ? 1 : (p.atrousIterations > 5 ? 5 : p.atrousIterations);
        for (int i = 0; i < np; ++i) {  // AtrousDenoiser.denoise (AtrousDenoiser.java:60-125)
            ops[n] = 100;
            steps[n++] = 1 << i;
        }
    }
    for (int k = 1; k <= 16; ++k) {
        if (k == 5 || k == 6 || !p.enable[k]) continue;  // p05/p06 pass through (PostProcessStack.java:184)
        ops[n] = k;
        steps[n++] = 0;
    }
    const float4 *out = nullptr;
    e = launch_post_sweeps(ctx->stream, src, ops, steps, n, w, h, io.a, io.b, io.display, p, t, dt, io.gbuf,
                           io.galbedo, io.variance, io.total, &out, &l);
    if (e != hipSuccess) return hip_fail(ctx, e, "post sweep launch");
    io.out = out;
    ctx->postLaunches = nl + l;
    return RC_OK;
}

static int run_post_chain(rc_ctx *ctx, const float4 *src, int w, int h, const rc_post_params &p,
                          const rc_temporal &t, int total, bool withGbuf) {
    Images &im = ctx->img;
    ChainIO io;
    io.src = src;
    io.gbuf = im.gbuf;
    io.galbedo = im.galbedo;
    io.variance = im.variance;
    io.w = w;
    io.h = h;
    io.total = total;
    io.a = im.postA;
    io.b = im.postB;
    io.display = im.display;
    int rc = run_chain(ctx, io, p, t, withGbuf);
    if (rc == RC_OK) im.postOut = io.out;
    return rc;
}

static int validate_post(rc_ctx *ctx, const rc_post_params *p, const rc_temporal *t) {
    if (!p || !t) return set_error(ctx, RC_ERR_INVALID_ARG, "post params or temporal state is NULL");
    if (p->ccdSmearLength < 0 || p->headSwitchLines < 0)
        return set_error(ctx, RC_ERR_INVALID_ARG, "ccdSmearLength and headSwitchLines must be >= 0");
    return RC_OK;
}

template <typename T>
static int dalloc(rc_ctx *ctx, T *&p, size_t count, const char *what) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T) > 0 ? count * sizeof(T) : 16);
    if (e != hipSuccess) return hip_fail(ctx, e, what);
    p = (T *)q;
    return RC_OK;
}

static int upload(rc_ctx *ctx, void *dst, const void *src, size_t bytes, const char *what) {
    if (bytes == 0) return RC_OK;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, what);
    return RC_OK;
}

static int ensure_images(rc_ctx *ctx, int w, int h) {
    Images &im = ctx->img;
    bool need64 = ctx->accumMode == RC_ACCUM_FP64;
    if (im.width == w && im.height == h && (!need64 || im.accum64)) return RC_OK;
    if (int rc = capture_block(ctx, "allocating images for a new resolution")) return rc;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
    ++ctx->freeGen;
    free_images(im);
    size_t n = (size_t)w * (size_t)h;
    int rc;
    if ((rc = dalloc(ctx, im.accum, n, "hipMalloc(accum)"))) return rc;
    if ((rc = dalloc(ctx, im.gbuf, n, "hipMalloc(gbuffer)"))) return rc;
    if ((rc = dalloc(ctx, im.galbedo, n, "hipMalloc(galbedo)"))) return rc;
    if ((rc = dalloc(ctx, im.variance, n, "hipMalloc(variance)"))) return rc;
    if ((rc = dalloc(ctx, im.primId, n, "hipMalloc(primId)"))) return rc;
    if ((rc = dalloc(ctx, im.gbufId, n, "hipMalloc(gbufId)"))) return rc;
    if (need64 && (rc = dalloc(ctx, im.accum64, n, "hipMalloc(accum64)"))) return rc;
    // Zero-initialised like a fresh GL texture allocation with NULL data.
    size_t sizes[] = {n * 16, n * 8, n * 8, n * 4, n * 4, n * 4};
    void *ptrs[] = {im.accum, im.gbuf, im.galbedo, im.variance, im.primId, im.gbufId};
    for (int i = 0; i < 6; ++i) {
        e = hipMemsetAsync(ptrs[i], 0, sizes[i], ctx->stream);
        if (e != hipSuccess) return hip_fail(ctx, e, "hipMemsetAsync(images)");
    }
    if (im.accum64) {
        e = hipMemsetAsync(im.accum64, 0, n * 32, ctx->stream);
        if (e != hipSuccess) return hip_fail(ctx, e, "hipMemsetAsync(accum64)");
    }
    im.width = w;
    im.height = h;
    ctx->totalSamples = 0;
    return RC_OK;
}

static inline int32_t bits_i(float f) { int32_t i; std::memcpy(&i, &f, 4); return i; }
static inline float bits_f(int32_t i) { float f; std::memcpy(&f, &i, 4); return f; }

static int validate_camera(rc_ctx *ctx, const rc_camera *cam) {
    if (!cam) return set_error(ctx, RC_ERR_INVALID_ARG, "camera is NULL");
    if (cam->width <= 0 || cam->height <= 0 || cam->width > 16384 || cam->height > 16384)
        return set_error(ctx, RC_ERR_INVALID_ARG, "camera width/height must be in 1..16384");
    if (cam->lensRadius > 0.0f && cam->apertureBlades < 1)
        return set_error(ctx, RC_ERR_INVALID_ARG, "apertureBlades must be >= 1 when lensRadius > 0");
    return RC_OK;
}

}  // namespace rc

using namespace rc;

extern "C" {

int rc_create(int device, rc_ctx **out) {
    if (!out) return set_error(nullptr, RC_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        return set_error(nullptr, RC_ERR_NO_DEVICE, "no HIP device available");
    if (device < 0 || device >= n) return set_error(nullptr, RC_ERR_INVALID_ARG, "device index out of range");
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return set_error(nullptr, RC_ERR_HIP, "hipGetDeviceProperties failed");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        std::string m = "device is ";
        m += prop.gcnArchName;
        m += "; libretrocam_hip.so is built for gfx950 only";
        return set_error(nullptr, RC_ERR_NO_DEVICE, m);
    }
    rc_ctx *ctx = new (std::nothrow) rc_ctx();
    if (!ctx) return set_error(nullptr, RC_ERR_OOM, "host allocation failed");
    ctx->device = device;
    e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return set_error(nullptr, RC_ERR_HIP, "hipStreamCreate failed");
    }
    if (int rc = noalloc_setup(ctx, device)) {  // rc_post_image_noalloc's scratch address on this device
        std::string m = ctx->err;
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return set_error(nullptr, rc, m);
    }
    void *c = nullptr;
    e = hipMalloc(&c, 64 * sizeof(unsigned long long));
    if (e != hipSuccess) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return set_error(nullptr, RC_ERR_OOM, "hipMalloc(counters) failed");
    }
    ctx->counters = (unsigned long long *)c;
    (void)hipMemsetAsync(ctx->counters, 0, 64 * sizeof(unsigned long long), ctx->stream);
    e = hipMalloc((void **)&ctx->dTemporal, sizeof(rc_temporal) * kMaxTemporalSlots);
    if (e == hipSuccess)
        e = hipHostMalloc((void **)&ctx->hTemporal, sizeof(rc_temporal) * kMaxTemporalSlots, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc((void **)&ctx->dCams, sizeof(CamArgs) * (kMaxCameraSlots + 1));
    if (e == hipSuccess)
        e = hipHostMalloc((void **)&ctx->hCams, sizeof(CamArgs) * kMaxCameraSlots, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipFree(ctx->counters);
        if (ctx->dTemporal) (void)hipFree(ctx->dTemporal);
        if (ctx->hTemporal) (void)hipHostFree(ctx->hTemporal);
        if (ctx->dCams) (void)hipFree(ctx->dCams);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return set_error(nullptr, RC_ERR_OOM, "allocating the temporal-input slots failed");
    }
    *out = ctx;
    return RC_OK;
}

int rc_destroy(rc_ctx *ctx) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    (void)hipSetDevice(ctx->device);
    if (ctx->capturing) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(ctx->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        ctx->capturing = 0;
    }
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->frameExec) (void)hipGraphExecDestroy(ctx->frameExec);
    if (ctx->frameGraph) (void)hipGraphDestroy(ctx->frameGraph);
    if (ctx->dTemporal) (void)hipFree(ctx->dTemporal);
    if (ctx->hTemporal) (void)hipHostFree(ctx->hTemporal);
    if (ctx->dCams) (void)hipFree(ctx->dCams);
    if (ctx->hCams) (void)hipHostFree(ctx->hCams);
    free_scene(ctx->scene);
    for (auto &d : ctx->extraScenes) free_scene(d);
    free_views(ctx);
    free_images(ctx->img);
    free_sppm(ctx->sppm);
    free_wavefront(ctx->wf);
    free_lane_buffers(ctx->lanes);
    free_net(ctx->net);
    dfree(ctx->counters);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return RC_OK;
}

const char *rc_last_error(const rc_ctx *ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

void *rc_stream(rc_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int rc_sync(rc_ctx *ctx) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (int rc = capture_block(ctx, "rc_sync")) return rc;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
    return RC_OK;
}

}  // extern "C"

static int upload_scene_into(rc_ctx *ctx, DeviceScene &D, const float *tris, int32_t nTris, const float *nodes,
                             int32_t nNodes, const float *mats, int32_t nMats, const float *lights, int32_t nLights,
                             float totalLightPower) {
    if (nTris <= 0 || nNodes <= 0 || nMats <= 0 || nLights < 0)
        return set_error(ctx, RC_ERR_INVALID_ARG, "nTris, nNodes and nMats must be > 0, nLights >= 0");
    if (!tris || !nodes || !mats || (nLights > 0 && !lights))
        return set_error(ctx, RC_ERR_INVALID_ARG, "NULL scene stream");
    if (nTris >= (1 << 26))
        return set_error(ctx, RC_ERR_LIMIT, "nTris must be < 2^26");
    if (int rc = capture_block(ctx, "uploading a scene")) return rc;
    (void)hipSetDevice(ctx->device);

    // --- validate the BVH: children in range, strictly after the parent, each
    //     node reached once (a tree), leaves in range.
    std::vector<int> parents(nNodes, 0);
    std::vector<int32_t> L(nNodes), R(nNodes), PS(nNodes), PC(nNodes);
    for (int i = 0; i < nNodes; ++i) {
        const float *n = nodes + 12 * (size_t)i;
        L[i] = bits_i(n[8]); R[i] = bits_i(n[9]); PS[i] = bits_i(n[10]); PC[i] = bits_i(n[11]);
        if (PC[i] > 0) {
            if (PS[i] < 0 || (int64_t)PS[i] + PC[i] > nTris)
                return set_error(ctx, RC_ERR_INVALID_ARG, "BVH leaf primitive range out of bounds");
        } else {
            if (L[i] <= i || R[i] <= i || L[i] >= nNodes || R[i] >= nNodes || L[i] == R[i])
                return set_error(ctx, RC_ERR_INVALID_ARG, "BVH child index invalid (must be > parent, < nNodes)");
            parents[L[i]]++;
            parents[R[i]]++;
        }
    }
    for (int i = 1; i < nNodes; ++i)
        if (parents[i] > 1) return set_error(ctx, RC_ERR_INVALID_ARG, "BVH node has more than one parent");
    for (int i = 0; i < nTris; ++i) {
        int32_t m = bits_i(tris[32 * (size_t)i + 30]);
        if (m < 0 || m >= nMats) return set_error(ctx, RC_ERR_INVALID_ARG, "triangle matIndex out of range");
    }

    // --- inner-node numbering (DFS order preserved) and child refs
    std::vector<int> innerIdx(nNodes, -1);
    int nInner = 0, nLeaves = 0;
    for (int i = 0; i < nNodes; ++i) {
        if (PC[i] > 0) ++nLeaves;
        else innerIdx[i] = nInner++;
    }
    std::vector<int> leafBig(nTris, 0);
    bool anyBig = false;
    auto refOf = [&](int node) -> int32_t {
        if (PC[node] > 0) {
            uint32_t cnt = PC[node] <= 31 ? (uint32_t)PC[node] : 0u;
            if (cnt == 0) { leafBig[PS[node]] = PC[node]; anyBig = true; }
            return (int32_t)~(((uint32_t)PS[node] << 5) | cnt);
        }
        return innerIdx[node];
    };
    // inner node records of 14 words (56 B, packed): both child boxes, then the refs
    std::vector<float> inner((size_t)nInner * 14);
    for (int i = 0; i < nNodes; ++i) {
        if (PC[i] > 0) continue;
        const float *l = nodes + 12 * (size_t)L[i];
        const float *r = nodes + 12 * (size_t)R[i];
        float *o = &inner[(size_t)innerIdx[i] * 14];
        const float rec[14] = {l[0], l[1], l[2], l[4], l[5], l[6], r[0], r[1], r[2], r[4], r[5], r[6],
                               bits_f(refOf(L[i])), bits_f(refOf(R[i]))};
        for (int k = 0; k < 14; ++k) o[k] = rec[k];
    }
    int32_t rootRef = refOf(0);
    // depth
    int maxDepth = 0;
    {
        std::vector<std::pair<int, int>> st;
        st.push_back({0, 0});
        while (!st.empty()) {
            auto pr = st.back();
            st.pop_back();
            if (pr.second > maxDepth) maxDepth = pr.second;
            if (PC[pr.first] <= 0) {
                st.push_back({L[pr.first], pr.second + 1});
                st.push_back({R[pr.first], pr.second + 1});
            }
        }
    }

    // triangle tests read v0, e1 = v1 - v0, e2 = v2 - v0 packed in 36 bytes: more
    // triangles per cache line than 16-byte-aligned rows would give
    std::vector<float> pos((size_t)nTris * 9);
    std::vector<float4> shade((size_t)nTris * 3);
    for (int i = 0; i < nTris; ++i) {
        const float *t = tris + 32 * (size_t)i;
        const float rec[9] = {t[0], t[1], t[2], t[4] - t[0], t[5] - t[1], t[6] - t[2],
                              t[8] - t[0], t[9] - t[1], t[10] - t[2]};
        for (int k = 0; k < 9; ++k) pos[9 * (size_t)i + k] = rec[k];
        shade[3 * i + 0] = make_float4(t[12], t[13], t[14], t[30]);
        shade[3 * i + 1] = make_float4(t[16], t[17], t[18], 0.0f);
        shade[3 * i + 2] = make_float4(t[20], t[21], t[22], 0.0f);
    }
    std::vector<float> cdf((size_t)(nLights > 0 ? nLights : 1), 1.0f);
    for (int i = 0; i < nLights; ++i) cdf[i] = lights[20 * (size_t)i + 19];
    // a non-decreasing, NaN-free CDF (as uploadLights writes it) lets the light
    // pick bisect instead of scanning: the first entry above u is the same
    int cdfSorted = 1;
    for (int i = 0; i < nLights; ++i)
        if (cdf[i] != cdf[i] || (i > 0 && !(cdf[i - 1] <= cdf[i]))) cdfSorted = 0;

    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
    ++ctx->freeGen;
    free_scene(D);
    int rc;
    if ((rc = dalloc(ctx, D.inner, inner.size() ? inner.size() : 14, "hipMalloc(nodes)"))) return rc;
    if ((rc = dalloc(ctx, D.triPos, pos.size(), "hipMalloc(triangles)"))) return rc;
    if ((rc = dalloc(ctx, D.triShade, shade.size(), "hipMalloc(shading)"))) return rc;
    if ((rc = dalloc(ctx, D.mats, (size_t)nMats * 4, "hipMalloc(materials)"))) return rc;
    if ((rc = dalloc(ctx, D.lights, (size_t)(nLights > 0 ? nLights : 1) * 5, "hipMalloc(lights)"))) return rc;
    if ((rc = dalloc(ctx, D.cdf, cdf.size(), "hipMalloc(cdf)"))) return rc;
    if ((rc = dalloc(ctx, D.leafBigCount, (size_t)nTris, "hipMalloc(leafBigCount)"))) return rc;
    if ((rc = upload(ctx, D.inner, inner.data(), inner.size() * sizeof(float), "upload(nodes)"))) return rc;
    if ((rc = upload(ctx, D.triPos, pos.data(), pos.size() * sizeof(float), "upload(triangles)"))) return rc;
    if ((rc = upload(ctx, D.triShade, shade.data(), shade.size() * sizeof(float4), "upload(shading)"))) return rc;
    if ((rc = upload(ctx, D.mats, mats, (size_t)nMats * 64, "upload(materials)"))) return rc;
    if (nLights > 0 && (rc = upload(ctx, D.lights, lights, (size_t)nLights * 80, "upload(lights)"))) return rc;
    if ((rc = upload(ctx, D.cdf, cdf.data(), cdf.size() * sizeof(float), "upload(cdf)"))) return rc;
    if ((rc = upload(ctx, D.leafBigCount, leafBig.data(), leafBig.size() * sizeof(int), "upload(leaves)"))) return rc;
    (void)anyBig;
    e = hipStreamSynchronize(ctx->stream);  // host vectors go out of scope
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
    D.nTris = nTris; D.nInner = nInner; D.nLeaves = nLeaves; D.nMats = nMats; D.nLights = nLights;
    D.totalLightPower = totalLightPower;
    for (int k = 0; k < 3; ++k) { D.rootMin[k] = nodes[k]; D.rootMax[k] = nodes[4 + k]; }
    D.rootRef = rootRef;
    D.maxDepth = maxDepth;
    D.cdfSorted = cdfSorted;
    D.valid = true;
    return RC_OK;
}

DeviceScene *rc::scene_slot(rc_ctx *ctx, int slot) {
    if (slot == 0) return &ctx->scene;
    if (slot >= 1 && slot < RC_MAX_SCENE_SLOTS) return &ctx->extraScenes[slot - 1];
    return nullptr;
}

extern "C" {

int rc_upload_scene(rc_ctx *ctx, const float *tris, int32_t nTris, const float *nodes, int32_t nNodes,
                    const float *mats, int32_t nMats, const float *lights, int32_t nLights,
                    float totalLightPower) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    return upload_scene_into(ctx, ctx->scene, tris, nTris, nodes, nNodes, mats, nMats, lights, nLights,
                             totalLightPower);
}

int rc_upload_scene_slot(rc_ctx *ctx, int32_t slot, const float *tris, int32_t nTris, const float *nodes,
                         int32_t nNodes, const float *mats, int32_t nMats, const float *lights, int32_t nLights,
                         float totalLightPower) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    DeviceScene *D = scene_slot(ctx, slot);
    if (!D) return set_error(ctx, RC_ERR_INVALID_ARG, "scene slot out of range");
    return upload_scene_into(ctx, *D, tris, nTris, nodes, nNodes, mats, nMats, lights, nLights, totalLightPower);
}

static int render_common(rc_ctx *ctx, const rc_camera *cam, const rc_pt_params *p, int32_t n) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    int rc;
    if ((rc = validate_camera(ctx, cam))) return rc;
    if (!p) return set_error(ctx, RC_ERR_INVALID_ARG, "params is NULL");
    if (p->maxBounces < 1 || p->maxBounces > 16)
        return set_error(ctx, RC_ERR_INVALID_ARG, "maxBounces must be in 1..16");
    if (n < 1) return set_error(ctx, RC_ERR_INVALID_ARG, "nSamples must be >= 1");
    if (!ctx->scene.valid) return set_error(ctx, RC_ERR_STATE, "no scene uploaded");
    if ((int64_t)ctx->totalSamples + n > 0x7fffffff / 8)
        return set_error(ctx, RC_ERR_LIMIT, "sample count overflow");
    if ((rc = capture_needs_reset(ctx))) return rc;
    if (ctx->ptMode == RC_PT_WAVEFRONT && (rc = capture_block(ctx, "the wavefront path tracer (host-driven queues)")))
        return rc;
    (void)hipSetDevice(ctx->device);
    if ((rc = ensure_images(ctx, cam->width, cam->height))) return rc;
    hipError_t e;
    const CamArgs *dcam = nullptr;
    if (ctx->ptMode != RC_PT_WAVEFRONT && (rc = camera_slot(ctx, cam_args(*cam), &dcam))) return rc;
    if (ctx->ptMode == RC_PT_WAVEFRONT) {
        e = launch_wavefront(ctx, *cam, *p, n);
    } else if (ctx->ptMode == RC_PT_LANES && lanes_supported(ctx->scene)) {
        PassArgs A = pass_args(ctx, *cam, *p, n);
        A.dcam = dcam;
        e = launch_lanes(ctx, A, ctx->laneThresh, ctx->laneVariant);
    } else {
        e = launch_pathtrace(ctx, *cam, *p, n, dcam);
    }
    if (e != hipSuccess) return hip_fail(ctx, e, "pathtrace launch");
    ctx->totalSamples += n;
    return RC_OK;
}

int rc_render_pass(rc_ctx *ctx, const rc_camera *cam, const rc_pt_params *p) {
    return render_common(ctx, cam, p, 1);
}

int rc_render_passes(rc_ctx *ctx, const rc_camera *cam, const rc_pt_params *p, int32_t n) {
    return render_common(ctx, cam, p, n);
}

int rc_reset(rc_ctx *ctx) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    ctx->totalSamples = 0;
    if (ctx->capturing) ctx->captureResetSeen = 1;
    // Renderer.reset re-creates the textures with NULL data (zeroed storage).
    Images &im = ctx->img;
    if (im.width > 0) {
        (void)hipSetDevice(ctx->device);
        size_t n = (size_t)im.width * (size_t)im.height;
        const size_t sizes[] = {n * 16, n * 8, n * 8, n * 4, n * 4, n * 4, n * 32};
        void *ptrs[] = {im.accum, im.gbuf, im.galbedo, im.variance, im.primId, im.gbufId, im.accum64};
        ZeroSegs z{};
        size_t words = 0;
        for (int i = 0; i < 7; ++i) {
            if (!ptrs[i]) continue;
            words += sizes[i] / 4;
            z.p[z.count] = (uint32_t *)ptrs[i];
            z.end[z.count++] = words;
        }
        if (z.count > 0) {
            hipLaunchKernelGGL(zero_segs_kernel, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, ctx->stream, z);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return hip_fail(ctx, e, "reset fill");
        }
    }
    return RC_OK;
}

int rc_debug_counters(rc_ctx *ctx, int32_t enable, uint64_t *out, int32_t n) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (int rc = capture_block(ctx, "rc_debug_counters")) return rc;
    if (n < 0 || n > 64 || (n > 0 && !out)) return set_error(ctx, RC_ERR_INVALID_ARG, "bad counter buffer");
    (void)hipSetDevice(ctx->device);
    hipError_t e = hipSuccess;
    if (n > 0) {
        e = hipMemcpyAsync(out, ctx->counters, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return hip_fail(ctx, e, "rc_debug_counters(read)");
    }
    e = hipMemsetAsync(ctx->counters, 0, 64 * sizeof(unsigned long long), ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "rc_debug_counters(clear)");
    ctx->countersEnabled = enable ? 1 : 0;
    return RC_OK;
}

int32_t rc_total_samples(const rc_ctx *ctx) { return ctx ? ctx->totalSamples : -1; }

int rc_get_tuning(const rc_ctx *ctx, rc_tuning *out) {
    if (!ctx || !out) return RC_ERR_INVALID_ARG;
    out->laneThreshold = ctx->laneThresh;
    out->laneVariant = ctx->laneVariant;
    out->laneBufferMiB = ctx->laneBufferMiB;
    out->laneTileWidth = 1 << ctx->laneTileShift;
    out->ptMode = ctx->ptMode;
    return RC_OK;
}

int rc_set_tuning(rc_ctx *ctx, const rc_tuning *t) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (!t) return set_error(ctx, RC_ERR_INVALID_ARG, "tuning is NULL");
    if (t->laneThreshold < 1 || t->laneThreshold > 64)
        return set_error(ctx, RC_ERR_INVALID_ARG, "laneThreshold must be in 1..64");
    if (t->laneVariant < 0 || t->laneVariant >= RC_LANE_VARIANTS)
        return set_error(ctx, RC_ERR_INVALID_ARG, "laneVariant out of range");
    if (t->laneBufferMiB < 64 || t->laneBufferMiB > 65536)
        return set_error(ctx, RC_ERR_INVALID_ARG, "laneBufferMiB must be in 64..65536");
    int shift = -1;
    for (int k = 2; k <= 5; ++k)
        if (t->laneTileWidth == (1 << k)) shift = k;
    if (shift < 0) return set_error(ctx, RC_ERR_INVALID_ARG, "laneTileWidth must be 4, 8, 16 or 32");
    if (t->ptMode != RC_PT_MEGAKERNEL && t->ptMode != RC_PT_WAVEFRONT && t->ptMode != RC_PT_LANES)
        return set_error(ctx, RC_ERR_INVALID_ARG, "ptMode must be RC_PT_LANES, RC_PT_MEGAKERNEL or RC_PT_WAVEFRONT");
    ctx->laneTileShift = shift;
    ctx->ptMode = t->ptMode;
    ctx->laneThresh = t->laneThreshold;
    ctx->laneVariant = t->laneVariant;
    ctx->laneBufferMiB = t->laneBufferMiB;
    return RC_OK;
}

int rc_set_pt_mode(rc_ctx *ctx, int32_t mode) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (mode != RC_PT_MEGAKERNEL && mode != RC_PT_WAVEFRONT && mode != RC_PT_LANES)
        return set_error(ctx, RC_ERR_INVALID_ARG, "mode must be RC_PT_LANES, RC_PT_MEGAKERNEL or RC_PT_WAVEFRONT");
    ctx->ptMode = mode;
    return RC_OK;
}

int rc_set_accum_mode(rc_ctx *ctx, int32_t mode) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (int rc = capture_block(ctx, "rc_set_accum_mode")) return rc;
    if (mode != RC_ACCUM_FP32 && mode != RC_ACCUM_FP64)
        return set_error(ctx, RC_ERR_INVALID_ARG, "accumulation mode must be RC_ACCUM_FP32 or RC_ACCUM_FP64");
    ctx->accumMode = mode;
    ctx->totalSamples = 0;
    if (ctx->img.width > 0) {
        int w = ctx->img.width, h = ctx->img.height;
        if (mode == RC_ACCUM_FP64 && !ctx->img.accum64) {
            (void)hipSetDevice(ctx->device);
            ctx->img.width = 0;  // force reallocation with the FP64 buffer
            return ensure_images(ctx, w, h);
        }
    }
    return RC_OK;
}

size_t rc_buffer_bytes(const rc_ctx *ctx, int32_t which) {
    if (!ctx) return 0;
    size_t n = (size_t)ctx->img.width * (size_t)ctx->img.height;
    switch (which) {
        case RC_BUF_ACCUM: return n * 16;
        case RC_BUF_GBUFFER: return n * 8;
        case RC_BUF_GALBEDO: return n * 8;
        case RC_BUF_VARIANCE: return n * 4;
        case RC_BUF_PRIMARY_ID: return n * 4;
        case RC_BUF_GBUFFER_ID: return n * 4;
        case RC_BUF_ACCUM64: return ctx->img.accum64 ? n * 32 : 0;
        case RC_BUF_POST_HDR:
            return ctx->img.postOut ? (size_t)ctx->img.postW * (size_t)ctx->img.postH * 16 : 0;
        case RC_BUF_DISPLAY:
            return ctx->img.postOut ? (size_t)ctx->img.postW * (size_t)ctx->img.postH * 4 : 0;
        default: return 0;
    }
}

void rc_post_defaults(rc_post_params *p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    // RenderSettings.java:26-191 defaults
    p->exposure = 1.0f;
    p->atrousEnabled = 0;
    p->atrousIterations = 4;
    p->atrousMaxSpp = 0;
    p->atrousSigmaColor = 1.0f;
    p->atrousSigmaNormal = 64.0f;
    p->atrousSigmaDepth = 0.15f;
    p->lumaBlurSigma = 1.5f;
    p->chromaResSigma = 7.0f;
    p->colorTempBias = 0.0f;
    p->colorHueRot = 0.0f;
    p->colorSaturation = 0.95f;
    p->colorDriftSpeed = 0.05f;
    p->drBlackLift = 0.01f;
    p->drShadowCrush = 0.03f;
    p->drHighlightKnee = 0.78f;
    p->drHighlightClip = 0.95f;
    p->ccdNoiseLumaBase = 0.015f;
    p->ccdNoiseChromaScale = 1.6f;
    p->ccdSmearThreshold = 0.85f;
    p->ccdSmearIntensity = 0.009f;
    p->ccdSmearLength = 20;
    p->chromaBleedFactor = 0.35f;
    p->timebaseAmplitudePx = 1.5f;
    p->timebaseFreq = 0.02f;
    p->timebaseSpeed = 0.70f;
    p->dotCrawlIntensity = 0.25f;
    p->dotCrawlSubcarrier = 0.14f;
    p->dotCrawlEdgeThresh = 0.04f;
    p->tapeAge = 0.3f;
    p->dropoutBright = 0.92f;
    p->dropoutBurstRate = 0.45f;
    p->headSwitchLines = 12;
    p->headSwitchJitter = 4.0f;
    p->headSwitchRipple = 0.10f;
    p->headSwitchChroma = 0.35f;
    p->trackingSeverity = 0.1f;
    p->trackingMaxDisplacePx = 28.0f;
    p->trackingFringeWidthPx = 4.0f;
    p->edgeEnhanceAmount = 1.5f;
    p->edgeEnhanceCoreThreshold = 0.025f;
    p->interlaceFieldOffsetPx = 0.5f;
    p->interlaceCombStrength = 1.5f;
    p->interlaceCombEdgeThresh = 0.04f;
    p->interlaceLineWeighting = 0.05f;
}

int rc_post(rc_ctx *ctx, const rc_post_params *p, const rc_temporal *t) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    int rc;
    if ((rc = validate_post(ctx, p, t))) return rc;
    if (ctx->img.width <= 0 || ctx->totalSamples <= 0)
        return set_error(ctx, RC_ERR_STATE, "nothing accumulated: render before rc_post");
    if ((rc = capture_needs_reset(ctx))) return rc;
    (void)hipSetDevice(ctx->device);
    if ((rc = ensure_post(ctx, ctx->img.width, ctx->img.height))) return rc;
    return run_post_chain(ctx, ctx->img.accum, ctx->img.width, ctx->img.height, *p, *t, ctx->totalSamples, true);
}

int rc_set_post_mode(rc_ctx *ctx, int32_t mode) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (mode != RC_POST_FUSED && mode != RC_POST_PER_PASS) return set_error(ctx, RC_ERR_INVALID_ARG, "bad post mode");
    if (int rc = capture_block(ctx, "rc_set_post_mode")) return rc;
    ctx->postMode = mode;
    return RC_OK;
}

int rc_post_launches(const rc_ctx *ctx, int32_t *launches) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (launches) *launches = ctx->postLaunches;
    return RC_OK;
}

int rc_post_image(rc_ctx *ctx, const float *rgba, int32_t width, int32_t height, const rc_post_params *p,
                  const rc_temporal *t) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    int rc;
    if ((rc = validate_post(ctx, p, t))) return rc;
    if (!rgba || width <= 0 || height <= 0 || width > 16384 || height > 16384)
        return set_error(ctx, RC_ERR_INVALID_ARG, "image pointer/size invalid");
    if ((rc = capture_block(ctx, "rc_post_image"))) return rc;
    (void)hipSetDevice(ctx->device);
    if ((rc = ensure_post(ctx, width, height))) return rc;
    hipError_t e = hipMemcpyAsync(ctx->img.postIn, rgba, (size_t)width * (size_t)height * 16, hipMemcpyHostToDevice,
                                  ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "upload(image)");
    // PostProcessStack.runOnImage (PostProcessStack.java:130-136): p00 with N = 1 and
    // exposure 1, no a-trous (no G-buffer).
    rc_post_params q = *p;
    q.exposure = 1.0f;
    rc = run_post_chain(ctx, ctx->img.postIn, width, height, q, *t, 1, false);
    if (rc) return rc;
    e = hipStreamSynchronize(ctx->stream);  // the host image may be freed on return
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
    return RC_OK;
}

int rc_read(rc_ctx *ctx, int32_t which, void *dst, size_t bytes) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (int rc = capture_block(ctx, "rc_read")) return rc;
    if (!dst) return set_error(ctx, RC_ERR_INVALID_ARG, "dst is NULL");
    size_t need = rc_buffer_bytes(ctx, which);
    if (need == 0) return set_error(ctx, RC_ERR_STATE, "buffer not available (bad id, no render yet, or FP64 off)");
    if (bytes != need) return set_error(ctx, RC_ERR_INVALID_ARG, "bytes does not match the buffer size");
    const void *src = nullptr;
    switch (which) {
        case RC_BUF_ACCUM: src = ctx->img.accum; break;
        case RC_BUF_GBUFFER: src = ctx->img.gbuf; break;
        case RC_BUF_GALBEDO: src = ctx->img.galbedo; break;
        case RC_BUF_VARIANCE: src = ctx->img.variance; break;
        case RC_BUF_PRIMARY_ID: src = ctx->img.primId; break;
        case RC_BUF_GBUFFER_ID: src = ctx->img.gbufId; break;
        case RC_BUF_ACCUM64: src = ctx->img.accum64; break;
        case RC_BUF_POST_HDR: src = ctx->img.postOut; break;
        case RC_BUF_DISPLAY: src = ctx->img.display; break;
    }
    (void)hipSetDevice(ctx->device);
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "rc_read");
    return RC_OK;
}

// ---- captured frames ----------------------------------------------------

static void release_frame(rc_ctx *ctx) {
    if (ctx->frameExec || ctx->frameGraph) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->frameExec) (void)hipGraphExecDestroy(ctx->frameExec);
    if (ctx->frameGraph) (void)hipGraphDestroy(ctx->frameGraph);
    ctx->frameExec = nullptr;
    ctx->frameGraph = nullptr;
    ctx->graphUsesViews = 0;
}

int rc_frame_capture_begin(rc_ctx *ctx) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (ctx->capturing) return set_error(ctx, RC_ERR_STATE, "a frame is already being captured");
    (void)hipSetDevice(ctx->device);
    release_frame(ctx);
    ctx->nTemporal = 0;
    ctx->nCams = 0;
    ctx->replayCamSet = 0;
    ctx->slotsExhausted = 0;
    ctx->captureResetSeen = 0;
    ctx->captureViewsResetSeen = 0;
    hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamBeginCapture");
    ctx->capturing = 1;
    return RC_OK;
}

int rc_frame_capture_end(rc_ctx *ctx) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (!ctx->capturing) return set_error(ctx, RC_ERR_STATE, "no frame is being captured");
    (void)hipSetDevice(ctx->device);
    ctx->capturing = 0;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        ctx->graphUsesViews = 0;
        return hip_fail(ctx, e != hipSuccess ? e : hipErrorUnknown, "hipStreamEndCapture");
    }
    e = hipGraphInstantiate(&ctx->frameExec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        ctx->frameExec = nullptr;
        (void)hipGraphDestroy(g);
        ctx->graphUsesViews = 0;
        return hip_fail(ctx, e, "hipGraphInstantiate");
    }
    ctx->frameGraph = g;
    ctx->graphFreeGen = ctx->freeGen;
    ctx->captured.totalSamples = ctx->totalSamples;
    ctx->captured.radius = ctx->sppm.radius;
    ctx->captured.radius64 = ctx->sppm.radius64;
    ctx->captured.iteration = ctx->sppm.iteration;
    ctx->captured.photonCount = ctx->sppm.photonCount;
    ctx->captured.hashSize = ctx->sppm.hashSize;
    ctx->captured.viewTotals = views_totals(ctx);
    return RC_OK;
}

int rc_frame_replay(rc_ctx *ctx, const rc_temporal *t) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (ctx->capturing) return set_error(ctx, RC_ERR_STATE, "rc_frame_replay while capturing");
    if (!ctx->frameExec) return set_error(ctx, RC_ERR_STATE, "no captured frame");
    if (ctx->graphFreeGen != ctx->freeGen)
        return set_error(ctx, RC_ERR_STATE,
                         "buffers the captured frame uses were reallocated since (new resolution, scene upload, larger "
                         "workspace): capture it again");
    if (ctx->graphUsesViews && ctx->graphViewTablesSig != ctx->viewTablesSig)
        return set_error(ctx, RC_ERR_STATE,
                         "the view tables changed since the frame was captured (other views were rendered): capture "
                         "it again");
    if (t && ctx->slotsExhausted)
        return set_error(ctx, RC_ERR_LIMIT, "the captured frame has more post chains than temporal slots");
    (void)hipSetDevice(ctx->device);
    hipError_t e = hipSuccess;
    if (ctx->nTemporal > 0) {
        if (t) {
            hipLaunchKernelGGL(fill_temporal_kernel, dim3((ctx->nTemporal + 255) / 256), dim3(256), 0, ctx->stream,
                               ctx->dTemporal, ctx->nTemporal, *t);
            e = hipGetLastError();
        } else {
            e = hipMemcpyAsync(ctx->dTemporal, ctx->hTemporal, sizeof(rc_temporal) * ctx->nTemporal,
                               hipMemcpyHostToDevice, ctx->stream);
        }
        if (e != hipSuccess) return hip_fail(ctx, e, "temporal inputs of the captured frame");
    }
    if (ctx->nCams > 0) {
        if (ctx->replayCamSet) {
            hipLaunchKernelGGL(fill_cam_kernel, dim3((ctx->nCams + 255) / 256), dim3(256), 0, ctx->stream,
                               ctx->dCams, ctx->nCams, ctx->replayCam);
            e = hipGetLastError();
        } else {
            e = hipMemcpyAsync(ctx->dCams, ctx->hCams, sizeof(CamArgs) * ctx->nCams, hipMemcpyHostToDevice,
                               ctx->stream);
        }
        if (e != hipSuccess) return hip_fail(ctx, e, "camera of the captured frame");
    }
    e = hipGraphLaunch(ctx->frameExec, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipGraphLaunch");
    ctx->totalSamples = ctx->captured.totalSamples;
    ctx->sppm.radius = ctx->captured.radius;
    ctx->sppm.radius64 = ctx->captured.radius64;
    ctx->sppm.iteration = ctx->captured.iteration;
    ctx->sppm.photonCount = ctx->captured.photonCount;
    ctx->sppm.hashSize = ctx->captured.hashSize;
    views_set_totals(ctx, ctx->captured.viewTotals);
    return RC_OK;
}

int rc_frame_set_camera(rc_ctx *ctx, const rc_camera *cam) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (ctx->capturing) return set_error(ctx, RC_ERR_STATE, "rc_frame_set_camera while capturing");
    if (!ctx->frameExec) return set_error(ctx, RC_ERR_STATE, "no captured frame");
    if (!cam) {
        ctx->replayCamSet = 0;
        return RC_OK;
    }
    int rc;
    if ((rc = validate_camera(ctx, cam))) return rc;
    for (int i = 0; i < ctx->nCams; ++i)
        if (ctx->hCams[i].width != cam->width || ctx->hCams[i].height != cam->height)
            return set_error(ctx, RC_ERR_INVALID_ARG, "rc_frame_set_camera: the recorded frame has another resolution");
    ctx->replayCam = cam_args(*cam);
    ctx->replayCamSet = 1;
    return RC_OK;
}

int rc_frame_release(rc_ctx *ctx) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (ctx->capturing) return set_error(ctx, RC_ERR_STATE, "rc_frame_release while capturing");
    (void)hipSetDevice(ctx->device);
    release_frame(ctx);
    return RC_OK;
}

}  // extern "C"
