// rc_batch.hip — batched views: many render requests (each with its own scene
// slot, camera, resolution and sample count) traced in ONE persistent launch
// of the path tracer, then post-processed per view. Project-new (the
// reference renders one view per Renderer); every view's pixels are computed
// exactly as a single-view rc_render_passes call would compute them.
//
// Layout: one arena per image kind, views packed one after another (pixel
// offset pixOff[v]); a device table of per-view launch arguments sorted by
// work (samples x tiles, heaviest first) and the prefix sum of their 16x16
// tile counts, which the persistent kernel walks with an atomic tile counter.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "rc_internal.h"
#include "rc_shade.h"

namespace rc {

struct Views {
    int n = 0;
    std::vector<int> w, h, total;
    std::vector<size_t> pixOff;
    size_t pixCap = 0;
    bool fp64 = false;
    int viewCap = 0;
    float4 *accum = nullptr;
    double4 *accum64 = nullptr;
    ushort4 *gbuf = nullptr, *galbedo = nullptr;
    float *variance = nullptr;
    int *primId = nullptr, *gbufId = nullptr;
    float4 *hdrA = nullptr, *hdrB = nullptr;
    uchar4 *display = nullptr;
    std::vector<const float4 *> postOut;
    PassArgs *dArgs = nullptr;
    int *dTileOff = nullptr;
    int *dCounter = nullptr;
    PassArgs *hArgs = nullptr;  // pinned staging of the table
    int *hTileOff = nullptr;
    // lane-scheduler tables (rc_pathtrace_lanes.hip): chunk x view entries
    int chunkCap = 0;
    LaneChunk *dChunks = nullptr, *hChunks = nullptr;
    long long *dPixOff = nullptr, *hPixOff = nullptr;
    int uploadedN = -1;              // table sizes of the last upload
    long long uploadedEntries = -1;
};

namespace {

template <typename T>
void vfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

void free_arenas(Views &v) {
    vfree(v.accum); vfree(v.accum64); vfree(v.gbuf); vfree(v.galbedo); vfree(v.variance);
    vfree(v.primId); vfree(v.gbufId); vfree(v.hdrA); vfree(v.hdrB); vfree(v.display);
    v.pixCap = 0;
}

void free_lane_tables(Views &v) {
    vfree(v.dChunks); vfree(v.dPixOff);
    if (v.hChunks) (void)hipHostFree(v.hChunks);
    if (v.hPixOff) (void)hipHostFree(v.hPixOff);
    v.hChunks = nullptr;
    v.hPixOff = nullptr;
    v.chunkCap = 0;
}

void free_tables(Views &v) {
    free_lane_tables(v);
    v.uploadedN = -1;
    v.uploadedEntries = -1;
    vfree(v.dArgs); vfree(v.dTileOff); vfree(v.dCounter);
    if (v.hArgs) (void)hipHostFree(v.hArgs);
    if (v.hTileOff) (void)hipHostFree(v.hTileOff);
    v.hArgs = nullptr;
    v.hTileOff = nullptr;
    v.viewCap = 0;
}

template <typename T>
hipError_t valloc(T *&p, size_t count) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, (count ? count : 1) * sizeof(T));
    p = (T *)q;
    return e;
}

int validate_view(rc_ctx *ctx, const rc_view &v) {
    DeviceScene *S = scene_slot(ctx, v.scene);
    if (!S || !S->valid) return set_error(ctx, RC_ERR_STATE, "view names an empty or invalid scene slot");
    const rc_camera &c = v.cam;
    if (c.width <= 0 || c.height <= 0 || c.width > 16384 || c.height > 16384)
        return set_error(ctx, RC_ERR_INVALID_ARG, "view width/height must be in 1..16384");
    if (c.lensRadius > 0.0f && c.apertureBlades < 1)
        return set_error(ctx, RC_ERR_INVALID_ARG, "apertureBlades must be >= 1 when lensRadius > 0");
    if (v.params.maxBounces < 1 || v.params.maxBounces > 16)
        return set_error(ctx, RC_ERR_INVALID_ARG, "maxBounces must be in 1..16");
    if (v.nSamples < 1 || v.nSamples > (1 << 20)) return set_error(ctx, RC_ERR_INVALID_ARG, "nSamples in 1..2^20");
    return RC_OK;
}

// (Re)lays out the arenas for these view sizes; returns true when the layout changed.
int layout(rc_ctx *ctx, Views &V, const rc_view *views, int n, bool fp64, bool &changed) {
    changed = V.n != n || V.fp64 != fp64;
    for (int i = 0; i < n && !changed; ++i) changed = V.w[i] != views[i].cam.width || V.h[i] != views[i].cam.height;
    if (!changed) return RC_OK;
    if (int rc = capture_block(ctx, "a new set of view sizes")) return rc;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
    V.w.resize(n); V.h.resize(n); V.total.assign(n, 0); V.pixOff.resize(n); V.postOut.assign(n, nullptr);
    size_t px = 0;
    for (int i = 0; i < n; ++i) {
        V.w[i] = views[i].cam.width;
        V.h[i] = views[i].cam.height;
        V.pixOff[i] = px;
        px += (size_t)V.w[i] * V.h[i];
    }
    if (px > V.pixCap || (fp64 && !V.accum64)) {
        ++ctx->freeGen;
        free_arenas(V);
        if (e == hipSuccess) e = valloc(V.accum, px);
        if (e == hipSuccess && fp64) e = valloc(V.accum64, px);
        if (e == hipSuccess) e = valloc(V.gbuf, px);
        if (e == hipSuccess) e = valloc(V.galbedo, px);
        if (e == hipSuccess) e = valloc(V.variance, px);
        if (e == hipSuccess) e = valloc(V.primId, px);
        if (e == hipSuccess) e = valloc(V.gbufId, px);
        if (e == hipSuccess) e = valloc(V.hdrA, px);
        if (e == hipSuccess) e = valloc(V.hdrB, px);
        if (e == hipSuccess) e = valloc(V.display, px);
        if (e != hipSuccess) {
            free_arenas(V);
            V.n = 0;
            return hip_fail(ctx, e, "hipMalloc(view images)");
        }
        V.pixCap = px;
    }
    if (n > V.viewCap) {
        ++ctx->freeGen;
        free_tables(V);
        if (e == hipSuccess) e = valloc(V.dArgs, n);
        if (e == hipSuccess) e = valloc(V.dTileOff, n + 1);
        if (e == hipSuccess) e = valloc(V.dCounter, 1);
        if (e == hipSuccess) e = hipHostMalloc((void **)&V.hArgs, sizeof(PassArgs) * n, hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void **)&V.hTileOff, sizeof(int) * (n + 1), hipHostMallocDefault);
        if (e != hipSuccess) {
            free_tables(V);
            V.n = 0;
            return hip_fail(ctx, e, "view tables");
        }
        V.viewCap = n;
    }
    V.n = n;
    V.fp64 = fp64;
    return RC_OK;
}

// Host tables of one batch: the per-view launch arguments (sorted heaviest
// first), the tile prefix sums of the megakernel form and, for the lane
// scheduler, the chunk x view table and the pixel prefix sums.
struct ViewTables {
    std::vector<PassArgs> args;
    std::vector<int> tileOff;
    bool lanes = false;
    int nChunks = 0;
    std::vector<LaneChunk> chunks;
    std::vector<long long> pixOff;
    std::vector<long long> units;  // per chunk
    long long sampleBytes = 0;
};

// Lane-scheduler chunk plan: the largest chunk K (samples per view and launch)
// whose sample buffer fits rc_ctx::laneBufferMiB.
int plan_lanes(rc_ctx *ctx, ViewTables &T) {
    const int n = (int)T.args.size();
    std::vector<long long> tilePx(n);
    int maxS = 0;
    for (int k = 0; k < n; ++k) {
        const CamArgs &c = T.args[k].cam;
        tilePx[k] = LaneTile(c.width, c.height, T.args[k].tileShift).pixels();
        maxS = std::max(maxS, T.args[k].nSamples);
    }
    auto bytesFor = [&](long long K) {
        long long b = 0;
        for (int k = 0; k < n; ++k) b += std::min<long long>(K, T.args[k].nSamples) * tilePx[k] * 16;
        return b;
    };
    const long long cap = (long long)ctx->laneBufferMiB << 20;
    long long K = maxS;
    while (K > 1 && bytesFor(K) > cap) K = (K + 1) / 2;
    for (int k = 0; k < n; ++k)
        while (K > 1 && tilePx[k] * (K + 1) > (1ll << 30)) K = (K + 1) / 2;  // + the G-buffer rays
    for (int k = 0; k < n; ++k)
        if (tilePx[k] * (K + 1) > (1ll << 30)) return set_error(ctx, RC_ERR_LIMIT, "views: view too large");
    T.nChunks = (int)((maxS + K - 1) / K);
    const long long entries = (long long)T.nChunks * n;
    if (entries > (1 << 24)) return set_error(ctx, RC_ERR_LIMIT, "views: too many chunks");
    T.sampleBytes = bytesFor(K);
    hipError_t e = ensure_lane_buffers(ctx, (size_t)T.sampleBytes, (int)entries);
    if (e != hipSuccess) return hip_fail(ctx, e, "lane buffers");
    T.chunks.assign((size_t)entries, LaneChunk{});
    T.pixOff.assign(n + 1, 0);
    T.units.assign(T.nChunks, 0);
    long long px = 0, off = 0;
    for (int k = 0; k < n; ++k) {
        T.pixOff[k] = px;
        px += (long long)T.args[k].cam.width * T.args[k].cam.height;
        for (int c = 0; c < T.nChunks; ++c) {
            LaneChunk &L = T.chunks[(size_t)c * n + k];
            const long long s0 = (long long)c * K;
            const long long ns = std::max(0ll, std::min<long long>(K, T.args[k].nSamples - s0));
            L.samples = ctx->lanes.samples + off;
            L.counter = ctx->lanes.counters + (size_t)c * n + k;
            L.s0 = (int)s0;
            L.nSamples = (int)ns;
            L.nTilePx = (int)tilePx[k];
            T.units[c] += ns * tilePx[k];
            if (ns > 0 && T.args[k].frame0 + s0 == 0) T.units[c] += tilePx[k];  // G-buffer rays
        }
        off += std::min<long long>(K, T.args[k].nSamples) * tilePx[k];
    }
    T.pixOff[n] = px;
    return RC_OK;
}

// Uploads the tables (eager calls), or, while a frame is being captured,
// checks that they equal the tables already on the device: a recorded frame
// may not copy host memory, so it reuses the last upload.
int upload_tables(rc_ctx *ctx, Views &V, const ViewTables &T) {
    const int n = (int)T.args.size();
    const size_t entries = T.chunks.size();
    if (ctx->capturing) {
        bool same = V.uploadedN == n && V.uploadedEntries == (long long)entries &&
                    std::memcmp(V.hArgs, T.args.data(), sizeof(PassArgs) * n) == 0 &&
                    std::memcmp(V.hTileOff, T.tileOff.data(), sizeof(int) * (n + 1)) == 0;
        if (same && T.lanes)
            same = V.hChunks && V.hPixOff &&
                   std::memcmp(V.hChunks, T.chunks.data(), sizeof(LaneChunk) * entries) == 0 &&
                   std::memcmp(V.hPixOff, T.pixOff.data(), sizeof(long long) * (n + 1)) == 0;
        if (!same)
            return set_error(ctx, RC_ERR_STATE,
                             "captured rc_views_render: these views differ from the last rendered ones; render them "
                             "once before capturing");
        ctx->graphUsesViews = 1;
        ctx->graphViewTablesSig = ctx->viewTablesSig;
        return RC_OK;
    }
    // The pinned staging may still be read by a previous launch's copy.
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
    if (T.lanes && ((long long)entries > V.chunkCap || !V.dPixOff)) {
        ++ctx->freeGen;
        free_lane_tables(V);
        e = valloc(V.dChunks, entries);
        if (e == hipSuccess) e = valloc(V.dPixOff, (size_t)V.viewCap + 1);
        if (e == hipSuccess) e = hipHostMalloc((void **)&V.hChunks, sizeof(LaneChunk) * entries, hipHostMallocDefault);
        if (e == hipSuccess)
            e = hipHostMalloc((void **)&V.hPixOff, sizeof(long long) * (V.viewCap + 1), hipHostMallocDefault);
        if (e != hipSuccess) {
            free_lane_tables(V);
            return hip_fail(ctx, e, "view lane tables");
        }
        V.chunkCap = (int)entries;
    }
    std::memcpy(V.hArgs, T.args.data(), sizeof(PassArgs) * n);
    std::memcpy(V.hTileOff, T.tileOff.data(), sizeof(int) * (n + 1));
    e = hipMemcpyAsync(V.dArgs, V.hArgs, sizeof(PassArgs) * n, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(V.dTileOff, V.hTileOff, sizeof(int) * (n + 1), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && T.lanes) {
        std::memcpy(V.hChunks, T.chunks.data(), sizeof(LaneChunk) * entries);
        std::memcpy(V.hPixOff, T.pixOff.data(), sizeof(long long) * (n + 1));
        e = hipMemcpyAsync(V.dChunks, V.hChunks, sizeof(LaneChunk) * entries, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(V.dPixOff, V.hPixOff, sizeof(long long) * (n + 1), hipMemcpyHostToDevice,
                               ctx->stream);
    }
    if (e != hipSuccess) return hip_fail(ctx, e, "views table upload");
    V.uploadedN = n;
    V.uploadedEntries = (long long)entries;
    uint64_t sig = 1469598103934665603ull;  // FNV-1a over the uploaded bytes
    auto mix = [&sig](const void *p, size_t bytes) {
        const unsigned char *c = (const unsigned char *)p;
        for (size_t i = 0; i < bytes; ++i) sig = (sig ^ c[i]) * 1099511628211ull;
    };
    mix(T.args.data(), sizeof(PassArgs) * n);
    mix(T.tileOff.data(), sizeof(int) * (n + 1));
    mix(T.chunks.data(), sizeof(LaneChunk) * entries);
    mix(T.pixOff.data(), sizeof(long long) * T.pixOff.size());
    ctx->viewTablesSig = sig;
    return RC_OK;
}

Views *get_views(rc_ctx *ctx) {
    if (!ctx->views) ctx->views = new (std::nothrow) Views();
    return ctx->views;
}

}  // namespace

std::vector<int> views_totals(const rc_ctx *ctx) {
    return ctx->views ? ctx->views->total : std::vector<int>();
}

void views_set_totals(rc_ctx *ctx, const std::vector<int> &totals) {
    if (ctx->views && ctx->views->total.size() == totals.size()) ctx->views->total = totals;
}

void free_views(rc_ctx *ctx) {
    if (!ctx->views) return;
    free_arenas(*ctx->views);
    free_tables(*ctx->views);
    delete ctx->views;
    ctx->views = nullptr;
}

}  // namespace rc

using namespace rc;

extern "C" {

int rc_views_render(rc_ctx *ctx, const rc_view *views, int32_t n, int32_t flags) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (!views || n < 1 || n > 4096) return set_error(ctx, RC_ERR_INVALID_ARG, "views: 1..4096 views required");
    int rc;
    for (int i = 0; i < n; ++i)
        if ((rc = validate_view(ctx, views[i]))) return rc;
    if (ctx->capturing) {
        if (flags & RC_VIEWS_RESET) ctx->captureViewsResetSeen = 1;
        if (!ctx->captureViewsResetSeen)
            return set_error(ctx, RC_ERR_STATE,
                             "a captured frame's first rc_views_render must reset the views (RC_VIEWS_RESET)");
    }
    (void)hipSetDevice(ctx->device);
    Views *V = get_views(ctx);
    if (!V) return set_error(ctx, RC_ERR_OOM, "host allocation failed");
    const bool fp64 = (flags & RC_VIEWS_FP64) != 0;
    bool changed = false;
    if ((rc = layout(ctx, *V, views, n, fp64, changed))) return rc;
    std::vector<int> total = V->total;
    if (flags & RC_VIEWS_RESET) total.assign(n, 0);
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    auto tiles = [&](int i) { return ((V->w[i] + 15) / 16) * ((V->h[i] + 15) / 16); };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return (long long)views[a].nSamples * tiles(a) > (long long)views[b].nSamples * tiles(b);
    });
    ViewTables T;
    T.args.assign(n, PassArgs{});
    T.tileOff.assign(n + 1, 0);
    long long tot = 0;
    for (int k = 0; k < n; ++k) {
        const int i = order[k];
        const rc_view &v = views[i];
        PassArgs &A = T.args[k];
        A.S = scene_view(*scene_slot(ctx, v.scene));
        A.cam = cam_args(v.cam);
        A.dcam = &V->dArgs[k].cam;
        A.nee = v.params.neeEnabled;
        A.neeClamp = v.params.neeFireflyClamp;
        A.maxBounces = v.params.maxBounces;
        A.frame0 = total[i];
        A.nSamples = v.nSamples;
        const size_t o = V->pixOff[i];
        A.accum = V->accum + o;
        A.accum64 = V->accum64 ? V->accum64 + o : nullptr;
        A.gbuf = V->gbuf + o;
        A.galbedo = V->galbedo + o;
        A.variance = V->variance + o;
        A.primId = V->primId + o;
        A.gbufId = V->gbufId + o;
        A.counters = ctx->counters;
        A.tileShift = ctx->laneTileShift;
        T.tileOff[k] = (int)tot;
        tot += tiles(i);
    }
    if (tot >= (1ll << 30)) return set_error(ctx, RC_ERR_LIMIT, "views: too many tiles");
    T.tileOff[n] = (int)tot;
    T.lanes = ctx->ptMode == RC_PT_LANES;
    for (int i = 0; i < n && T.lanes; ++i) T.lanes = lanes_supported(*scene_slot(ctx, views[i].scene));
    if (T.lanes && (rc = plan_lanes(ctx, T))) return rc;
    if ((rc = upload_tables(ctx, *V, T))) return rc;
    hipError_t e = hipSuccess;
    if (T.lanes) {
        for (int c = 0; c < T.nChunks && e == hipSuccess; ++c)
            e = launch_lanes_views(ctx, V->dArgs, V->dChunks + (size_t)c * n, V->dPixOff, n, T.pixOff[n], T.units[c],
                                   ctx->laneThresh, ctx->laneVariant, fp64, c == 0 ? (int)T.chunks.size() : 0);
    } else {
        e = launch_pathtrace_views(ctx, V->dArgs, V->dTileOff, n, V->dCounter, (int)tot, fp64);
    }
    if (e != hipSuccess) return hip_fail(ctx, e, "views launch");
    for (int i = 0; i < n; ++i) {
        V->total[i] = total[i] + views[i].nSamples;
        V->postOut[i] = nullptr;
    }
    return RC_OK;
}

int rc_views_post(rc_ctx *ctx, int32_t view, const rc_post_params *p, const rc_temporal *t) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    Views *V = ctx->views;
    if (!V || view < 0 || view >= V->n || V->total[view] <= 0)
        return set_error(ctx, RC_ERR_STATE, "view index out of range or nothing rendered");
    if (!p || !t) return set_error(ctx, RC_ERR_INVALID_ARG, "post params or temporal state is NULL");
    if (p->ccdSmearLength < 0 || p->headSwitchLines < 0)
        return set_error(ctx, RC_ERR_INVALID_ARG, "ccdSmearLength and headSwitchLines must be >= 0");
    (void)hipSetDevice(ctx->device);
    const size_t o = V->pixOff[view];
    ChainIO io;
    io.src = V->accum + o;
    io.gbuf = V->gbuf + o;
    io.galbedo = V->galbedo + o;
    io.variance = V->variance + o;
    io.w = V->w[view];
    io.h = V->h[view];
    io.total = V->total[view];
    io.a = V->hdrA + o;
    io.b = V->hdrB + o;
    io.display = V->display + o;
    int rc = run_chain(ctx, io, *p, *t, true);
    if (rc == RC_OK) V->postOut[view] = io.out;
    return rc;
}

size_t rc_views_buffer_bytes(const rc_ctx *ctx, int32_t view, int32_t which) {
    if (!ctx || !ctx->views || view < 0 || view >= ctx->views->n) return 0;
    const size_t px = (size_t)ctx->views->w[view] * ctx->views->h[view];
    switch (which) {
        case RC_BUF_ACCUM: return px * 16;
        case RC_BUF_GBUFFER: return px * 8;
        case RC_BUF_GALBEDO: return px * 8;
        case RC_BUF_VARIANCE: return px * 4;
        case RC_BUF_PRIMARY_ID: return px * 4;
        case RC_BUF_GBUFFER_ID: return px * 4;
        case RC_BUF_ACCUM64: return ctx->views->accum64 ? px * 32 : 0;
        case RC_BUF_POST_HDR: return ctx->views->postOut[view] ? px * 16 : 0;
        case RC_BUF_DISPLAY: return ctx->views->postOut[view] ? px * 4 : 0;
        default: return 0;
    }
}

int rc_views_read(rc_ctx *ctx, int32_t view, int32_t which, void *dst, size_t bytes) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (int rc = capture_block(ctx, "rc_views_read")) return rc;
    Views *V = ctx->views;
    if (!V || view < 0 || view >= V->n) return set_error(ctx, RC_ERR_STATE, "view index out of range");
    const size_t need = rc_views_buffer_bytes(ctx, view, which);
    if (need == 0) return set_error(ctx, RC_ERR_STATE, "buffer not available for this view");
    if (!dst || bytes < need) return set_error(ctx, RC_ERR_INVALID_ARG, "destination too small");
    const size_t o = V->pixOff[view];
    const void *src = nullptr;
    switch (which) {
        case RC_BUF_ACCUM: src = V->accum + o; break;
        case RC_BUF_GBUFFER: src = V->gbuf + o; break;
        case RC_BUF_GALBEDO: src = V->galbedo + o; break;
        case RC_BUF_VARIANCE: src = V->variance + o; break;
        case RC_BUF_PRIMARY_ID: src = V->primId + o; break;
        case RC_BUF_GBUFFER_ID: src = V->gbufId + o; break;
        case RC_BUF_ACCUM64: src = V->accum64 + o; break;
        case RC_BUF_POST_HDR: src = V->postOut[view]; break;
        case RC_BUF_DISPLAY: src = V->display + o; break;
        default: return set_error(ctx, RC_ERR_INVALID_ARG, "unknown buffer");
    }
    (void)hipSetDevice(ctx->device);
    hipError_t e = hipMemcpyAsync(dst, src, need, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "view readback");
    return RC_OK;
}

int32_t rc_views_total_samples(const rc_ctx *ctx, int32_t view) {
    if (!ctx || !ctx->views || view < 0 || view >= ctx->views->n) return 0;
    return ctx->views->total[view];
}

}  // extern "C"
