// rc_sppm.hip — SPPM on gfx950: photon emission + specular transport + hash
// insert (photon_trace.comp), the per-pixel radiance gather (photon_gather.comp)
// and the radius schedule (SPPMManager.java:170-173). Shader paths are relative
// to src/main/resources/shaders/.
//
// Photon map in HBM, as PhotonMap.java:41-75 lays it out: photon records
// (position, power, direction; 48 B), hash heads (int32 per cell, -1 = empty),
// next links (int32 per photon) and one atomic counter. Insertion keeps the
// reference's lock-free list: idx = atomicAdd(counter), then
// next[idx] = atomicExch(heads[cell], idx) — list order is the atomic order.
#include <cmath>

#include "rc_device.h"
#include "rc_internal.h"

namespace rc {

__device__ __forceinline__ int cell_hash(int cx, int cy, int cz, int hashSize) {
    uint32_t h = ((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u) ^ ((uint32_t)cz * 83492791u);
    return (int)(h & (uint32_t)(hashSize - 1));
}

__global__ void __launch_bounds__(256) fill_i32_kernel(int *p, int n, int v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

struct TraceArgs {
    SceneView S;
    int photonCount, seed, hashSize;
    float radius;
    float4 *photons;
    int *heads, *next;
    unsigned int *counter;
    unsigned int *occupied;  // one bit per cell: set when a photon is stored in it
    int *work;               // photon_trace_lanes: next photon index to take
};

// Photon gid leaving the light (photon_trace.comp:168-200): light choice by
// the power CDF, a uniform point on the light triangle and a cosine-weighted
// direction, all drawn from the photon's own seed.
__device__ __forceinline__ void emit_photon(const TraceArgs &A, uint32_t gid, uint32_t &seed, v3 &ro, v3 &rd,
                                            v3 &power) {
    const SceneView &S = A.S;
    seed = gid * 1664525u + (uint32_t)A.seed * 1013904223u + 2531011u;
    pcg(seed);
    pcg(seed);
    float u = rnd(seed);
    int sel = S.nLights - 1;
    for (int i = 0; i < S.nLights; ++i)
        if (u < S.cdf[i]) { sel = i; break; }
    const float4 *L = S.lights + 5 * (size_t)sel;
    float4 L0 = L[0], L1 = L[1], L2 = L[2], L3 = L[3], L4 = L[4];
    float r1 = rnd(seed);
    float r2 = rnd(seed);
    if (r1 + r2 > 1.0f) { r1 = 1.0f - r1; r2 = 1.0f - r2; }
    v3 lv0 = xyz(L0);
    v3 sp = (lv0 + (xyz(L1) - lv0) * r1) + (xyz(L2) - lv0) * r2;
    v3 en = normalize(xyz(L4));
    float ux = rnd(seed);
    float uy = rnd(seed);
    float rr = sqrtf(ux), phi = 6.28318530718f * uy;
    v3 loc = mk(rr * rc_cosf(phi), rr * rc_sinf(phi), sqrtf(fmaxf(0.0f, 1.0f - ux)));
    v3 up = fabsf(en.y) < 0.9999f ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f);
    v3 tt = normalize(cross(up, en));
    v3 bb = cross(en, tt);
    rd = normalize((tt * loc.x + bb * loc.y) + en * loc.z);
    v3 em = xyz(L3);
    float lum = (0.299f * em.x + 0.587f * em.y) + 0.114f * em.z;
    float pp = S.totalLightPower / (float)A.photonCount;
    power = lum > 1e-6f ? mk((em.x / lum) * pp, (em.y / lum) * pp, (em.z / lum) * pp) : mk(pp, pp, pp);
    ro = sp + en * 1e-3f;
}

// One segment of a photon's path (photon_trace.comp:201-251): the closest hit,
// then glass (Fresnel-chosen reflection or refraction) or metal (reflection,
// power times albedo) continue the path; a diffuse hit after a specular bounce
// stores the photon (storePhoton, :152-165) and ends it, as does a miss. Returns
// whether the path continues.
__device__ __forceinline__ bool photon_segment(const TraceArgs &A, const StackRef &st, uint32_t &seed, v3 &ro, v3 &rd,
                                               v3 &power, bool &spec) {
    const SceneView &S = A.S;
    Hit h = closest_hit_unordered(S, ro, rd, st);
    if (h.tri < 0) return false;
    const float4 *sh = S.triShade + 3 * (size_t)h.tri;
    float4 na = sh[0], nb = sh[1], nc = sh[2];
    float w = (1.0f - h.bu) - h.bv;
    v3 hn = normalize((xyz(na) * w + xyz(nb) * h.bu) + xyz(nc) * h.bv);
    v3 hp = ro + rd * h.t;
    const float4 *M = S.mats + 4 * (size_t)rc_f2i_bits(na.w);
    float4 m0 = M[0], m2 = M[2];
    v3 n = dot(rd, hn) < 0.0f ? hn : neg(hn);
    if (m2.z > 0.5f) {
        bool entering = dot(rd, hn) < 0.0f;
        float ior = m2.y;
        float eta = entering ? (1.0f / ior) : ior;
        float cosI = fabsf(dot(rd, n));
        float r0 = (1.0f - ior) / (1.0f + ior);
        r0 = r0 * r0;
        float F = r0 + (1.0f - r0) * rc_powf(rc_clampf(1.0f - cosI, 0.0f, 1.0f), 5.0f);
        if (rnd(seed) < F) {
            rd = reflect(rd, n);
        } else {
            v3 rf = refract(rd, n, eta);
            if (dot(rf, rf) < 0.001f) rf = reflect(rd, n);
            rd = normalize(rf);
        }
        ro = hp + rd * 2e-3f;
        spec = true;
        return true;
    }
    if (m0.w > 0.5f) {
        power = power * xyz(m0);
        rd = reflect(rd, n);
        ro = hp + n * 2e-3f;
        spec = true;
        return true;
    }
    if (spec) {  // storePhoton (photon_trace.comp:152-165)
        unsigned int idx = atomicAdd(A.counter, 1u);
        if (idx < (unsigned int)A.photonCount) {
            int cx = (int)floorf(hp.x / A.radius), cy = (int)floorf(hp.y / A.radius), cz = (int)floorf(hp.z / A.radius);
            const int hc = cell_hash(cx, cy, cz, A.hashSize);
            int old = atomicExch(&A.heads[hc], (int)idx);
            atomicOr(&A.occupied[hc >> 5], 1u << (hc & 31));
            A.next[idx] = old;
            float4 *rec = A.photons + 3 * (size_t)idx;
            rec[0] = make_float4(hp.x, hp.y, hp.z, 0.0f);
            rec[1] = make_float4(power.x, power.y, power.z, 0.0f);
            rec[2] = make_float4(rd.x, rd.y, rd.z, 0.0f);
        }
    }
    return false;
}

// The photons taken a segment at a time by persistent lanes: a lane whose path
// ended takes the next photon index (one atomic per wave and round), so a wave
// never waits for its longest path (photons that enter glass bounce several
// times, most end at their first diffuse hit). Each photon's draws come from its
// own seed, so the stored photons are those of one work-item per photon (the
// reference's dispatch); only the order of insertion, an atomic order in both,
// can differ.
__global__ void __launch_bounds__(kBlock) photon_trace_lanes(TraceArgs A) {
    __shared__ int2 stk[kShortStack][kBlock];
    const StackRef st{stk, (int)threadIdx.x};
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    bool alive = false, done = false, spec = false;
    int bounce = 0;
    uint32_t seed = 0;
    v3 ro = mk(0.0f, 0.0f, 0.0f), rd = mk(0.0f, 0.0f, 1.0f), power = ro;
    while (true) {
        const unsigned long long need = __ballot(!alive && !done);
        if (need != 0ull) {
            const int leader = __ffsll((unsigned long long)need) - 1;
            int base = 0;
            if (lane == leader) base = atomicAdd(A.work, __popcll(need));
            base = __shfl(base, leader, 64);
            if (!alive && !done) {
                const int gid = base + __popcll(need & below);
                if (gid < A.photonCount) {
                    emit_photon(A, (uint32_t)gid, seed, ro, rd, power);
                    bounce = 0;
                    spec = false;
                    alive = true;
                } else {
                    done = true;
                }
            }
        }
        if (__ballot(alive) == 0ull) break;
        if (alive) {
            const bool more = photon_segment(A, st, seed, ro, rd, power, spec);
            alive = more && ++bounce < 8;
        }
    }
}


struct GatherArgs {
    SceneView S;
    const CamArgs *cam;  // device copy of the camera: position, basis, tanHalf
    int width, height, hashSize;
    float radius;
    const float4 *photons;
    const int *heads, *next;
    const unsigned int *occupied;  // cells holding at least one photon
    float4 *accum;
    double4 *accum64;
};

// photon_gather.comp:128-221, one work-item per pixel.
template <bool FP64>
__global__ void __launch_bounds__(kBlock) photon_gather_kernel(GatherArgs A) {
    __shared__ int2 stk[kShortStack][kBlock];
    const StackRef st{stk, (int)threadIdx.x};
    const SceneView &S = A.S;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int x = blockIdx.x * 16 + (wave & 1) * 8 + (lane & 7);
    const int y = blockIdx.y * 16 + (wave >> 1) * 8 + (lane >> 3);
    if (x >= A.width || y >= A.height) return;
    float aspect = (float)A.width / (float)A.height;
    float nx = (((float)x + 0.5f) / (float)A.width) * 2.0f - 1.0f;
    float ny = (((float)y + 0.5f) / (float)A.height) * 2.0f - 1.0f;
    const CamArgs &c = *A.cam;
    v3 F = mk(c.fwd[0], c.fwd[1], c.fwd[2]), R = mk(c.right[0], c.right[1], c.right[2]);
    v3 U = mk(c.up[0], c.up[1], c.up[2]);
    v3 rd = normalize((F + R * ((nx * aspect) * c.tanHalf)) + U * (ny * c.tanHalf));
    v3 ro = mk(c.pos[0], c.pos[1], c.pos[2]);
    bool found = false;
    v3 hp = mk(0, 0, 0);
    float4 alb = make_float4(0, 0, 0, 0);
    for (int s = 0; s <= 6; ++s) {
        Hit h = closest_hit_unordered(S, ro, rd, st);
        if (h.tri < 0) break;
        const float4 *sh = S.triShade + 3 * (size_t)h.tri;
        float4 na = sh[0], nb = sh[1], nc = sh[2];
        const float4 *M = S.mats + 4 * (size_t)rc_f2i_bits(na.w);
        float4 m0 = M[0], m2 = M[2];
        if (m0.w <= 0.5f && m2.z <= 0.5f) {
            found = true;
            hp = ro + rd * h.t;
            alb = m0;
            break;
        }
        float w = (1.0f - h.bu) - h.bv;
        v3 hn = normalize((xyz(na) * w + xyz(nb) * h.bu) + xyz(nc) * h.bv);
        v3 hpos = ro + rd * h.t;
        v3 n = dot(rd, hn) < 0.0f ? hn : neg(hn);
        if (m2.z > 0.5f) {
            bool entering = dot(rd, hn) < 0.0f;
            float eta = entering ? (1.0f / m2.y) : m2.y;
            v3 rf = refract(rd, n, eta);
            if (dot(rf, rf) < 0.001f) rf = reflect(rd, n);
            rd = normalize(rf);
        } else {
            rd = reflect(rd, n);
        }
        ro = hpos + rd * 2e-3f;
    }
    if (!found) return;
    float r = A.radius, r2 = r * r;
    int cx = (int)floorf(hp.x / r), cy = (int)floorf(hp.y / r), cz = (int)floorf(hp.z / r);
    v3 sum = mk(0, 0, 0);
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                // a cell whose bit is clear has head -1: skipping it visits the same photons
                // in the same order without fetching its head
                const int h = cell_hash(cx + dx, cy + dy, cz + dz, A.hashSize);
                if (!((A.occupied[h >> 5] >> (h & 31)) & 1u)) continue;
                int idx = A.heads[h];
                int safety = 0;
                while (idx >= 0 && safety < 512) {
                    const float4 *ph = A.photons + 3 * (size_t)idx;
                    float4 pp = ph[0];
                    v3 d = hp - xyz(pp);
                    float d2 = dot(d, d);
                    if (d2 <= r2) {
                        float w = 1.0f - d2 / r2;
                        sum = sum + xyz(ph[1]) * w;
                    }
                    idx = A.next[idx];
                    ++safety;
                }
            }
    float den = 3.14159265358979f * r2;
    v3 L = mk((alb.x * sum.x) / den, (alb.y * sum.y) / den, (alb.z * sum.z) / den);
    if (isnan(L.x) || isnan(L.y) || isnan(L.z) || isinf(L.x) || isinf(L.y) || isinf(L.z)) L = mk(0, 0, 0);
    size_t pi = (size_t)y * (size_t)A.width + (size_t)x;
    if (FP64) {
        double4 a = A.accum64[pi];
        a.x += (double)L.x; a.y += (double)L.y; a.z += (double)L.z;
        A.accum64[pi] = a;
        float4 f = A.accum[pi];
        A.accum[pi] = make_float4((float)a.x, (float)a.y, (float)a.z, f.w);
    } else {
        float4 a = A.accum[pi];
        A.accum[pi] = make_float4(a.x + L.x, a.y + L.y, a.z + L.z, a.w);
    }
}

void free_sppm(Sppm &s) {
    if (s.photons) (void)hipFree(s.photons);
    if (s.heads) (void)hipFree(s.heads);
    if (s.next) (void)hipFree(s.next);
    if (s.counter) (void)hipFree(s.counter);
    if (s.occupied) (void)hipFree(s.occupied);
    s.photons = nullptr; s.heads = nullptr; s.next = nullptr; s.counter = nullptr; s.occupied = nullptr;
    s.capPhotons = s.capHash = 0;
}

static int ensure_sppm(rc_ctx *ctx, int photons, int hashSize) {
    Sppm &s = ctx->sppm;
    if (s.capPhotons >= photons && s.capHash >= hashSize && s.counter) return RC_OK;
    if (int rc = capture_block(ctx, "allocating the photon map")) return rc;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
    int cp = s.capPhotons > photons ? s.capPhotons : photons;
    int ch = s.capHash > hashSize ? s.capHash : hashSize;
    ++ctx->freeGen;
    free_sppm(s);
    e = hipMalloc((void **)&s.photons, (size_t)cp * 48);
    if (e == hipSuccess) e = hipMalloc((void **)&s.next, (size_t)cp * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&s.heads, (size_t)ch * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&s.counter, 64);
    if (e == hipSuccess) e = hipMalloc((void **)&s.occupied, (size_t)((ch + 31) / 32) * 4);
    if (e != hipSuccess) {
        free_sppm(s);
        return hip_fail(ctx, e, "hipMalloc(photon map)");
    }
    s.capPhotons = cp;
    s.capHash = ch;
    return RC_OK;
}

}  // namespace rc

using namespace rc;

// Photon capacity limit: 16x the reference's 2,000,000 (PhotonMap.java:44); the
// photon map then takes < 1.3 GB of the 288 GB HBM.
static const int kMaxPhotons = 32000000;
static const int kMaxHash = 1 << 26;

extern "C" {

int rc_sppm_reset(rc_ctx *ctx, float initRadius) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (!(initRadius > 0.0f)) return set_error(ctx, RC_ERR_INVALID_ARG, "initRadius must be > 0");
    ctx->sppm.radius = initRadius;
    ctx->sppm.radius64 = (double)initRadius;
    ctx->sppm.iteration = 0;
    return RC_OK;
}

int rc_sppm_set_precision(rc_ctx *ctx, int32_t fp64Radius) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    ctx->sppm.fp64Radius = fp64Radius ? 1 : 0;
    ctx->sppm.radius64 = (double)ctx->sppm.radius;
    return RC_OK;
}

int rc_sppm_trace(rc_ctx *ctx, int32_t photonCount, int32_t hashSize) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (photonCount <= 0 || hashSize <= 0 || (hashSize & (hashSize - 1)))
        return set_error(ctx, RC_ERR_INVALID_ARG, "photonCount must be > 0 and hashSize a power of two");
    if (photonCount > kMaxPhotons || hashSize > kMaxHash)
        return set_error(ctx, RC_ERR_LIMIT, "photonCount <= 32,000,000 and hashSize <= 2^26");
    if (!ctx->scene.valid) return set_error(ctx, RC_ERR_STATE, "no scene uploaded");
    (void)hipSetDevice(ctx->device);
    int rc = ensure_sppm(ctx, photonCount, hashSize);
    if (rc) return rc;
    Sppm &s = ctx->sppm;
    // PhotonMap.clear (PhotonMap.java:84-106): heads := -1, counter := 0 (fill kernels)
    hipLaunchKernelGGL(fill_i32_kernel, dim3((hashSize + 255) / 256), dim3(256), 0, ctx->stream, s.heads, hashSize, -1);
    hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(256), 0, ctx->stream, (int *)s.counter, 2, 0);  // + work index
    const int words = (hashSize + 31) / 32;
    hipLaunchKernelGGL(fill_i32_kernel, dim3((words + 255) / 256), dim3(256), 0, ctx->stream, (int *)s.occupied, words,
                       0);
    s.photonCount = photonCount;
    s.hashSize = hashSize;
    if (ctx->scene.nLights > 0) {
        TraceArgs A;
        A.S = scene_view(ctx->scene);
        A.photonCount = photonCount;
        A.seed = s.iteration * 7919 + 6271;  // SPPMManager.java:88
        A.hashSize = hashSize;
        A.radius = s.radius;
        A.photons = s.photons;
        A.heads = s.heads;
        A.next = s.next;
        A.counter = s.counter;
        A.occupied = s.occupied;
        A.work = (int *)s.counter + 1;
        const int grid = persistent_blocks((const void *)photon_trace_lanes, photonCount);
        hipLaunchKernelGGL(photon_trace_lanes, dim3(grid), dim3(kBlock), 0, ctx->stream, A);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(ctx, e, "photon trace launch");
    return RC_OK;
}

int rc_sppm_gather(rc_ctx *ctx, const rc_camera *cam) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (!cam) return set_error(ctx, RC_ERR_INVALID_ARG, "camera is NULL");
    Sppm &s = ctx->sppm;
    if (!s.counter || s.photonCount <= 0) return set_error(ctx, RC_ERR_STATE, "rc_sppm_trace has not run");
    if (ctx->img.width <= 0 || ctx->totalSamples <= 0)
        return set_error(ctx, RC_ERR_STATE, "no accumulation to gather into (render first)");
    if (cam->width != ctx->img.width || cam->height != ctx->img.height)
        return set_error(ctx, RC_ERR_INVALID_ARG, "camera resolution differs from the accumulation buffer");
    if (ctx->capturing && !ctx->captureResetSeen)
        return set_error(ctx, RC_ERR_STATE, "a captured frame must begin with rc_reset before gathering into it");
    (void)hipSetDevice(ctx->device);
    GatherArgs A;
    A.S = scene_view(ctx->scene);
    A.width = cam->width;
    A.height = cam->height;
    A.hashSize = s.hashSize;
    A.radius = s.radius;
    A.photons = s.photons;
    A.heads = s.heads;
    A.next = s.next;
    A.occupied = s.occupied;
    A.accum = ctx->img.accum;
    A.accum64 = ctx->img.accum64;
    if (int rc = camera_slot(ctx, cam_args(*cam), &A.cam)) return rc;
    dim3 grid((cam->width + 15) / 16, (cam->height + 15) / 16);
    if (ctx->accumMode == RC_ACCUM_FP64)
        hipLaunchKernelGGL(photon_gather_kernel<true>, grid, dim3(kBlock), 0, ctx->stream, A);
    else
        hipLaunchKernelGGL(photon_gather_kernel<false>, grid, dim3(kBlock), 0, ctx->stream, A);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(ctx, e, "photon gather launch");
    return RC_OK;
}

int rc_sppm_update_radius(rc_ctx *ctx, float alpha) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    Sppm &s = ctx->sppm;
    // SPPMManager.updateRadius: r *= (float) Math.sqrt((iteration + alpha) / (iteration + 1.0))
    float num = (float)s.iteration + alpha;
    double q = (double)num / ((double)s.iteration + 1.0);
    s.radius64 *= std::sqrt(((double)s.iteration + (double)alpha) / ((double)s.iteration + 1.0));
    s.radius = s.fp64Radius ? (float)s.radius64 : s.radius * (float)std::sqrt(q);
    s.iteration++;
    return RC_OK;
}

int rc_sppm_get_state(const rc_ctx *ctx, float *radius, int32_t *iteration, double *radius64) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (radius) *radius = ctx->sppm.radius;
    if (iteration) *iteration = ctx->sppm.iteration;
    if (radius64) *radius64 = ctx->sppm.radius64;
    return RC_OK;
}

int rc_sppm_read(rc_ctx *ctx, float *photons, int32_t *heads, int32_t *next, uint32_t *counter) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (int rc = capture_block(ctx, "rc_sppm_read")) return rc;
    Sppm &s = ctx->sppm;
    if (!s.counter || s.photonCount <= 0) return set_error(ctx, RC_ERR_STATE, "rc_sppm_trace has not run");
    (void)hipSetDevice(ctx->device);
    hipError_t e = hipSuccess;
    if (photons) e = hipMemcpyAsync(photons, s.photons, (size_t)s.photonCount * 48, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && heads)
        e = hipMemcpyAsync(heads, s.heads, (size_t)s.hashSize * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && next)
        e = hipMemcpyAsync(next, s.next, (size_t)s.photonCount * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && counter) e = hipMemcpyAsync(counter, s.counter, 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "rc_sppm_read");
    return RC_OK;
}

}  // extern "C"
