// rc_shade.h — path-tracing shading shared by the megakernel (rc_pathtrace.hip)
// and the wavefront kernels (rc_wavefront.hip): camera rays, BSDF sampling and
// next-event estimation, restating pathtrace.comp:252-439 and 475-581 with the
// reference's operand order (bit-identical to the CPU oracle).
#pragma once

#include "rc_device.h"
#include "rc_internal.h"
#include "retrocam_hip.h"

namespace rc {

// One progressive launch: nSamples passes from frame0 on, over the images given.
struct PassArgs {
    SceneView S;
    CamArgs cam;
    int nee;
    float neeClamp;
    int maxBounces;
    int frame0;
    int nSamples;
    float4 *accum;
    double4 *accum64;
    ushort4 *gbuf;
    ushort4 *galbedo;
    float *variance;
    int *primId;
    int *gbufId;
    unsigned long long *counters;
    int tileShift;  // lane scheduler: unit tiles of (1 << tileShift) x (64 >> tileShift) pixels
    // The camera the kernels read: a device copy of `cam` (a camera slot of the
    // context, or the entry of a device view table), never null.
    const CamArgs *dcam;
};

// The camera of a launch. Its device copy is written before the launch and
// never during it, so it is addressed through the constant address space, as
// kernel arguments are: its words are invariant scalar loads that the compiler
// reissues at each use instead of keeping 24 values live across the path loop.
__device__ __forceinline__ const CamArgs &launch_camera(const PassArgs &A) {
    typedef const __attribute__((address_space(4))) CamArgs *ConstCamPtr;
    return *(const CamArgs *)(ConstCamPtr)A.dcam;
}

// Unit tiles of the lane scheduler: 64 pixels each, tw = 1 << shift wide and
// th = 64 >> shift high; a view's units are numbered tile by tile.
struct LaneTile {
    int shift, tilesX, tilesY;
    __host__ __device__ LaneTile(int W, int H, int s)
        : shift(s), tilesX((W + (1 << s) - 1) >> s), tilesY((H + (64 >> s) - 1) >> (6 - s)) {}
    __host__ __device__ int pixels() const { return tilesX * tilesY * 64; }
    // position of pixel (px, py) in the view's tile order
    __host__ __device__ int slot(int px, int py) const {
        return ((((py >> (6 - shift)) * tilesX + (px >> shift)) << 6) | ((py & ((64 >> shift) - 1)) << shift) |
                (px & ((1 << shift) - 1)));
    }
    __host__ __device__ void pixel(int tile, int within, int &px, int &py) const {
        px = (tile % tilesX) * (1 << shift) + (within & ((1 << shift) - 1));
        py = (tile / tilesX) * (64 >> shift) + (within >> shift);
    }
};

// Lane-scheduler form (rc_pathtrace_lanes.hip): one chunk of a view's samples.
// samples: [nSamples][nTilePx] float4, nTilePx = pixels of the view rounded up
// to 8x8 tiles; counter: the view's unit counter, zeroed before the launch.
struct LaneChunk {
    float4 *samples;
    int *counter;
    int s0;        // first sample of the chunk, relative to PassArgs::frame0
    int nSamples;  // samples in this chunk (0: the view has none here)
    int nTilePx;
};

// Per-pixel frame-order accumulation of one chunk (lanes_reduce_kernel).
struct ReduceArgs {
    const float4 *samples;
    int nTilePx, nSamples, frame0, width, height, tileShift;
    float4 *accum;
    double4 *accum64;
    float *variance;
};

__device__ __forceinline__ v3 sample_vndf(v3 wo, float alpha, float ux, float uy) {
    v3 v = normalize(mk(wo.x * alpha, wo.y, wo.z * alpha));
    v3 t1 = (v.y < 0.9999f) ? normalize(cross(v, mk(0.0f, 1.0f, 0.0f))) : mk(1.0f, 0.0f, 0.0f);
    v3 t2 = cross(t1, v);
    float a = 1.0f / (1.0f + v.y);
    float r = sqrtf(ux);
    float phi = (uy < a) ? (uy / a) * RC_PI_F : RC_PI_F + ((uy - a) / (1.0f - a)) * RC_PI_F;
    float p1 = r * rc_cosf(phi);
    float p2 = (r * rc_sinf(phi)) * ((uy < a) ? 1.0f : v.y);
    float sz = sqrtf(fmaxf(0.0f, (1.0f - p1 * p1) - p2 * p2));
    v3 n = (t1 * p1 + t2 * p2) + v * sz;
    return normalize(mk(alpha * n.x, fmaxf(0.0f, n.y), alpha * n.z));
}

__device__ __forceinline__ float smith_g1(float nv, float alpha) {
    float a2 = alpha * alpha;
    float denom = nv + sqrtf(a2 + ((1.0f - a2) * nv) * nv);
    return (2.0f * nv) / fmaxf(denom, 1e-7f);
}

__device__ __forceinline__ float smith_g2(float nl, float nv, float alpha) {
    float a2 = alpha * alpha;
    float da = nv * sqrtf(a2 + ((1.0f - a2) * nl) * nl);
    float db = nl * sqrtf(a2 + ((1.0f - a2) * nv) * nv);
    return ((2.0f * nl) * nv) / fmaxf(da + db, 1e-7f);
}

__device__ __forceinline__ v3 sample_cosine(float ux, float uy, v3 n) {
    float r = sqrtf(ux);
    float phi = 6.28318530718f * uy;
    v3 l = mk(r * rc_cosf(phi), r * rc_sinf(phi), sqrtf(fmaxf(0.0f, 1.0f - ux)));
    v3 up = fabsf(n.y) < 0.9999f ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f);
    v3 t = normalize(cross(up, n));
    v3 b = cross(n, t);
    return normalize((t * l.x + b * l.y) + n * l.z);
}

// Camera ray (pathtrace.comp:401-439, sampleNGon :310-321).
__device__ __forceinline__ void gen_ray(const CamArgs &c, int px, int py, uint32_t &seed, float fOff,
                                        v3 &ro, v3 &rd) {
    float aspect = (float)c.width / (float)c.height;
    float jx = rnd(seed) - 0.5f;
    float jy = rnd(seed) - 0.5f;
    float ux = (((float)px + 0.5f) + jx) / (float)c.width;
    float uy = (((float)py + 0.5f) + jy) / (float)c.height;
    float sx = ux * 2.0f - 1.0f, sy = uy * 2.0f - 1.0f;
    v3 F = mk(c.fwd[0], c.fwd[1], c.fwd[2]);
    v3 R = mk(c.right[0], c.right[1], c.right[2]);
    v3 U = mk(c.up[0], c.up[1], c.up[2]);
    v3 P = mk(c.pos[0], c.pos[1], c.pos[2]);
    float a = (sx * aspect) * c.tanHalf;
    float b = sy * c.tanHalf;
    v3 pin = normalize((F + R * a) + U * b);
    float lx = 0.0f, ly = 0.0f;
    if (c.lensRadius > 0.0f) {
        float r1 = rnd(seed);
        float r2 = rnd(seed);
        if (r1 + r2 > 1.0f) { r1 = 1.0f - r1; r2 = 1.0f - r2; }
        int sector = (int)(rnd(seed) * (float)c.blades);
        float a0 = (6.28318530718f * (float)sector) / (float)c.blades;
        float a1 = (6.28318530718f * (float)(sector + 1)) / (float)c.blades;
        float c0 = rc_cosf(a0), s0 = rc_sinf(a0), c1 = rc_cosf(a1), s1 = rc_sinf(a1);
        lx = (r1 * c0 + r2 * c1) * c.lensRadius;
        ly = (r1 * s0 + r2 * s1) * c.lensRadius;
    }
    float nr2 = 0.0f;
    if (c.lensRadius > 1e-6f) nr2 = rc_clampf((lx * lx + ly * ly) / (c.lensRadius * c.lensRadius), 0.0f, 1.0f);
    float ef = (c.focalDist + fOff) + c.saStrength * nr2;
    ef = fmaxf(ef, 0.01f);
    float tf = ef / dot(pin, F);
    v3 fp = P + pin * tf;
    v3 o = (P + R * lx) + U * ly;
    ro = o;
    rd = normalize(fp - o);
}

// Interpolated shading normal and material of a hit (pathtrace.comp:166-172).
__device__ __forceinline__ v3 hit_normal(const SceneView &S, const Hit &h, int &mat) {
    const float4 *sh = S.triShade + 3 * (size_t)h.tri;
    float4 a = sh[0], b = sh[1], c = sh[2];
    float w = (1.0f - h.bu) - h.bv;
    mat = rc_f2i_bits(a.w);
    return normalize((xyz(a) * w + xyz(b) * h.bu) + xyz(c) * h.bv);
}

__device__ __forceinline__ void wave_add_counters(unsigned long long *dst, const Counters &C) {
    const uint32_t *c = (const uint32_t *)&C;
#pragma unroll
    for (int k = 0; k < (int)(sizeof(Counters) / 4); ++k) {
        uint32_t v = c[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if ((threadIdx.x & 63) == 0 && v) atomicAdd(&dst[k], (unsigned long long)v);
    }
}

// A next-event-estimation sample whose visibility is still to be tested: if
// the shadow ray (o, d, maxDist) is unoccluded, radiance += weight.
struct NeeSample {
    bool shoot;
    v3 o, d;
    float maxDist;
    v3 weight;  // throughput * clamped NEE contribution (pathtrace.comp:385-394, :573)
};

// One path vertex after a valid hit (pathtrace.comp:475-581): emission, Russian
// roulette, and the BSDF branch (glass, GGX metal, or Lambertian with NEE). On
// return, `ro`/`rd` hold the next ray unless the path ended or this was the last
// bounce (`last`, whose sampled direction is never traced and whose random
// numbers are never reused, so it is not sampled). Returns true if Russian
// roulette ended the path.
template <bool STATS>
__device__ __forceinline__ bool shade_vertex(const SceneView &S, int neeEnabled, float neeClamp, bool last,
                                             int bounce, float4 m0, float4 m1, float4 m2, v3 hn, v3 hp,
                                             v3 &ro, v3 &rd, v3 &thr, v3 &rad, uint32_t &seed,
                                             bool &lastDiffuse, NeeSample &ns, Counters &C) {
    ns.shoot = false;
    v3 albedo = xyz(m0);
    float metallic = m0.w, estr = m1.w, rough = m2.x, ior = m2.y, trans = m2.z;
    v3 n = dot(rd, hn) < 0.0f ? hn : neg(hn);
    if (!lastDiffuse) rad = rad + (thr * xyz(m1)) * estr;
    if (bounce >= 2) {
        float surv = fmaxf(albedo.x, fmaxf(albedo.y, albedo.z));
        if (metallic > 0.5f || trans > 0.5f) surv = 0.95f;
        surv = rc_clampf(surv, 0.01f, 0.95f);
        if (rnd(seed) > surv) return true;
        thr = thr / surv;
    }
    if (trans > 0.5f) {
        if (!last) {
            bool entering = dot(rd, hn) < 0.0f;
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
        }
        thr = thr * albedo;
        lastDiffuse = false;
    } else if (metallic > 0.5f) {
        if (!last) {
            v3 wo = neg(rd);
            float alpha = fmaxf(rough, 1e-3f);
            v3 up = fabsf(n.y) < 0.9999f ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f);
            v3 T = normalize(cross(up, n));
            v3 B = cross(n, T);
            v3 wl = mk(dot(wo, T), dot(wo, n), dot(wo, B));
            float u1 = rnd(seed);
            float u2 = rnd(seed);
            v3 hl = sample_vndf(wl, alpha, u1, u2);
            v3 hv = normalize((T * hl.x + n * hl.y) + B * hl.z);
            rd = reflect(rd, hv);
            if (dot(rd, n) <= 0.0f) rd = reflect(neg(wo), n);
            float nl = fmaxf(dot(n, rd), 1e-4f);
            float nv = fmaxf(dot(n, wo), 1e-4f);
            float hdv = fmaxf(dot(hv, wo), 1e-4f);
            float pw = rc_powf(1.0f - hdv, 5.0f);
            v3 F = mk(albedo.x + (1.0f - albedo.x) * pw, albedo.y + (1.0f - albedo.y) * pw,
                      albedo.z + (1.0f - albedo.z) * pw);
            float g = smith_g2(nl, nv, alpha) / fmaxf(smith_g1(nv, alpha), 1e-7f);
            thr = thr * (F * g);
            ro = hp + n * 2e-3f;
        }
        lastDiffuse = false;
    } else {
        if (neeEnabled != 0 && S.nLights > 0 && S.totalLightPower > 0.0f) {
            // sampleDirectLight (pathtrace.comp:341-395). An early-out or a zero
            // pdf adds thr * 0 in the reference, which leaves radiance unchanged.
            float us = rnd(seed);
            int sel = S.nLights - 1;
            if (S.cdfSorted) {
                // first i with us < cdf[i] (else the last light) by bisection: with a
                // non-decreasing CDF the predicate is monotone, so this is the entry
                // the reference's scan stops at; counted as that scan's length
                int lo = 0, hi = S.nLights;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (us < S.cdf[mid]) hi = mid;
                    else lo = mid + 1;
                }
                if (lo < S.nLights) sel = lo;
                if (STATS) C.lightScans += (uint32_t)(lo < S.nLights ? lo + 1 : S.nLights);
            } else {
                for (int i = 0; i < S.nLights; ++i) {
                    if (STATS) C.lightScans++;
                    if (us < S.cdf[i]) { sel = i; break; }
                }
            }
            if (STATS) C.lightReads++;
            const float4 *L = S.lights + 5 * (size_t)sel;
            float4 L0 = L[0], L3 = L[3];
            float pickPdf = L3.w / S.totalLightPower;
            if (pickPdf > 0.0f) {
                float4 L1 = L[1], L2 = L[2], L4 = L[4];
                float r1 = rnd(seed);
                float r2 = rnd(seed);
                if (r1 + r2 > 1.0f) { r1 = 1.0f - r1; r2 = 1.0f - r2; }
                v3 lv0 = xyz(L0);
                v3 lp = (lv0 + (xyz(L1) - lv0) * r1) + (xyz(L2) - lv0) * r2;
                v3 tl = lp - hp;
                float d2 = dot(tl, tl);
                float d = sqrtf(d2);
                v3 sd = tl / d;
                float cosHit = dot(n, sd);
                float cosLight = fabsf(dot(xyz(L4), sd));
                if (cosHit > 0.0f && cosLight > 0.0f) {
                    float area = L0.w;
                    float pdf = pickPdf * (fmaxf(d2, area) / (area * cosLight));
                    if (pdf > 0.0f) {
                        v3 em = xyz(L3);
                        float cl = ((0.299f * em.x + 0.587f * em.y) + 0.114f * em.z) * neeClamp;
                        v3 c = mk(fminf((((albedo.x / RC_PI_F) * em.x) * cosHit) / pdf, cl),
                                  fminf((((albedo.y / RC_PI_F) * em.y) * cosHit) / pdf, cl),
                                  fminf((((albedo.z / RC_PI_F) * em.z) * cosHit) / pdf, cl));
                        ns.shoot = true;
                        ns.o = hp + n * 2e-3f;
                        ns.d = sd;
                        ns.maxDist = d - 4e-3f;
                        ns.weight = thr * c;
                    }
                }
            }
        }
        if (!last) {
            float u1 = rnd(seed);
            float u2 = rnd(seed);
            rd = sample_cosine(u1, u2, n);
            ro = hp + n * 1e-3f;
        }
        thr = thr * albedo;
        lastDiffuse = true;
    }
    return false;
}

// Launch arguments of nSamples passes from frame ctx->totalSamples on the
// context's scene and images (rc_pathtrace.hip).
PassArgs pass_args(rc_ctx *ctx, const rc_camera &cam, const rc_pt_params &p, int nSamples);
// Persistent launch over nViews views (rc_pathtrace.hip); views and tileOff in
// device memory, counter a device int the launch resets first.
hipError_t launch_pathtrace_views(rc_ctx *ctx, const PassArgs *views, const int *tileOff, int nViews, int *counter,
                                  int maxBlocks, bool fp64);

// Lane-scheduler path tracer (rc_pathtrace_lanes.hip). Usable when the scene's
// BVH is shallower than 30 levels (lanes_supported); same images as the
// megakernel. launch_lanes runs A.nSamples samples in chunks of at most
// lanes_chunk_samples (sample buffer <= rc_ctx::laneBufferMiB); launch_lanes_views
// runs one chunk of many views.
bool lanes_supported(const DeviceScene &D);
int lanes_chunk_samples(int width, int height, int tileShift, int nSamples, long long capBytes);
hipError_t ensure_lane_buffers(rc_ctx *ctx, size_t sampleBytes, int counters);
void free_lane_buffers(LaneBuffers &L);
hipError_t launch_lanes(rc_ctx *ctx, const PassArgs &A, int thresh, int variant);
// zeroCounters > 0: first zero that many unit counters (all chunks' at once).
hipError_t launch_lanes_views(rc_ctx *ctx, const PassArgs *dViews, LaneChunk *dChunks, const long long *dPixOff,
                              int nViews, long long totalPx, long long totalUnits, int thresh, int variant,
                              bool fp64, int zeroCounters);

}  // namespace rc
