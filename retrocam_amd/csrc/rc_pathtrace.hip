// rc_pathtrace.hip — progressive path-trace kernel for gfx950.
//
// One work-item per pixel; a launch runs nSamples consecutive progressive
// passes (frame indices f0 .. f0+nSamples-1) and keeps the running sums in
// registers, so the accumulation buffers are read and written once per launch.
// The per-sample arithmetic restates pathtrace.comp:401-655 (raygen,
// traceRadiance, NEE, accumulate) with the same operation order.
//
// Scheduling: each lane owns one pixel and regenerates its path state as soon
// as a path ends (next LCA channel, else next sample), so the lanes of a wave
// stay busy until every lane has finished all its samples, instead of waiting at
// each sample for the longest path of the wave. Samples of a pixel are still
// accumulated in frame order, so every sum is the sequential one of the
// reference. Traversal stacks live in LDS (rc_device.h).
//
// Launch shape: 256-thread workgroups covering 16x16 pixel tiles, each wave64
// an 8x8 sub-tile so camera rays of one wave are spatially coherent.
#include "rc_device.h"
#include "rc_internal.h"

namespace rc {

struct CamArgs {
    float pos[3], right[3], up[3], fwd[3];
    float focalDist, lensRadius, tanHalf, saStrength;
    float lca[3];
    int blades, width, height;
};

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

template <bool STATS, bool FP64>
__global__ void __launch_bounds__(kBlock) pathtrace_kernel(PassArgs A) {
    __shared__ int2 stk[kShortStack][kBlock];
    const StackRef st{stk, (int)threadIdx.x};
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int px = blockIdx.x * 16 + (wave & 1) * 8 + (lane & 7);
    const int py = blockIdx.y * 16 + (wave >> 1) * 8 + (lane >> 3);
    Counters C;
    if (STATS) C = Counters{};
    const SceneView &S = A.S;
    const CamArgs &cam = A.cam;
    if (px < cam.width && py < cam.height) {
        const size_t pi = (size_t)py * (size_t)cam.width + (size_t)px;
        const bool lca = (fabsf(cam.lca[0]) + fabsf(cam.lca[2])) > 1e-4f;
        float4 acc = make_float4(0, 0, 0, 0);
        double4 acc64 = make_double4(0, 0, 0, 0);
        float var = 0.0f;
        if (A.frame0 > 0) {
            if (FP64) acc64 = A.accum64[pi];
            else acc = A.accum[pi];
            var = A.variance[pi];
        }
        v3 ro, rd;
        if (A.frame0 == 0) {
            // G-buffer from the seed-99991 camera ray (pathtrace.comp:590-595, 643-646)
            uint32_t gs = init_seed(px, py, 99991);
            gen_ray(cam, px, py, gs, 0.0f, ro, rd);
            Hit h = closest_hit<STATS>(S, ro, rd, st, C);
            float g0 = 0.0f, g1 = 1.0f, g2 = 0.0f, g3 = 1e4f;
            if (h.tri >= 0) {
                int m;
                v3 nn = hit_normal(S, h, m);
                v3 n = dot(rd, nn) < 0.0f ? nn : neg(nn);
                g0 = n.x; g1 = n.y; g2 = n.z; g3 = h.t;
            }
            A.gbuf[pi] = make_ushort4(rc_f32_to_f16(g0), rc_f32_to_f16(g1), rc_f32_to_f16(g2), rc_f32_to_f16(g3));
            A.gbufId[pi] = h.tri;
        }

        // path state
        int k = 0, ch = 0, bounce = 0;
        uint32_t seed = 0;
        v3 thr = mk(1, 1, 1), rad = mk(0, 0, 0);
        float sR = 0.0f, sG = 0.0f, sB = 0.0f;  // radiance of the current sample
        bool lastDiffuse = false, fresh = true;
        int firstTri = -1, lastTri = -1;
        while (true) {
            if (fresh) {
                if (k >= A.nSamples) break;
                const int frame = A.frame0 + k;
                float fOff = 0.0f;
                if (lca) {
                    seed = init_seed(px, py, frame * 3 + ch);
                    fOff = cam.lca[ch];
                } else {
                    seed = init_seed(px, py, frame);
                }
                gen_ray(cam, px, py, seed, fOff, ro, rd);
                thr = mk(1.0f, 1.0f, 1.0f);
                rad = mk(0.0f, 0.0f, 0.0f);
                bounce = 0;
                lastDiffuse = false;
                fresh = false;
                if (STATS) C.paths++;
            }
            bool end = false;
            Hit h = closest_hit<STATS>(S, ro, rd, st, C);
            if (h.tri < 0) {
                end = true;  // black sky (pathtrace.comp:470-473)
                if (bounce == 0 && (!lca || ch == 1) && A.frame0 + k == 0) {
                    const uint16_t one = rc_f32_to_f16(1.0f);  // miss: albedo stays (1,1,1) (:455)
                    A.galbedo[pi] = make_ushort4(one, one, one, one);
                }
            } else {
                if (STATS) C.shadeHits++;
                int mi;
                v3 hn = hit_normal(S, h, mi);
                v3 hp = ro + rd * h.t;
                const float4 *M = S.mats + 4 * (size_t)mi;
                float4 m0 = M[0], m1 = M[1], m2 = M[2];
                v3 albedo = xyz(m0);
                float metallic = m0.w, estr = m1.w, rough = m2.x, ior = m2.y, trans = m2.z;
                v3 n = dot(rd, hn) < 0.0f ? hn : neg(hn);
                if (bounce == 0 && (!lca || ch == 1)) {
                    // first-hit albedo (pathtrace.comp:481-483); the G channel's under LCA
                    firstTri = h.tri;
                    if (A.frame0 + k == 0) {
                        v3 fa = (trans > 0.5f) ? mk(1.0f, 1.0f, 1.0f) : albedo;
                        A.galbedo[pi] = make_ushort4(rc_f32_to_f16(fa.x), rc_f32_to_f16(fa.y),
                                                     rc_f32_to_f16(fa.z), rc_f32_to_f16(1.0f));
                    }
                }
                if (!lastDiffuse) rad = rad + (thr * xyz(m1)) * estr;
                if (bounce >= 2) {
                    float surv = fmaxf(albedo.x, fmaxf(albedo.y, albedo.z));
                    if (metallic > 0.5f || trans > 0.5f) surv = 0.95f;
                    surv = rc_clampf(surv, 0.01f, 0.95f);
                    if (rnd(seed) > surv) end = true;
                    else thr = thr / surv;
                }
                // The direction sampled at the last bounce is never traced and its
                // random numbers are never reused (each path re-seeds): skipped.
                const bool last = bounce + 1 >= A.maxBounces;
                if (!end) {
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
                        if (A.nee != 0) {
                            // sampleDirectLight (pathtrace.comp:341-395)
                            v3 nee = mk(0.0f, 0.0f, 0.0f);
                            if (S.nLights > 0 && S.totalLightPower > 0.0f) {
                                float us = rnd(seed);
                                int sel = S.nLights - 1;
                                for (int i = 0; i < S.nLights; ++i) {
                                    if (STATS) C.lightScans++;
                                    if (us < S.cdf[i]) { sel = i; break; }
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
                                    if (cosHit > 0.0f && cosLight > 0.0f &&
                                        unoccluded<STATS>(S, hp + n * 2e-3f, sd, d - 4e-3f, st, C)) {
                                        float area = L0.w;
                                        float pdf = pickPdf * (fmaxf(d2, area) / (area * cosLight));
                                        if (pdf > 0.0f) {
                                            v3 em = xyz(L3);
                                            float cl = ((0.299f * em.x + 0.587f * em.y) + 0.114f * em.z) * A.neeClamp;
                                            nee = mk(fminf((((albedo.x / RC_PI_F) * em.x) * cosHit) / pdf, cl),
                                                     fminf((((albedo.y / RC_PI_F) * em.y) * cosHit) / pdf, cl),
                                                     fminf((((albedo.z / RC_PI_F) * em.z) * cosHit) / pdf, cl));
                                        }
                                    }
                                }
                            }
                            rad = rad + thr * nee;
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
                }
                ++bounce;
                if (last) end = true;
            }
            if (end) {
                if (!lca) {
                    sR = rad.x; sG = rad.y; sB = rad.z;
                } else if (ch == 0) {
                    sR = rad.x;
                } else if (ch == 1) {
                    sG = rad.y;
                } else {
                    sB = rad.z;
                }
                fresh = true;
                if (lca && ch < 2) {
                    ++ch;
                } else {
                    // sample complete (pathtrace.comp:633-654)
                    ch = 0;
                    if (isnan(sR) || isnan(sG) || isnan(sB) || isinf(sR) || isinf(sG) || isinf(sB))
                        sR = sG = sB = 0.0f;
                    float luma = (sR * 0.2126f + sG * 0.7152f) + sB * 0.0722f;
                    float l2 = luma * luma;
                    if (A.frame0 + k == 0) {
                        acc = make_float4(sR, sG, sB, 1.0f);
                        acc64 = make_double4(sR, sG, sB, 1.0);
                        var = l2;
                    } else {
                        if (FP64) {
                            acc64.x += (double)sR; acc64.y += (double)sG; acc64.z += (double)sB; acc64.w += 1.0;
                        } else {
                            acc.x += sR; acc.y += sG; acc.z += sB; acc.w += 1.0f;
                        }
                        var += l2;
                    }
                    lastTri = firstTri;
                    firstTri = -1;
                    ++k;
                }
            }
        }
        if (FP64) {
            A.accum64[pi] = acc64;
            acc = make_float4((float)acc64.x, (float)acc64.y, (float)acc64.z, (float)acc64.w);
        }
        A.accum[pi] = acc;
        A.variance[pi] = var;
        A.primId[pi] = lastTri;
    }
    if (STATS) wave_add_counters(A.counters, C);
}

hipError_t launch_pathtrace(rc_ctx *ctx, const rc_camera &cam, const rc_pt_params &p, int nSamples) {
    PassArgs A;
    A.S = scene_view(ctx->scene);
    for (int k = 0; k < 3; ++k) {
        A.cam.pos[k] = cam.camPos[k]; A.cam.right[k] = cam.camRight[k];
        A.cam.up[k] = cam.camUp[k]; A.cam.fwd[k] = cam.camForward[k];
        A.cam.lca[k] = cam.lcaDelta[k];
    }
    A.cam.focalDist = cam.focalDist;
    A.cam.lensRadius = cam.lensRadius;
    A.cam.tanHalf = cam.fovTanHalf;
    A.cam.saStrength = cam.saStrength;
    A.cam.blades = cam.apertureBlades;
    A.cam.width = cam.width;
    A.cam.height = cam.height;
    A.nee = p.neeEnabled;
    A.neeClamp = p.neeFireflyClamp;
    A.maxBounces = p.maxBounces;
    A.frame0 = ctx->totalSamples;
    A.nSamples = nSamples;
    A.accum = ctx->img.accum;
    A.accum64 = ctx->img.accum64;
    A.gbuf = ctx->img.gbuf;
    A.galbedo = ctx->img.galbedo;
    A.variance = ctx->img.variance;
    A.primId = ctx->img.primId;
    A.gbufId = ctx->img.gbufId;
    A.counters = ctx->counters;
    dim3 grid((cam.width + 15) / 16, (cam.height + 15) / 16);
    const bool stats = ctx->countersEnabled && ctx->counters;
    const bool fp64 = ctx->accumMode == RC_ACCUM_FP64;
    if (stats && fp64) hipLaunchKernelGGL((pathtrace_kernel<true, true>), grid, dim3(kBlock), 0, ctx->stream, A);
    else if (stats) hipLaunchKernelGGL((pathtrace_kernel<true, false>), grid, dim3(kBlock), 0, ctx->stream, A);
    else if (fp64) hipLaunchKernelGGL((pathtrace_kernel<false, true>), grid, dim3(kBlock), 0, ctx->stream, A);
    else hipLaunchKernelGGL((pathtrace_kernel<false, false>), grid, dim3(kBlock), 0, ctx->stream, A);
    return hipGetLastError();
}

}  // namespace rc
