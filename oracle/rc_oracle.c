/*
 * rc_oracle.c — CPU oracle: scalar restatement of pathtrace.comp.
 *
 * TEST INFRASTRUCTURE ONLY (see rc_oracle.h). Built with -O3 -ffp-contract=off.
 * Paths cited below are relative to src/main/resources/shaders/.
 */
#include "rc_oracle.h"
#include "rc_fmath.h"

#include <math.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { float x, y, z; } v3;

static inline v3 V(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 muls(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 divs(v3 a, float s) { return V(a.x / s, a.y / s, a.z / s); }
static inline float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* GLSL cross: (x1*y2 - y1*x2, x2*y0 - y2*x0, x0*y1 - y0*x1). */
static inline v3 cross3(v3 a, v3 b) {
    return V(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
/* normalize(v) = v * (1 / sqrt(dot(v, v))). */
static inline v3 normalize3(v3 a) { float inv = 1.0f / sqrtf(dot3(a, a)); return muls(a, inv); }
/* reflect(I, N) = I - 2 dot(N, I) N. */
static inline v3 reflect3(v3 i, v3 n) { float d = 2.0f * dot3(n, i); return sub(i, muls(n, d)); }
/* refract(I, N, eta) (GLSL 4.30 spec, 8.5). */
static inline v3 refract3(v3 i, v3 n, float eta) {
    float d = dot3(n, i);
    float k = 1.0f - (eta * eta) * (1.0f - d * d);
    if (k < 0.0f) return V(0.0f, 0.0f, 0.0f);
    float c = eta * d + sqrtf(k);
    return sub(muls(i, eta), muls(n, c));
}
static inline v3 ld3(const float *p) { return V(p[0], p[1], p[2]); }
static inline int32_t ldi(const float *p) { int32_t i; memcpy(&i, p, 4); return i; }

/* ---- RNG (pathtrace.comp:80-98) ----------------------------------------- */
uint32_t rco_pcg(uint32_t *state) {
    uint32_t s = *state * 747796405u + 2891336453u;
    *state = s;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}
float rco_rand(uint32_t *state) { return (float)rco_pcg(state) / 4294967296.0f; }
uint32_t rco_init_seed(int32_t x, int32_t y, int32_t frame) {
    uint32_t s = (uint32_t)x + (uint32_t)y * 7919u + (uint32_t)frame * 6271u;
    rco_pcg(&s);
    rco_pcg(&s);
    return s;
}

typedef struct { rco_stats st; } ctr;

/* ---- intersection (pathtrace.comp:102-130) ------------------------------ */
static int aabb_hit(v3 ro, v3 inv, const float *bmin, const float *bmax, float *tmin) {
    float t0x = (bmin[0] - ro.x) * inv.x, t0y = (bmin[1] - ro.y) * inv.y, t0z = (bmin[2] - ro.z) * inv.z;
    float t1x = (bmax[0] - ro.x) * inv.x, t1y = (bmax[1] - ro.y) * inv.y, t1z = (bmax[2] - ro.z) * inv.z;
    float lox = fminf(t0x, t1x), loy = fminf(t0y, t1y), loz = fminf(t0z, t1z);
    float hix = fmaxf(t0x, t1x), hiy = fmaxf(t0y, t1y), hiz = fmaxf(t0z, t1z);
    float tn = fmaxf(fmaxf(lox, loy), loz);
    float tx = fminf(fminf(hix, hiy), hiz);
    *tmin = tn;
    return tx >= fmaxf(tn, 0.0f);
}

static int tri_hit(const rco_scene *s, v3 ro, v3 rd, int idx, float *t, float *bu, float *bv) {
    const float *T = s->tris + 32 * (size_t)idx;
    v3 v0 = ld3(T), v1 = ld3(T + 4), v2 = ld3(T + 8);
    v3 e1 = sub(v1, v0), e2 = sub(v2, v0);
    v3 h = cross3(rd, e2);
    float a = dot3(e1, h);
    if (fabsf(a) < 1e-8f) return 0;
    float f = 1.0f / a;
    v3 sv = sub(ro, v0);
    float u = f * dot3(sv, h);
    if (u < 0.0f || u > 1.0f) return 0;
    v3 q = cross3(sv, e1);
    float v = f * dot3(rd, q);
    if (v < 0.0f || u + v > 1.0f) return 0;
    float tt = f * dot3(e2, q);
    if (tt < 1e-4f) return 0;
    *t = tt; *bu = u; *bv = v;
    return 1;
}

typedef struct { int valid; float t; v3 pos, normal; int mat; int tri; } hitrec;

/* traverseBVH (pathtrace.comp:142-203). */
static hitrec traverse(const rco_scene *s, v3 ro, v3 rd, rco_stats *st) {
    hitrec rec;
    rec.valid = 0; rec.t = 1e30f; rec.tri = -1; rec.mat = 0;
    rec.pos = V(0, 0, 0); rec.normal = V(0, 0, 0);
    v3 inv = V(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    int stack[64];
    int top = 0;
    stack[top++] = 0;
    if (st) st->rays++;
    while (top > 0) {
        int idx = stack[--top];
        const float *N = s->nodes + 12 * (size_t)idx;
        if (st) st->nodePops++;
        float at;
        if (!aabb_hit(ro, inv, N, N + 4, &at)) continue;
        if (at > rec.t) continue;
        int left = ldi(N + 8), right = ldi(N + 9), ps = ldi(N + 10), pc = ldi(N + 11);
        if (pc > 0) {
            for (int i = ps; i < ps + pc; ++i) {
                float t, bu, bv;
                if (st) st->triTests++;
                if (tri_hit(s, ro, rd, i, &t, &bu, &bv) && t < rec.t) {
                    const float *T = s->tris + 32 * (size_t)i;
                    rec.t = t;
                    rec.pos = add(ro, muls(rd, t));
                    float w = (1.0f - bu) - bv;
                    v3 n0 = ld3(T + 12), n1 = ld3(T + 16), n2 = ld3(T + 20);
                    rec.normal = normalize3(add(add(muls(n0, w), muls(n1, bu)), muls(n2, bv)));
                    rec.mat = ldi(T + 30);
                    rec.valid = 1;
                    rec.tri = i;
                    if (st) st->hitUpdates++;
                }
            }
        } else if (top < 62) {
            const float *L = s->nodes + 12 * (size_t)left;
            const float *R = s->nodes + 12 * (size_t)right;
            float tl, tr;
            if (st) st->childTests++;
            int hl = aabb_hit(ro, inv, L, L + 4, &tl);
            int hr = aabb_hit(ro, inv, R, R + 4, &tr);
            if (hl && hr) {
                if (tl < tr) { stack[top++] = right; stack[top++] = left; }
                else { stack[top++] = left; stack[top++] = right; }
            } else if (hl) {
                stack[top++] = left;
            } else if (hr) {
                stack[top++] = right;
            }
        }
    }
    return rec;
}

/* shadowRay (pathtrace.comp:211-242). */
static int shadow(const rco_scene *s, v3 ro, v3 rd, float maxDist, rco_stats *st) {
    v3 inv = V(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    int stack[32];
    int top = 0;
    stack[top++] = 0;
    if (st) st->shadowRays++;
    while (top > 0) {
        int idx = stack[--top];
        const float *N = s->nodes + 12 * (size_t)idx;
        if (st) st->shadowPops++;
        float at;
        if (!aabb_hit(ro, inv, N, N + 4, &at)) continue;
        if (at > maxDist) continue;
        int left = ldi(N + 8), right = ldi(N + 9), ps = ldi(N + 10), pc = ldi(N + 11);
        if (pc > 0) {
            for (int i = ps; i < ps + pc; ++i) {
                float t, bu, bv;
                if (st) st->shadowTriTests++;
                if (tri_hit(s, ro, rd, i, &t, &bu, &bv) && t < maxDist) return 0;
            }
        } else if (top < 30) {
            stack[top++] = right;
            stack[top++] = left;
        }
    }
    return 1;
}

int32_t rco_closest_hit(const rco_scene *s, const float *ro, const float *rd, float *tOut, float *nOut) {
    hitrec h = traverse(s, ld3(ro), ld3(rd), NULL);
    if (tOut) *tOut = h.t;
    if (nOut) { nOut[0] = h.normal.x; nOut[1] = h.normal.y; nOut[2] = h.normal.z; }
    return h.valid ? h.tri : -1;
}

int32_t rco_brute_closest(const rco_scene *s, const float *ro, const float *rd, float *tOut) {
    float best = 1e30f;
    int bi = -1;
    for (int i = 0; i < s->nTris; ++i) {
        float t, bu, bv;
        if (tri_hit(s, ld3(ro), ld3(rd), i, &t, &bu, &bv) && t < best) { best = t; bi = i; }
    }
    if (tOut) *tOut = best;
    return bi;
}

int32_t rco_shadow_ray(const rco_scene *s, const float *ro, const float *rd, float maxDist) {
    return shadow(s, ld3(ro), ld3(rd), maxDist, NULL);
}

/* ---- sampling (pathtrace.comp:252-321) ---------------------------------- */
static v3 sample_vndf(v3 wo, float alpha, float ux, float uy) {
    v3 v = normalize3(V(wo.x * alpha, wo.y, wo.z * alpha));
    v3 t1 = (v.y < 0.9999f) ? normalize3(cross3(v, V(0.0f, 1.0f, 0.0f))) : V(1.0f, 0.0f, 0.0f);
    v3 t2 = cross3(t1, v);
    float a = 1.0f / (1.0f + v.y);
    float r = sqrtf(ux);
    float phi = (uy < a) ? (uy / a) * RC_PI_F : RC_PI_F + ((uy - a) / (1.0f - a)) * RC_PI_F;
    float p1 = r * rc_cosf(phi);
    float p2 = (r * rc_sinf(phi)) * ((uy < a) ? 1.0f : v.y);
    float sz = sqrtf(fmaxf(0.0f, (1.0f - p1 * p1) - p2 * p2));
    v3 n = add(add(muls(t1, p1), muls(t2, p2)), muls(v, sz));
    return normalize3(V(alpha * n.x, fmaxf(0.0f, n.y), alpha * n.z));
}

static float smith_g1(float nv, float alpha) {
    float a2 = alpha * alpha;
    float denom = nv + sqrtf(a2 + ((1.0f - a2) * nv) * nv);
    return (2.0f * nv) / fmaxf(denom, 1e-7f);
}

static float smith_g2(float nl, float nv, float alpha) {
    float a2 = alpha * alpha;
    float da = nv * sqrtf(a2 + ((1.0f - a2) * nl) * nl);
    float db = nl * sqrtf(a2 + ((1.0f - a2) * nv) * nv);
    return ((2.0f * nl) * nv) / fmaxf(da + db, 1e-7f);
}

static v3 sample_cosine(float ux, float uy, v3 n) {
    float r = sqrtf(ux);
    float phi = 6.28318530718f * uy;
    v3 l = V(r * rc_cosf(phi), r * rc_sinf(phi), sqrtf(fmaxf(0.0f, 1.0f - ux)));
    v3 up = fabsf(n.y) < 0.9999f ? V(0.0f, 1.0f, 0.0f) : V(1.0f, 0.0f, 0.0f);
    v3 t = normalize3(cross3(up, n));
    v3 b = cross3(n, t);
    return normalize3(add(add(muls(t, l.x), muls(b, l.y)), muls(n, l.z)));
}

static void sample_ngon(uint32_t *seed, float radius, int n, float *ox, float *oy) {
    float r1 = rco_rand(seed);
    float r2 = rco_rand(seed);
    if (r1 + r2 > 1.0f) { r1 = 1.0f - r1; r2 = 1.0f - r2; }
    int sector = (int)(rco_rand(seed) * (float)n);
    float a0 = (6.28318530718f * (float)sector) / (float)n;
    float a1 = (6.28318530718f * (float)(sector + 1)) / (float)n;
    float c0 = rc_cosf(a0), s0 = rc_sinf(a0), c1 = rc_cosf(a1), s1 = rc_sinf(a1);
    *ox = (r1 * c0 + r2 * c1) * radius;
    *oy = (r1 * s0 + r2 * s1) * radius;
}

/* ---- NEE (pathtrace.comp:341-395) --------------------------------------- */
static v3 direct_light(const rco_scene *s, const rc_pt_params *p, v3 hp, v3 n, v3 albedo,
                       uint32_t *seed, rco_stats *st) {
    if (s->nLights <= 0 || s->totalLightPower <= 0.0f) return V(0, 0, 0);
    float us = rco_rand(seed);
    int sel = s->nLights - 1;
    for (int i = 0; i < s->nLights; ++i) {
        if (st) st->lightScans++;
        if (us < s->lights[20 * (size_t)i + 19]) { sel = i; break; }
    }
    const float *L = s->lights + 20 * (size_t)sel;
    if (st) st->lightReads++;
    float pickPdf = L[15] / s->totalLightPower;
    if (pickPdf <= 0.0f) return V(0, 0, 0);
    float r1 = rco_rand(seed);
    float r2 = rco_rand(seed);
    if (r1 + r2 > 1.0f) { r1 = 1.0f - r1; r2 = 1.0f - r2; }
    v3 lv0 = ld3(L), lv1 = ld3(L + 4), lv2 = ld3(L + 8);
    v3 lp = add(add(lv0, muls(sub(lv1, lv0), r1)), muls(sub(lv2, lv0), r2));
    v3 tl = sub(lp, hp);
    float d2 = dot3(tl, tl);
    float d = sqrtf(d2);
    v3 sd = divs(tl, d);
    float cosHit = dot3(n, sd);
    if (cosHit <= 0.0f) return V(0, 0, 0);
    float cosLight = fabsf(dot3(ld3(L + 16), sd));
    if (cosLight <= 0.0f) return V(0, 0, 0);
    float d2s = fmaxf(d2, L[3]);
    v3 so = add(hp, muls(n, 2e-3f));
    if (!shadow(s, so, sd, d - 4e-3f, st)) return V(0, 0, 0);
    float pdfSA = d2s / (L[3] * cosLight);
    float pdf = pickPdf * pdfSA;
    if (pdf <= 0.0f) return V(0, 0, 0);
    v3 em = ld3(L + 12);
    v3 c = V((((albedo.x / RC_PI_F) * em.x) * cosHit) / pdf,
             (((albedo.y / RC_PI_F) * em.y) * cosHit) / pdf,
             (((albedo.z / RC_PI_F) * em.z) * cosHit) / pdf);
    float emLum = (0.299f * em.x + 0.587f * em.y) + 0.114f * em.z;
    float cl = emLum * p->neeFireflyClamp;
    return V(fminf(c.x, cl), fminf(c.y, cl), fminf(c.z, cl));
}

/* ---- ray generation (pathtrace.comp:401-439) ---------------------------- */
static void gen_ray(const rc_camera *c, int px, int py, uint32_t *seed, float fOff, v3 *ro, v3 *rd) {
    float aspect = (float)c->width / (float)c->height;
    float jx = rco_rand(seed) - 0.5f;
    float jy = rco_rand(seed) - 0.5f;
    float ux = (((float)px + 0.5f) + jx) / (float)c->width;
    float uy = (((float)py + 0.5f) + jy) / (float)c->height;
    float sx = ux * 2.0f - 1.0f, sy = uy * 2.0f - 1.0f;
    v3 F = ld3(c->camForward), R = ld3(c->camRight), U = ld3(c->camUp), P = ld3(c->camPos);
    float a = (sx * aspect) * c->fovTanHalf;
    float b = sy * c->fovTanHalf;
    v3 pin = normalize3(add(add(F, muls(R, a)), muls(U, b)));
    float lx = 0.0f, ly = 0.0f;
    if (c->lensRadius > 0.0f) sample_ngon(seed, c->lensRadius, c->apertureBlades, &lx, &ly);
    float nr2 = 0.0f;
    if (c->lensRadius > 1e-6f)
        nr2 = rc_clampf((lx * lx + ly * ly) / (c->lensRadius * c->lensRadius), 0.0f, 1.0f);
    float ef = (c->focalDist + fOff) + c->saStrength * nr2;
    ef = fmaxf(ef, 0.01f);
    float tf = ef / dot3(pin, F);
    v3 fp = add(P, muls(pin, tf));
    v3 o = add(add(P, muls(R, lx)), muls(U, ly));
    *ro = o;
    *rd = normalize3(sub(fp, o));
}

/* ---- specular scattering of traceRadiance -------------------------------- */
/* Glass (pathtrace.comp:504-523): Schlick F on r0 = ((1-ior)/(1+ior))^2 at the
 * facing normal n, one rand picks reflection (u < F) or refraction (eta = 1/ior
 * entering, ior leaving; total internal reflection reflects); the new origin is
 * 2e-3 along the new direction; throughput *= albedo. geomN is the hit normal
 * before facing, which decides entering. *Fout (may be NULL) receives F. */
static void glass_step(float ior, v3 albedo, v3 pos, v3 geomN, v3 n, uint32_t *seed, v3 *ro, v3 *rd, v3 *thr,
                       float *Fout) {
    v3 d = *rd;
    int entering = dot3(d, geomN) < 0.0f;
    float eta = entering ? (1.0f / ior) : ior;
    float cosI = fabsf(dot3(d, n));
    float r0 = (1.0f - ior) / (1.0f + ior);
    r0 = r0 * r0;
    float F = r0 + (1.0f - r0) * rc_powf(rc_clampf(1.0f - cosI, 0.0f, 1.0f), 5.0f);
    if (Fout) *Fout = F;
    if (rco_rand(seed) < F) {
        d = reflect3(d, n);
        *ro = add(pos, muls(d, 2e-3f));
    } else {
        v3 rf = refract3(d, n, eta);
        if (dot3(rf, rf) < 0.001f) rf = reflect3(d, n);
        d = normalize3(rf);
        *ro = add(pos, muls(d, 2e-3f));
    }
    *rd = d;
    *thr = mul(*thr, albedo);
}

/* Metal (pathtrace.comp:525-566): alpha = max(roughness, 1e-3) used as the GGX
 * alpha, a half vector from the visible normals in the frame (T, n, B), the
 * mirror direction about n when the reflection falls below the surface,
 * throughput *= F * G2 / G1 with Schlick F on the albedo; new origin 2e-3 along n. */
static void metal_step(float rough, v3 albedo, v3 pos, v3 n, uint32_t *seed, v3 *ro, v3 *rd, v3 *thr) {
    v3 wo = muls(*rd, -1.0f);
    float alpha = fmaxf(rough, 1e-3f);
    v3 up = fabsf(n.y) < 0.9999f ? V(0.0f, 1.0f, 0.0f) : V(1.0f, 0.0f, 0.0f);
    v3 T = normalize3(cross3(up, n));
    v3 B = cross3(n, T);
    v3 wl = V(dot3(wo, T), dot3(wo, n), dot3(wo, B));
    float u1 = rco_rand(seed);
    float u2 = rco_rand(seed);
    v3 hl = sample_vndf(wl, alpha, u1, u2);
    v3 hv = normalize3(add(add(muls(T, hl.x), muls(n, hl.y)), muls(B, hl.z)));
    v3 mwo = muls(wo, -1.0f);
    v3 d = reflect3(mwo, hv);
    if (dot3(d, n) <= 0.0f) d = reflect3(mwo, n);
    float nl = fmaxf(dot3(n, d), 1e-4f);
    float nv = fmaxf(dot3(n, wo), 1e-4f);
    float hv_ = fmaxf(dot3(hv, wo), 1e-4f);
    float pw = rc_powf(1.0f - hv_, 5.0f);
    v3 F = V(albedo.x + (1.0f - albedo.x) * pw, albedo.y + (1.0f - albedo.y) * pw,
             albedo.z + (1.0f - albedo.z) * pw);
    float g1 = smith_g1(nv, alpha);
    float g2 = smith_g2(nl, nv, alpha);
    float g = g2 / fmaxf(g1, 1e-7f);
    *thr = mul(*thr, muls(F, g));
    *ro = add(pos, muls(n, 2e-3f));
    *rd = d;
}

/* ---- traceRadiance (pathtrace.comp:454-585) ----------------------------- */
static v3 trace_radiance(const rco_scene *s, const rc_pt_params *p, v3 ro, v3 rd, uint32_t *seed,
                         v3 *firstAlbedo, int *firstTri, rco_stats *st) {
    *firstAlbedo = V(1.0f, 1.0f, 1.0f);
    *firstTri = -1;
    v3 rad = V(0, 0, 0), thr = V(1.0f, 1.0f, 1.0f);
    int lastDiffuse = 0;
    for (int bounce = 0; bounce < p->maxBounces; ++bounce) {
        hitrec h = traverse(s, ro, rd, st);
        if (!h.valid) break;
        const float *M = s->mats + 16 * (size_t)h.mat;
        if (st) st->shadeHits++;
        v3 albedo = ld3(M);
        float metallic = M[3], estr = M[7], rough = M[8], ior = M[9], trans = M[10];
        v3 emis = ld3(M + 4);
        v3 n = dot3(rd, h.normal) < 0.0f ? h.normal : muls(h.normal, -1.0f);
        if (bounce == 0) {
            *firstAlbedo = (trans > 0.5f) ? V(1.0f, 1.0f, 1.0f) : albedo;
            *firstTri = h.tri;
        }
        if (!lastDiffuse) rad = add(rad, muls(mul(thr, emis), estr));
        if (bounce >= 2) {
            float surv = fmaxf(albedo.x, fmaxf(albedo.y, albedo.z));
            if (metallic > 0.5f || trans > 0.5f) surv = 0.95f;
            surv = rc_clampf(surv, 0.01f, 0.95f);
            if (rco_rand(seed) > surv) break;
            thr = divs(thr, surv);
        }
        if (trans > 0.5f) {
            glass_step(ior, albedo, h.pos, h.normal, n, seed, &ro, &rd, &thr, NULL);
            lastDiffuse = 0;
        } else if (metallic > 0.5f) {
            metal_step(rough, albedo, h.pos, n, seed, &ro, &rd, &thr);
            lastDiffuse = 0;
        } else {
            if (p->neeEnabled != 0) rad = add(rad, mul(thr, direct_light(s, p, h.pos, n, albedo, seed, st)));
            float u1 = rco_rand(seed);
            float u2 = rco_rand(seed);
            rd = sample_cosine(u1, u2, n);
            thr = mul(thr, albedo);
            ro = add(h.pos, muls(n, 1e-3f));
            lastDiffuse = 1;
        }
    }
    return rad;
}

/* ---- main (pathtrace.comp:601-655) -------------------------------------- */
static void pixel(const rco_scene *s, const rc_camera *c, const rc_pt_params *p, int frame,
                  int x, int y, rco_frame *o, rco_stats *st) {
    size_t pi = (size_t)y * (size_t)c->width + (size_t)x;
    v3 rad, alb;
    int tri;
    int lca = (fabsf(c->lcaDelta[0]) + fabsf(c->lcaDelta[2])) > 1e-4f;
    v3 ro, rd;
    if (lca) {
        uint32_t sr = rco_init_seed(x, y, frame * 3 + 0);
        uint32_t sg = rco_init_seed(x, y, frame * 3 + 1);
        uint32_t sb = rco_init_seed(x, y, frame * 3 + 2);
        v3 aR, aG, aB;
        int tR, tG, tB;
        gen_ray(c, x, y, &sr, c->lcaDelta[0], &ro, &rd);
        float r = trace_radiance(s, p, ro, rd, &sr, &aR, &tR, st).x;
        gen_ray(c, x, y, &sg, c->lcaDelta[1], &ro, &rd);
        float g = trace_radiance(s, p, ro, rd, &sg, &aG, &tG, st).y;
        gen_ray(c, x, y, &sb, c->lcaDelta[2], &ro, &rd);
        float b = trace_radiance(s, p, ro, rd, &sb, &aB, &tB, st).z;
        rad = V(r, g, b);
        alb = aG;
        tri = tG;
        if (st) st->paths += 3;
    } else {
        uint32_t sd = rco_init_seed(x, y, frame);
        gen_ray(c, x, y, &sd, 0.0f, &ro, &rd);
        rad = trace_radiance(s, p, ro, rd, &sd, &alb, &tri, st);
        if (st) st->paths += 1;
    }
    if (isnan(rad.x) || isnan(rad.y) || isnan(rad.z) || isinf(rad.x) || isinf(rad.y) || isinf(rad.z))
        rad = V(0, 0, 0);
    float luma = (rad.x * 0.2126f + rad.y * 0.7152f) + rad.z * 0.0722f;
    float l2 = luma * luma;
    if (o->primId) o->primId[pi] = tri;
    if (frame == 0) {
        if (o->accum) { o->accum[4 * pi] = rad.x; o->accum[4 * pi + 1] = rad.y; o->accum[4 * pi + 2] = rad.z; o->accum[4 * pi + 3] = 1.0f; }
        if (o->accum64) { o->accum64[4 * pi] = rad.x; o->accum64[4 * pi + 1] = rad.y; o->accum64[4 * pi + 2] = rad.z; o->accum64[4 * pi + 3] = 1.0; }
        if (o->variance) o->variance[pi] = l2;
        uint32_t gs = rco_init_seed(x, y, 99991);
        gen_ray(c, x, y, &gs, 0.0f, &ro, &rd);
        hitrec h = traverse(s, ro, rd, st);
        float g[4];
        if (!h.valid) { g[0] = 0.0f; g[1] = 1.0f; g[2] = 0.0f; g[3] = 1e4f; }
        else {
            v3 n = dot3(rd, h.normal) < 0.0f ? h.normal : muls(h.normal, -1.0f);
            g[0] = n.x; g[1] = n.y; g[2] = n.z; g[3] = h.t;
        }
        if (o->gbuf) for (int k = 0; k < 4; ++k) o->gbuf[4 * pi + k] = rc_f32_to_f16(g[k]);
        if (o->gbufId) o->gbufId[pi] = h.valid ? h.tri : -1;
        if (o->galbedo) {
            o->galbedo[4 * pi] = rc_f32_to_f16(alb.x);
            o->galbedo[4 * pi + 1] = rc_f32_to_f16(alb.y);
            o->galbedo[4 * pi + 2] = rc_f32_to_f16(alb.z);
            o->galbedo[4 * pi + 3] = rc_f32_to_f16(1.0f);
        }
    } else {
        if (o->accum) {
            o->accum[4 * pi] += rad.x; o->accum[4 * pi + 1] += rad.y;
            o->accum[4 * pi + 2] += rad.z; o->accum[4 * pi + 3] += 1.0f;
        }
        if (o->accum64) {
            o->accum64[4 * pi] += rad.x; o->accum64[4 * pi + 1] += rad.y;
            o->accum64[4 * pi + 2] += rad.z; o->accum64[4 * pi + 3] += 1.0;
        }
        if (o->variance) o->variance[pi] += l2;
    }
}

static void stats_add(rco_stats *d, const rco_stats *a) {
    const uint64_t *pa = (const uint64_t *)a;
    uint64_t *pd = (uint64_t *)d;
    for (size_t i = 0; i < sizeof(rco_stats) / sizeof(uint64_t); ++i) pd[i] += pa[i];
}

int32_t rco_render_pass(const rco_scene *s, const rc_camera *c, const rc_pt_params *p,
                        int32_t frame, int32_t stride, rco_frame *o, rco_stats *stats) {
    if (!s || !c || !p || !o || stride < 1 || c->width <= 0 || c->height <= 0) return -1;
    if (s->nNodes <= 0 || s->nTris <= 0) return -1;
    int H = c->height, W = c->width;
#pragma omp parallel
    {
        rco_stats local;
        memset(&local, 0, sizeof(local));
#pragma omp for schedule(dynamic, 1)
        for (int y = 0; y < H; y += stride)
            for (int x = 0; x < W; x += stride) pixel(s, c, p, frame, x, y, o, stats ? &local : NULL);
        if (stats) {
#pragma omp critical
            stats_add(stats, &local);
        }
    }
    return 0;
}

int32_t rco_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Element-wise rc_fmath routines (for the accuracy tests of rc_fmath.h):
 * fn 0 sin, 1 cos, 2 exp, 3 log, 4 pow(x, y), 5 f32->f16->f32 round trip. */
int32_t rco_fmath(int32_t fn, const float *x, const float *y, float *out, int32_t n) {
    for (int32_t i = 0; i < n; ++i) {
        switch (fn) {
            case 0: out[i] = rc_sinf(x[i]); break;
            case 1: out[i] = rc_cosf(x[i]); break;
            case 2: out[i] = rc_expf(x[i]); break;
            case 3: out[i] = rc_logf(x[i]); break;
            case 4: out[i] = rc_powf(x[i], y[i]); break;
            case 5: out[i] = rc_f16_to_f32(rc_f32_to_f16(x[i])); break;
            default: return -1;
        }
    }
    return 0;
}

/* ---- sub-step probes (for the oracle's own closed-form pins) -------------- */
int32_t rco_gen_ray(const rc_camera *cam, int32_t px, int32_t py, uint32_t *seed, float focalOffset, float *ro,
                    float *rd) {
    if (!cam || !seed || !ro || !rd || cam->width <= 0 || cam->height <= 0) return -1;
    v3 o, d;
    gen_ray(cam, px, py, seed, focalOffset, &o, &d);
    ro[0] = o.x; ro[1] = o.y; ro[2] = o.z;
    rd[0] = d.x; rd[1] = d.y; rd[2] = d.z;
    return 0;
}

void rco_sample_ngon(uint32_t *seed, float radius, int32_t blades, float *xy) {
    sample_ngon(seed, radius, blades, &xy[0], &xy[1]);
}

void rco_sample_vndf(const float *wo, float alpha, float u1, float u2, float *h) {
    v3 r = sample_vndf(ld3(wo), alpha, u1, u2);
    h[0] = r.x; h[1] = r.y; h[2] = r.z;
}

float rco_smith_g1(float nv, float alpha) { return smith_g1(nv, alpha); }
float rco_smith_g2(float nl, float nv, float alpha) { return smith_g2(nl, nv, alpha); }

void rco_sample_cosine(float u1, float u2, const float *n, float *dir) {
    v3 r = sample_cosine(u1, u2, ld3(n));
    dir[0] = r.x; dir[1] = r.y; dir[2] = r.z;
}

int32_t rco_scatter(const float *mat, const float *rd, const float *geomNormal, const float *pos, uint32_t *seed,
                    float *roOut, float *rdOut, float *thrOut, float *fresnelOut) {
    if (!mat || !rd || !geomNormal || !pos || !seed || !roOut || !rdOut || !thrOut) return -1;
    v3 d = ld3(rd), gn = ld3(geomNormal), P = ld3(pos);
    v3 n = dot3(d, gn) < 0.0f ? gn : muls(gn, -1.0f);
    v3 albedo = ld3(mat), thr = V(1.0f, 1.0f, 1.0f), o = V(0, 0, 0);
    float F = -1.0f;
    if (mat[10] > 0.5f) glass_step(mat[9], albedo, P, gn, n, seed, &o, &d, &thr, &F);
    else if (mat[3] > 0.5f) metal_step(mat[8], albedo, P, n, seed, &o, &d, &thr);
    else return -1;
    roOut[0] = o.x; roOut[1] = o.y; roOut[2] = o.z;
    rdOut[0] = d.x; rdOut[1] = d.y; rdOut[2] = d.z;
    thrOut[0] = thr.x; thrOut[1] = thr.y; thrOut[2] = thr.z;
    if (fresnelOut) *fresnelOut = F;
    return 0;
}
