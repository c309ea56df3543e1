/*
 * rc_oracle_sppm.c — CPU oracle of SPPM: photon emission/trace/insert
 * (photon_trace.comp), radiance gather (photon_gather.comp) and the radius
 * schedule (SPPMManager.updateRadius).
 *
 * TEST INFRASTRUCTURE ONLY (see rc_oracle.h). Shader paths are relative to
 * src/main/resources/shaders/, Java paths to src/main/java/com/retrocam/.
 *
 * The photon shaders walk the BVH without ordering children: pop a node, test
 * its own box, push right then left (photon_trace.comp:101-124,
 * photon_gather.comp:85-118), with the 62-entry cap.
 */
#include "rc_oracle.h"
#include "rc_fmath.h"

#include <math.h>
#include <string.h>

typedef struct { float x, y, z; } v3;
static inline v3 V(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 muls(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
static inline float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross3(v3 a, v3 b) {
    return V(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
static inline v3 normalize3(v3 a) { float inv = 1.0f / sqrtf(dot3(a, a)); return muls(a, inv); }
static inline v3 reflect3(v3 i, v3 n) { float d = 2.0f * dot3(n, i); return sub(i, muls(n, d)); }
static inline v3 refract3(v3 i, v3 n, float eta) {
    float d = dot3(n, i);
    float k = 1.0f - (eta * eta) * (1.0f - d * d);
    if (k < 0.0f) return V(0.0f, 0.0f, 0.0f);
    float c = eta * d + sqrtf(k);
    return sub(muls(i, eta), muls(n, c));
}
static inline v3 ld3(const float *p) { return V(p[0], p[1], p[2]); }
static inline int32_t ldi(const float *p) { int32_t i; memcpy(&i, p, 4); return i; }

static inline uint32_t pcg(uint32_t *st) {
    uint32_t s = *st * 747796405u + 2891336453u;
    *st = s;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}
static inline float rnd(uint32_t *s) { return (float)pcg(s) / 4294967296.0f; }

/* initSeed (photon_trace.comp:68-72) */
uint32_t rco_photon_seed(uint32_t idx, int32_t seed) {
    uint32_t s = idx * 1664525u + (uint32_t)seed * 1013904223u + 2531011u;
    pcg(&s);
    pcg(&s);
    return s;
}

static int aabb_hit(v3 ro, v3 inv, const float *bmin, const float *bmax, float *tmin) {
    float t0x = (bmin[0] - ro.x) * inv.x, t0y = (bmin[1] - ro.y) * inv.y, t0z = (bmin[2] - ro.z) * inv.z;
    float t1x = (bmax[0] - ro.x) * inv.x, t1y = (bmax[1] - ro.y) * inv.y, t1z = (bmax[2] - ro.z) * inv.z;
    float tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z));
    float tx = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
    *tmin = tn;
    return tx >= fmaxf(tn, 0.0f);
}

static int tri_hit(const rco_scene *s, v3 ro, v3 rd, int idx, float *t, float *bu, float *bv) {
    const float *T = s->tris + 32 * (size_t)idx;
    v3 v0 = ld3(T), e1 = sub(ld3(T + 4), v0), e2 = sub(ld3(T + 8), v0);
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

typedef struct { int valid; float t; v3 pos, normal; int mat, tri; } hitrec;

/* Unordered closest hit (photon_trace.comp:101-124). */
static hitrec walk(const rco_scene *s, v3 ro, v3 rd) {
    hitrec rec;
    memset(&rec, 0, sizeof(rec));
    rec.t = 1e30f;
    rec.tri = -1;
    v3 inv = V(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    int stack[64], top = 0;
    stack[top++] = 0;
    while (top > 0) {
        const float *N = s->nodes + 12 * (size_t)stack[--top];
        float at;
        if (!aabb_hit(ro, inv, N, N + 4, &at)) continue;
        if (at > rec.t) continue;
        int ps = ldi(N + 10), pc = ldi(N + 11);
        if (pc > 0) {
            for (int i = ps; i < ps + pc; ++i) {
                float t, bu, bv;
                if (tri_hit(s, ro, rd, i, &t, &bu, &bv) && t < rec.t) {
                    const float *T = s->tris + 32 * (size_t)i;
                    rec.t = t;
                    rec.pos = add(ro, muls(rd, t));
                    float w = (1.0f - bu) - bv;
                    rec.normal = normalize3(add(add(muls(ld3(T + 12), w), muls(ld3(T + 16), bu)), muls(ld3(T + 20), bv)));
                    rec.mat = ldi(T + 30);
                    rec.valid = 1;
                    rec.tri = i;
                }
            }
        } else if (top < 62) {
            stack[top++] = ldi(N + 9);
            stack[top++] = ldi(N + 8);
        }
    }
    return rec;
}

int32_t rco_walk_closest(const rco_scene *s, const float *ro, const float *rd, float *tOut) {
    hitrec h = walk(s, ld3(ro), ld3(rd));
    if (tOut) *tOut = h.t;
    return h.valid ? h.tri : -1;
}

/* cellHash (photon_trace.comp:140-143): int32 products wrap. */
uint32_t rco_cell_hash(int32_t cx, int32_t cy, int32_t cz, int32_t hashSize) {
    uint32_t h = ((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u) ^ ((uint32_t)cz * 83492791u);
    return h & (uint32_t)(hashSize - 1);
}

static inline void cell_of(v3 p, float r, int32_t *c) {
    c[0] = (int32_t)floorf(p.x / r);
    c[1] = (int32_t)floorf(p.y / r);
    c[2] = (int32_t)floorf(p.z / r);
}

/* One photon (photon_trace.comp:168-251). Returns 1 and fills rec (12 floats)
 * when the photon is stored. */
static int trace_photon(const rco_scene *s, uint32_t gid, int32_t photonCount, int32_t seed, float *rec) {
    uint32_t st = rco_photon_seed(gid, seed);
    float u = rnd(&st);
    int sel = s->nLights - 1;
    for (int i = 0; i < s->nLights; ++i)
        if (u < s->lights[20 * (size_t)i + 19]) { sel = i; break; }
    const float *L = s->lights + 20 * (size_t)sel;
    float r1 = rnd(&st);
    float r2 = rnd(&st);
    if (r1 + r2 > 1.0f) { r1 = 1.0f - r1; r2 = 1.0f - r2; }
    v3 lv0 = ld3(L);
    v3 sp = add(add(lv0, muls(sub(ld3(L + 4), lv0), r1)), muls(sub(ld3(L + 8), lv0), r2));
    v3 en = normalize3(ld3(L + 16));
    /* cosineHemisphere (:131-137) */
    float ux = rnd(&st);
    float uy = rnd(&st);
    float rr = sqrtf(ux), phi = 6.28318530718f * uy;
    v3 loc = V(rr * rc_cosf(phi), rr * rc_sinf(phi), sqrtf(fmaxf(0.0f, 1.0f - ux)));
    v3 up = fabsf(en.y) < 0.9999f ? V(0.0f, 1.0f, 0.0f) : V(1.0f, 0.0f, 0.0f);
    v3 tt = normalize3(cross3(up, en));
    v3 bb = cross3(en, tt);
    v3 dir = normalize3(add(add(muls(tt, loc.x), muls(bb, loc.y)), muls(en, loc.z)));
    v3 em = ld3(L + 12);
    float lum = (0.299f * em.x + 0.587f * em.y) + 0.114f * em.z;
    float pp = s->totalLightPower / (float)photonCount;
    v3 power = lum > 1e-6f ? V((em.x / lum) * pp, (em.y / lum) * pp, (em.z / lum) * pp) : V(pp, pp, pp);
    v3 ro = add(sp, muls(en, 1e-3f));
    v3 rd = dir;
    int spec = 0;
    for (int bounce = 0; bounce < 8; ++bounce) {
        hitrec h = walk(s, ro, rd);
        if (!h.valid) break;
        const float *M = s->mats + 16 * (size_t)h.mat;
        v3 n = dot3(rd, h.normal) < 0.0f ? h.normal : muls(h.normal, -1.0f);
        if (M[10] > 0.5f) {
            int entering = dot3(rd, h.normal) < 0.0f;
            float ior = M[9];
            float eta = entering ? (1.0f / ior) : ior;
            float cosI = fabsf(dot3(rd, n));
            float r0 = (1.0f - ior) / (1.0f + ior);
            r0 = r0 * r0;
            float F = r0 + (1.0f - r0) * rc_powf(rc_clampf(1.0f - cosI, 0.0f, 1.0f), 5.0f);
            if (rnd(&st) < F) {
                rd = reflect3(rd, n);
            } else {
                v3 rf = refract3(rd, n, eta);
                if (dot3(rf, rf) < 0.001f) rf = reflect3(rd, n);
                rd = normalize3(rf);
            }
            ro = add(h.pos, muls(rd, 2e-3f));
            spec = 1;
            continue;
        }
        if (M[3] > 0.5f) {
            power = mul(power, ld3(M));
            rd = reflect3(rd, n);
            ro = add(h.pos, muls(n, 2e-3f));
            spec = 1;
            continue;
        }
        if (spec) {
            rec[0] = h.pos.x; rec[1] = h.pos.y; rec[2] = h.pos.z; rec[3] = 0.0f;
            rec[4] = power.x; rec[5] = power.y; rec[6] = power.z; rec[7] = 0.0f;
            rec[8] = rd.x; rec[9] = rd.y; rec[10] = rd.z; rec[11] = 0.0f;
            return 1;
        }
        break;
    }
    return 0;
}

int32_t rco_photon_trace(const rco_scene *s, int32_t photonCount, float radius, int32_t seed, int32_t hashSize,
                         float *photons, int32_t *heads, int32_t *next, int32_t *storedOut) {
    if (!s || photonCount <= 0 || hashSize <= 0 || (hashSize & (hashSize - 1)) || !photons || !heads || !next)
        return -1;
    for (int32_t i = 0; i < hashSize; ++i) heads[i] = -1;  /* PhotonMap.clear (PhotonMap.java:84-106) */
    int32_t count = 0;
    if (s->nLights > 0) {
        for (int32_t g = 0; g < photonCount; ++g) {
            float rec[12];
            if (!trace_photon(s, (uint32_t)g, photonCount, seed, rec)) continue;
            int32_t idx = count++;  /* atomicAdd order: photon index order on the CPU */
            if (idx >= photonCount) continue;
            int32_t c[3];
            cell_of(V(rec[0], rec[1], rec[2]), radius, c);
            uint32_t h = rco_cell_hash(c[0], c[1], c[2], hashSize);
            next[idx] = heads[h];
            heads[h] = idx;
            memcpy(photons + 12 * (size_t)idx, rec, sizeof(rec));
        }
    }
    if (storedOut) *storedOut = count;
    return 0;
}

/* photon_gather.comp:128-221 for every pixel; adds (L, 0) to accum (float4)
 * and, if accum64 is given, to the double sums too. */
int32_t rco_photon_gather(const rco_scene *s, const rc_camera *c, float radius, int32_t hashSize,
                          const float *photons, const int32_t *heads, const int32_t *next, float *accum,
                          double *accum64) {
    if (!s || !c || !photons || !heads || !next || !accum || hashSize <= 0) return -1;
    int W = c->width, H = c->height;
    float aspect = (float)W / (float)H;
    v3 F = ld3(c->camForward), R = ld3(c->camRight), U = ld3(c->camUp), P = ld3(c->camPos);
    float r2 = radius * radius;
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            float nx = (((float)x + 0.5f) / (float)W) * 2.0f - 1.0f;
            float ny = (((float)y + 0.5f) / (float)H) * 2.0f - 1.0f;
            v3 rd = normalize3(add(add(F, muls(R, (nx * aspect) * c->fovTanHalf)), muls(U, ny * c->fovTanHalf)));
            v3 ro = P;
            int found = 0;
            v3 hp = V(0, 0, 0);
            int hm = 0;
            for (int sIt = 0; sIt <= 6; ++sIt) {
                hitrec h = walk(s, ro, rd);
                if (!h.valid) break;
                const float *M = s->mats + 16 * (size_t)h.mat;
                if (M[3] <= 0.5f && M[10] <= 0.5f) { found = 1; hp = h.pos; hm = h.mat; break; }
                v3 n = dot3(rd, h.normal) < 0.0f ? h.normal : muls(h.normal, -1.0f);
                if (M[10] > 0.5f) {
                    int entering = dot3(rd, h.normal) < 0.0f;
                    float eta = entering ? (1.0f / M[9]) : M[9];
                    v3 rf = refract3(rd, n, eta);
                    if (dot3(rf, rf) < 0.001f) rf = reflect3(rd, n);
                    rd = normalize3(rf);
                } else {
                    rd = reflect3(rd, n);
                }
                ro = add(h.pos, muls(rd, 2e-3f));
            }
            if (!found) continue;
            int32_t cc[3];
            cell_of(hp, radius, cc);
            v3 sum = V(0, 0, 0);
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        int32_t idx = heads[rco_cell_hash(cc[0] + dx, cc[1] + dy, cc[2] + dz, hashSize)];
                        int safety = 0;
                        while (idx >= 0 && safety < 512) {
                            const float *ph = photons + 12 * (size_t)idx;
                            v3 d = sub(hp, ld3(ph));
                            float d2 = dot3(d, d);
                            if (d2 <= r2) {
                                float w = 1.0f - d2 / r2;
                                sum = add(sum, muls(ld3(ph + 4), w));
                            }
                            idx = next[idx];
                            ++safety;
                        }
                    }
            const float *M = s->mats + 16 * (size_t)hm;
            float den = 3.14159265358979f * r2;
            v3 L = V((M[0] * sum.x) / den, (M[1] * sum.y) / den, (M[2] * sum.z) / den);
            if (isnan(L.x) || isnan(L.y) || isnan(L.z) || isinf(L.x) || isinf(L.y) || isinf(L.z)) L = V(0, 0, 0);
            size_t pi = (size_t)y * (size_t)W + (size_t)x;
            accum[4 * pi] += L.x;
            accum[4 * pi + 1] += L.y;
            accum[4 * pi + 2] += L.z;
            if (accum64) {
                accum64[4 * pi] += (double)L.x;
                accum64[4 * pi + 1] += (double)L.y;
                accum64[4 * pi + 2] += (double)L.z;
            }
        }
    }
    return 0;
}

/* SPPMManager.updateRadius (SPPMManager.java:170-173) with Java semantics:
 * r *= (float) Math.sqrt((iteration + alpha) / (iteration + 1.0)). */
float rco_sppm_next_radius(float r, int32_t iteration, float alpha) {
    float num = (float)iteration + alpha;
    double q = (double)num / ((double)iteration + 1.0);
    return r * (float)sqrt(q);
}
