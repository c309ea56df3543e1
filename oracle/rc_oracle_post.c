/*
 * rc_oracle_post.c — CPU oracle of the post chain: p00 normalize, the a-trous
 * denoiser, the VHS passes p01..p16 and the display encode.
 *
 * TEST INFRASTRUCTURE ONLY (see rc_oracle.h). Shader paths are relative to
 * src/main/resources/shaders/, Java paths to src/main/java/com/retrocam/.
 *
 * Image convention: RGBA32F, row 0 = bottom, pixel (x, y) at y*W + x. A pass
 * writes (rgb, 1.0) like fragColor = vec4(rgb, 1.0).
 *
 * Texture sampling (fullscreen.vert:9-16 gives fragment-centre UVs
 * v_uv = ((x + 0.5)/W, (y + 0.5)/H)):
 *  - samples at v_uv plus whole-texel offsets land on texel centres, where
 *    GL_LINEAR returns the texel itself: restated as a texel fetch, clamped to
 *    the edge for the RGBA32F FBO textures (gl/Framebuffer.java:74-77) and
 *    wrapped for the Renderer's G-buffer textures (default GL_REPEAT,
 *    core/Renderer.java:163-183);
 *  - samples at fractional positions (p10, p12, p13, p14, p16) are bilinear
 *    with CLAMP_TO_EDGE: x = u*W - 0.5, weights from x - floor(x), float32.
 */
#include "rc_oracle.h"
#include "rc_fmath.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { float r, g, b; } c3;
typedef struct { const float *p; int w, h; } img;

static inline c3 C3(float r, float g, float b) { c3 c = {r, g, b}; return c; }
static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline c3 fetch(img s, int x, int y) {
    x = clampi(x, 0, s.w - 1);
    y = clampi(y, 0, s.h - 1);
    const float *q = s.p + 4 * ((size_t)y * (size_t)s.w + (size_t)x);
    return C3(q[0], q[1], q[2]);
}
static inline int wrapi(int v, int n) { int m = v % n; return m < 0 ? m + n : m; }

static c3 bilinear(img s, float u, float v) {
    float x = u * (float)s.w - 0.5f, y = v * (float)s.h - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    float ax = x - fx, ay = y - fy;
    int x0 = (int)fx, y0 = (int)fy;
    c3 t00 = fetch(s, x0, y0), t10 = fetch(s, x0 + 1, y0);
    c3 t01 = fetch(s, x0, y0 + 1), t11 = fetch(s, x0 + 1, y0 + 1);
    float bx = 1.0f - ax, by = 1.0f - ay;
    c3 r0 = C3(t00.r * bx + t10.r * ax, t00.g * bx + t10.g * ax, t00.b * bx + t10.b * ax);
    c3 r1 = C3(t01.r * bx + t11.r * ax, t01.g * bx + t11.g * ax, t01.b * bx + t11.b * ax);
    return C3(r0.r * by + r1.r * ay, r0.g * by + r1.g * ay, r0.b * by + r1.b * ay);
}

static inline float vu(int x, int w) { return ((float)x + 0.5f) / (float)w; }
static inline float clamp01(float v) { return rc_clampf(v, 0.0f, 1.0f); }
static inline float smoothstepf(float e0, float e1, float x) {
    float t = rc_clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return (t * t) * (3.0f - 2.0f * t);
}
static inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
static inline float fractf(float x) { return x - floorf(x); }
static inline float signf(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
static inline float luma709(c3 c) { return (c.r * 0.2126f + c.g * 0.7152f) + c.b * 0.0722f; }
static inline float luma601(c3 c) { return (0.299f * c.r + 0.587f * c.g) + 0.114f * c.b; }

/* BT.601 YCbCr with 5-decimal coefficients (p02/p07/p09/p13/p14/p15). */
static inline c3 ycc5(c3 c) {
    return C3((0.29900f * c.r + 0.58700f * c.g) + 0.11400f * c.b,
              ((-0.16874f * c.r - 0.33126f * c.g) + 0.50000f * c.b) + 0.5f,
              ((0.50000f * c.r - 0.41869f * c.g) - 0.08131f * c.b) + 0.5f);
}
static inline c3 rgb5(c3 y) {
    float cb = y.g - 0.5f, cr = y.b - 0.5f;
    return C3(y.r + 1.40200f * cr, (y.r - 0.34414f * cb) - 0.71414f * cr, y.r + 1.77200f * cb);
}
/* 3-decimal variant (p03/p04). */
static inline c3 ycc3(c3 c) {
    return C3((0.299f * c.r + 0.587f * c.g) + 0.114f * c.b,
              ((-0.169f * c.r - 0.331f * c.g) + 0.500f * c.b) + 0.5f,
              ((0.500f * c.r - 0.419f * c.g) - 0.081f * c.b) + 0.5f);
}
static inline c3 rgb3(c3 y) {
    float cb = y.g - 0.5f, cr = y.b - 0.5f;
    return C3(y.r + 1.402f * cr, (y.r - 0.344f * cb) - 0.714f * cr, y.r + 1.772f * cb);
}

static inline uint32_t pcg_hash(uint32_t v) {
    v = v * 747796405u + 2891336453u;
    uint32_t w = ((v >> ((v >> 28u) + 4u)) ^ v) * 277803737u;
    return (w >> 22u) ^ w;
}
/* p12's one-round integer hash (p12_dropout.frag:32-37). */
static inline uint32_t uhash1(uint32_t x) { x ^= x >> 16u; x *= 0x45d9f3bu; x ^= x >> 16u; return x; }
/* p13/p14's two-round integer hash (p13_headswitch.frag:40-45). */
static inline uint32_t uhash2(uint32_t x) {
    x = ((x >> 16u) ^ x) * 0x45d9f3bu;
    x = ((x >> 16u) ^ x) * 0x45d9f3bu;
    return (x >> 16u) ^ x;
}
static inline float fhash1(uint32_t a, uint32_t b) { return (float)uhash1(a ^ uhash1(b) * 2654435761u) / 4294967295.0f; }
static inline float fhash2(uint32_t a, uint32_t b) { return (float)uhash2(a ^ uhash2(b) * 2654435761u) / 4294967295.0f; }

/* ---- p00 (post/p00_normalize.frag:22-27) --------------------------------- */
static c3 p00(img s, int x, int y, int total, float exposure) {
    float n = (float)(total > 1 ? total : 1);
    c3 c = fetch(s, x, y);
    return C3((c.r / n) * exposure, (c.g / n) * exposure, (c.b / n) * exposure);
}

/* ---- a-trous pass (post/atrous_denoise.frag:50-122) ----------------------- */
typedef struct {
    const uint16_t *gbuf, *galbedo;
    const float *variance;
    int total;
    float exposure;
} gbufs;

static inline void gb_fetch(const gbufs *g, int w, int h, int x, int y, float *n, float *d, c3 *alb) {
    x = wrapi(x, w);
    y = wrapi(y, h);
    size_t i = (size_t)y * (size_t)w + (size_t)x;
    const uint16_t *q = g->gbuf + 4 * i, *a = g->galbedo + 4 * i;
    n[0] = rc_f16_to_f32(q[0]); n[1] = rc_f16_to_f32(q[1]); n[2] = rc_f16_to_f32(q[2]);
    *d = rc_f16_to_f32(q[3]);
    *alb = C3(rc_f16_to_f32(a[0]), rc_f16_to_f32(a[1]), rc_f16_to_f32(a[2]));
}

static inline c3 demod(c3 col, c3 alb) {
    if (luma709(alb) > 0.02f)
        return C3(col.r / fmaxf(alb.r, 0.01f), col.g / fmaxf(alb.g, 0.01f), col.b / fmaxf(alb.b, 0.01f));
    return col;
}

static c3 atrous(img s, const gbufs *g, const rc_post_params *p, int step, int x, int y) {
    static const float hk[5] = {1.0f / 16.0f, 1.0f / 4.0f, 3.0f / 8.0f, 1.0f / 4.0f, 1.0f / 16.0f};
    float cn[3], cd;
    c3 calb;
    gb_fetch(g, s.w, s.h, x, y, cn, &cd, &calb);
    c3 col = fetch(s, x, y);
    if (cd > 9000.0f) return col;
    int hasAlb = luma709(calb) > 0.02f;
    c3 cirr = demod(col, calb);
    float N = (float)(g->total > 1 ? g->total : 1);
    float sumL2 = g->variance[(size_t)y * (size_t)s.w + (size_t)x];
    float ex = fmaxf(p->exposure, 1e-4f);
    float meanL = luma709(C3(col.r / ex, col.g / ex, col.b / ex));
    float var = fmaxf(0.0f, sumL2 / N - meanL * meanL);
    float varDenom = (p->atrousSigmaColor * p->atrousSigmaColor) * var + 1e-6f;
    float cLog = rc_logf(luma709(cirr) + 1e-3f);
    float sd2 = p->atrousSigmaDepth * p->atrousSigmaDepth + 1e-6f;
    c3 sum = C3(0, 0, 0);
    float wsum = 0.0f;
    for (int j = -2; j <= 2; ++j) {
        for (int i = -2; i <= 2; ++i) {
            int x2 = x + i * step, y2 = y + j * step;
            float n2[3], d2;
            c3 a2;
            gb_fetch(g, s.w, s.h, x2, y2, n2, &d2, &a2);
            c3 col2 = fetch(s, x2, y2);
            c3 irr2 = demod(col2, a2);
            float hw = hk[i + 2] * hk[j + 2];
            float nd = (cn[0] * n2[0] + cn[1] * n2[1]) + cn[2] * n2[2];
            float wn = rc_powf(fmaxf(0.0f, nd), p->atrousSigmaNormal);
            float dd = (cd - d2) / (cd + 0.001f);
            float wd = rc_expf(-(dd * dd) / sd2);
            float ld = cLog - rc_logf(luma709(irr2) + 1e-3f);
            float wc = rc_expf(-(ld * ld) / varDenom);
            float w = ((hw * wn) * wd) * wc;
            sum = C3(sum.r + irr2.r * w, sum.g + irr2.g * w, sum.b + irr2.b * w);
            wsum += w;
        }
    }
    c3 res = wsum > 1e-6f ? C3(sum.r / wsum, sum.g / wsum, sum.b / wsum) : cirr;
    if (hasAlb) res = C3(res.r * calb.r, res.g * calb.g, res.b * calb.b);
    return res;
}

/* ---- p01 (post/p01_downsample.frag:36-53) -------------------------------- */
static inline float gaussf(float x, float sigma) { return rc_expf(-(x * x) / ((2.0f * sigma) * sigma)); }

static c3 p01(img s, const rc_post_params *p, int x, int y) {
    float sigma = fmaxf(p->lumaBlurSigma, 0.01f);
    c3 sum = C3(0, 0, 0);
    float wt = 0.0f;
    for (int i = -4; i <= 4; ++i) {
        float w = gaussf((float)i, sigma);
        c3 c = fetch(s, x + i, y);
        sum = C3(sum.r + c.r * w, sum.g + c.g * w, sum.b + c.b * w);
        wt += w;
    }
    return C3(sum.r / wt, sum.g / wt, sum.b / wt);
}

/* ---- p02 (post/p02_chroma_res.frag:55-85) -------------------------------- */
static c3 p02(img s, const rc_post_params *p, int x, int y) {
    float sigma = fmaxf(p->chromaResSigma, 0.5f);
    int radius = clampi((int)(sigma * 2.5f + 0.5f), 1, 32);
    float Y = ycc5(fetch(s, x, y)).r;
    float cb = 0.0f, cr = 0.0f, wt = 0.0f;
    for (int i = -radius; i <= radius; ++i) {
        float w = gaussf((float)i, sigma);
        c3 q = ycc5(fetch(s, x + i, y));
        cb += q.g * w;
        cr += q.b * w;
        wt += w;
    }
    return rgb5(C3(Y, cb / wt, cr / wt));
}

/* ---- p03 (post/p03_color_matrix.frag:56-90) ------------------------------ */
static c3 p03(img s, const rc_post_params *p, const rc_temporal *t, int x, int y) {
    c3 q = ycc3(fetch(s, x, y));
    float cx = (q.g - 0.5f) * p->colorSaturation, cy = (q.b - 0.5f) * p->colorSaturation;
    float rad = p->colorHueRot * 0.017453292519943295f;
    float ch = rc_cosf(rad), sh = rc_sinf(rad);
    float nx = ch * cx - sh * cy, ny = sh * cx + ch * cy;
    float drift = rc_sinf(t->time * p->colorDriftSpeed) * 0.015f;
    float shift = p->colorTempBias * 0.05f + drift;
    ny = ny + shift;
    nx = nx - shift * 0.75f;
    c3 o = rgb3(C3(q.r, nx + 0.5f, ny + 0.5f));
    return C3(clamp01(o.r), clamp01(o.g), clamp01(o.b));
}

/* ---- p04 (post/p04_dynamic_range.frag:52-98) ----------------------------- */
static float crush(float x, float cp) {
    if (x >= cp) return x;
    float t = x / fmaxf(cp, 0.0001f);
    float u = (t * t) * (3.0f - 2.0f * t);
    return u * cp;
}
static float knee(float x, float k, float clip) {
    if (x <= k) return x;
    float t = (x - k) / fmaxf(1.0f - k, 0.0001f);
    float rolled = k + (clip - k) * (1.0f - rc_expf(-3.5f * t));
    return fminf(rolled, clip);
}
static c3 p04(img s, const rc_post_params *p, int x, int y) {
    c3 q = ycc3(fetch(s, x, y));
    float yin = q.r;
    float Y = crush(yin, p->drShadowCrush);
    Y = p->drBlackLift + (1.0f - p->drBlackLift) * Y;
    Y = knee(Y, p->drHighlightKnee, p->drHighlightClip);
    float sc = (yin > 0.001f) ? (Y / yin) : 1.0f;
    c3 o = rgb3(C3(Y, (q.g - 0.5f) * sc + 0.5f, (q.b - 0.5f) * sc + 0.5f));
    return C3(clamp01(o.r), clamp01(o.g), clamp01(o.b));
}

/* ---- p07 (post/p07_ccd_noise.frag:45-115) -------------------------------- */
static inline float p07_rand(int px, int py, int seed) {
    uint32_t h = pcg_hash((uint32_t)px ^ ((uint32_t)py * 1619u) ^ ((uint32_t)seed * 2654435761u));
    return ((float)h / 4294967295.0f) * 2.0f - 1.0f;
}
static c3 p07(img s, const rc_post_params *p, const rc_temporal *t, int x, int y) {
    c3 q = ycc5(fetch(s, x, y));
    float lum = clamp01(q.r);
    float la = (p->ccdNoiseLumaBase + 0.04f * (1.0f - lum)) * t->agcGain;
    float ca = la * p->ccdNoiseChromaScale;
    int f = t->frameIndex;
    float nY = p07_rand(x, y, f) * la;
    float nCb = p07_rand(x, y, f + 137) * ca;
    float nCr = p07_rand(x, y, f + 271) * ca;
    int rowSeed = y * 7 + f / 8;
    float band = (p07_rand(0, rowSeed, 0) * 0.005f) * t->agcGain;
    return rgb5(C3(q.r + (nY + band), q.g + nCb, q.b + nCr));
}

/* ---- p08 (post/p08_ccd_smear.frag:63-112) -------------------------------- */
static c3 p08(img s, const rc_post_params *p, int x, int y) {
    c3 base = fetch(s, x, y);
    float pa = 0.0f;
    int pt = 0;
    for (int i = 1; i <= 150; ++i) {
        if (i > p->ccdSmearLength) break;
        if (y + i > s.h - 1) break;  /* sUv.y > 1.0 */
        float l = luma601(fetch(s, x, y + i));
        pa += fmaxf(0.0f, l - p->ccdSmearThreshold);
        ++pt;
    }
    float prim = pt > 0 ? (pa / (float)pt) * p->ccdSmearIntensity : 0.0f;
    int slen = (int)((float)p->ccdSmearLength * 0.30f);
    if (slen < 1) slen = 1;
    float sa = 0.0f;
    int st = 0;
    for (int i = 1; i <= 150; ++i) {
        if (i > slen) break;
        if (y - i < 0) break;  /* sUv.y < 0.0 */
        float l = luma601(fetch(s, x, y - i));
        sa += fmaxf(0.0f, l - p->ccdSmearThreshold) * rc_powf(0.70f, (float)i);
        ++st;
    }
    float sec = st > 0 ? ((sa / (float)st) * p->ccdSmearIntensity) * 0.25f : 0.0f;
    float tot = prim + sec;
    return C3(clamp01(base.r + tot), clamp01(base.g + tot), clamp01(base.b + tot));
}

/* ---- p09 (post/p09_chroma_bleed.frag:62-94) ------------------------------ */
static c3 p09(img s, const rc_post_params *p, int x, int y) {
    float k = rc_clampf(p->chromaBleedFactor, 0.0f, 0.98f);
    float Y = ycc5(fetch(s, x, y)).r;
    float cb = 0.0f, cr = 0.0f, wt = 0.0f, kp = 1.0f, sc = 1.0f - k;
    for (int i = 0; i < 24; ++i) {
        c3 q = ycc5(fetch(s, x - i, y));
        float w = sc * kp;
        cb += q.g * w;
        cr += q.b * w;
        wt += w;
        kp *= k;
    }
    return rgb5(C3(Y, cb / wt, cr / wt));
}

/* ---- p10 (post/p10_timebase.frag:44-79) ---------------------------------- */
static inline float p10_hash(float x) { return fractf(rc_sinf(x * 127.1f + 311.7f) * 43758.5453f); }
static inline float p10_noise(float x) {
    float i = floorf(x), f = fractf(x);
    float u = (f * f) * (3.0f - 2.0f * f);
    return mixf(p10_hash(i), p10_hash(i + 1.0f), u);
}
static c3 p10(img s, const rc_post_params *p, const rc_temporal *t, int x, int y) {
    float vx = vu(x, s.w), vy = vu(y, s.h);
    float sy = vy * (float)s.h;
    float fb = (float)(t->frameIndex / 4) * 0.137f;
    float sp = rc_sinf(sy * p->timebaseFreq + t->time * p->timebaseSpeed);
    float np = p10_noise(sy * 0.05f + fb) * 2.0f - 1.0f;
    float off = p->timebaseAmplitudePx * (sp * 0.5f + np * 0.5f);
    float uo = off / (float)s.w;
    return bilinear(s, clamp01(vx + uo), vy);
}

/* ---- p11 (post/p11_dot_crawl.frag:48-105) -------------------------------- */
static c3 p11(img s, const rc_post_params *p, const rc_temporal *t, int x, int y) {
    float vx = vu(x, s.w), vy = vu(y, s.h);
    c3 col = fetch(s, x, y);
    float pixX = vx * (float)s.w;
    float lineY = floorf(vy * (float)s.h);
    float lc = luma709(col), ll = luma709(fetch(s, x - 1, y)), lr = luma709(fetch(s, x + 1, y));
    float grad = fmaxf(fabsf(lc - ll), fabsf(lr - lc));
    float gate = smoothstepf(p->dotCrawlEdgeThresh, p->dotCrawlEdgeThresh * 3.0f, grad);
    float lp = lineY * 3.14159265f;
    float fp = (float)t->frameIndex * (3.14159265f * 0.5f);
    float hp = ((pixX * p->dotCrawlSubcarrier) * 2.0f) * 3.14159265f;
    float ph = (hp + lp) + fp;
    float I = (rc_sinf(ph) * p->dotCrawlIntensity) * gate;
    float Q = ((rc_cosf(ph) * p->dotCrawlIntensity) * gate) * 0.5f;
    c3 c = C3(1.0000f * I + 0.9562f * Q, -0.2721f * I - 0.6474f * Q, -1.1069f * I + 1.7046f * Q);
    return C3(clamp01(col.r + c.r), clamp01(col.g + c.g), clamp01(col.b + c.b));
}

/* ---- p12 (post/p12_dropout.frag:59-200) ---------------------------------- */
static c3 p12(img s, const rc_post_params *p, const rc_temporal *t, int x, int y) {
    float W = (float)s.w, H = (float)s.h;
    float vx = vu(x, s.w), vy = vu(y, s.h);
    c3 col = fetch(s, x, y);
    float fragX = vx * W, fragY = floorf(vy * H);
    uint32_t vf = (uint32_t)(t->frameIndex / 2);
    /* burstActivity (:39-57) */
    uint32_t win = vf / 30u, prev = win - 1u;
    float thr = 1.0f - p->dropoutBurstRate;
    float thisA = fhash1(win, 0xDEADBEEFu) < thr ? 0.0f : 1.0f;
    float prevA = fhash1(prev, 0xDEADBEEFu) < thr ? 0.0f : 1.0f;
    float wpos = (float)(vf % 30u) / 30.0f;
    float bleed = prevA * smoothstepf(0.7f, 1.0f, 1.0f - wpos);
    float act = fmaxf(thisA, bleed);
    if (act < 0.001f) return col;
    float age = 0.05f + p->tapeAge * 0.95f;
    int nEv = 1 + (int)roundf(((act * age) * 13.0f));
    float bw = 0.0f, bu = vx, bv = vy, bb = 0.0f;
    for (int i = 0; i < nEv; ++i) {
        uint32_t si = (uint32_t)i;
        float evY = floorf(fhash1(vf * 31u + si, si * 13u + 1u) * H);
        int evH = (fhash1(vf + si * 3u, si + 77u) > 0.88f) ? 2 : 1;
        if (fragY < evY || fragY >= evY + (float)evH) continue;
        float wr = fhash1(vf * 7u + si, si * 7u + 2u);
        float evW;
        if (fhash1(vf * 2u + si, si + 55u) > 0.88f) evW = 80.0f + wr * 140.0f;
        else evW = 8.0f + (wr * wr) * 32.0f;
        float maxStart = fmaxf(0.0f, W - evW);
        float x0 = fhash1(vf * 5u + si, si * 7u + 3u) * maxStart;
        if (fragX < x0 || fragX >= x0 + evW) continue;
        float dL = fragX - x0, dR = (x0 + evW) - fragX;
        float w = smoothstepf(0.0f, 1.5f, dL) * smoothstepf(0.0f, 5.0f, dR);
        float dxn = fhash1(vf * 11u + si, si * 7u + 4u) * 2.0f - 1.0f;
        float dx = dxn * (4.0f + p->tapeAge * 16.0f);
        float dyr = fhash1(vf * 13u + si, si * 7u + 5u);
        float dy = (dyr > 0.6f) ? (dyr > 0.8f ? -1.0f : 1.0f) : 0.0f;
        float str = fhash1(vf * 17u + si, si * 7u + 6u);
        float stretch = (str > 0.7f) ? (1.0f + (str * p->tapeAge) * 0.5f) : 1.0f;
        float cx = x0 + evW * 0.5f;
        float sx = (cx + (fragX - cx) * stretch) + dx;
        float du = clamp01(sx / W), dv = clamp01((fragY + dy) / H);
        float burst = smoothstepf(3.0f, 0.0f, dL) * p->dropoutBright;
        if (w > bw) { bw = w; bu = du; bv = dv; bb = burst; }
        break;
    }
    if (bw < 0.001f) return col;
    c3 r = bilinear(s, bu, bv);
    float g = 1.0f + bb * 0.3f;
    r = C3(r.r * g, r.g * g, r.b * g);
    float lr = luma709(r);
    float m = bb * 0.4f;
    r = C3(mixf(r.r, lr, m), mixf(r.g, lr, m), mixf(r.b, lr, m));
    return C3(mixf(col.r, r.r, bw), mixf(col.g, r.g, bw), mixf(col.b, r.b, bw));
}

/* ---- p13 (post/p13_headswitch.frag:78-191) ------------------------------- */
static inline float p13_noise(float x, uint32_t seed) {
    float i = floorf(x), f = fractf(x);
    float u = (f * f) * (3.0f - 2.0f * f);
    float a = fhash2((uint32_t)i, seed), b = fhash2((uint32_t)i + 1u, seed);
    return mixf(a, b, u) * 2.0f - 1.0f;
}
static inline void rot2(float *x, float *y, float th) {
    float c = rc_cosf(th), s = rc_sinf(th);
    float nx = *x * c - *y * s, ny = *x * s + *y * c;
    *x = nx; *y = ny;
}
static c3 p13(img s, const rc_post_params *p, const rc_temporal *t, int x, int y) {
    float vx = vu(x, s.w), vy = vu(y, s.h);
    float dist = floorf(vy * (float)s.h);
    float zone = (float)p->headSwitchLines;
    if (dist >= zone) return fetch(s, x, y);
    float prox = 1.0f - dist / zone;
    float settle = prox * prox;
    uint32_t fs = (uint32_t)t->frameIndex;
    float sn = p13_noise(dist * 0.8f, fs);
    float fn = p13_noise(dist * 3.0f, fs + 7u);
    float jit = (p->headSwitchJitter * settle) * (sn * 0.7f + fn * 0.3f);
    float uo = jit / (float)s.w;
    float pr = fhash2(fs, 0xCAFEBABEu) * 6.283185f;
    float rs = rc_sinf(dist * 1.570796f + pr);
    float ripple = (p->headSwitchRipple * settle) * rs;
    float sgn = (fhash2(fs, 0xBEEFu) > 0.5f) ? 1.0f : -1.0f;
    float ang = (p->headSwitchChroma * settle) * sgn;
    c3 q = ycc5(bilinear(s, clamp01(vx + uo), vy));
    float cb = q.g - 0.5f, cr = q.b - 0.5f;
    rot2(&cb, &cr, ang);
    return rgb5(C3(q.r + ripple, cb + 0.5f, cr + 0.5f));
}

/* ---- p14 (post/p14_tracking.frag:83-269) --------------------------------- */
static inline float p14_noise(float x, uint32_t seed) {
    float i = floorf(x), f = fractf(x);
    float u = (f * f) * (3.0f - 2.0f * f);
    return mixf(fhash2((uint32_t)i, seed), fhash2((uint32_t)i + 1u, seed), u);
}
static c3 p14(img s, const rc_post_params *p, const rc_temporal *t, int x, int y) {
    float W = (float)s.w, H = (float)s.h;
    float vx = vu(x, s.w), vy = vu(y, s.h);
    float fragY = floorf(vy * H);
    c3 col = fetch(s, x, y);
    if (p->trackingSeverity < 0.001f) return col;
    float sev = p->trackingSeverity;
    uint32_t frame = (uint32_t)t->frameIndex;
    uint32_t ce = frame / 30u;
    float tDisp = 0.0f, tFr = 0.0f, tNz = 0.0f, tHue = 0.0f, tW = 0.0f;
    for (int e = 0; e < 2; ++e) {
        uint32_t ei = (uint32_t)e;
        for (int ep = 0; ep < 3; ++ep) {
            uint32_t epoch = ce - (uint32_t)ep;
            float spawn = 0.08f + 0.62f * sev;
            if (fhash2(epoch * 7919u + ei * 3u, 0xDEADBEEFu) > spawn) continue;
            float dr = fhash2(epoch * 1009u + ei, 0xCAFEu);
            uint32_t dur = (uint32_t)(15.0f + dr * 105.0f);
            float sr = fhash2(epoch * 503u + ei, 0xBEEFu);
            uint32_t st = epoch * 30u + (uint32_t)(sr * 30.0f);
            uint32_t en = st + dur;
            if (frame < st || frame >= en) continue;
            float yr = fhash2(epoch * 2017u + ei, 0x1111u);
            float hr = fhash2(epoch * 3001u + ei, 0x2222u);
            float eh = 20.0f + (hr * 60.0f) * sev;
            float y0 = yr * (H - eh);
            float y1 = y0 + eh;
            if (fragY < y0 || fragY >= y1) continue;
            float bt = (fragY - y0) / fmaxf(eh - 1.0f, 1.0f);
            float ds = (fhash2(epoch * 4003u + ei, 0x3333u) > 0.5f) ? 1.0f : -1.0f;
            float base = (((ds * p->trackingMaxDisplacePx) * sev) * (bt - 0.5f)) * 2.0f;
            float fph = (float)frame * 0.37f + (float)ei * 1.57f;
            float wob = ((rc_sinf(bt * 6.28318f + fph) * p->trackingMaxDisplacePx) * 0.20f) * sev;
            float mj = ((p14_noise(fragY * 0.25f, frame + ei * 17u) * 2.0f - 1.0f) * 2.5f) * sev;
            float tot = (base + wob) + mj;
            float dT = fragY - y0, dB = y1 - fragY;
            float ew = smoothstepf(0.0f, 1.5f, dT) * smoothstepf(0.0f, 4.0f, dB);
            float edd = fminf(dT, dB);
            float fr = (1.0f - smoothstepf(0.0f, p->trackingFringeWidthPx, edd)) * ew;
            float nz = fr * 0.7f;
            float hue = ((ds * 0.436f) * sev) * (1.0f - fr);
            if (ew > tW) { tW = ew; tDisp = tot; tFr = fr; tNz = nz; tHue = hue; }
        }
    }
    if (tW < 0.001f) return col;
    float us = (tDisp / W) * tW;
    c3 q = ycc5(bilinear(s, clamp01(vx + us), vy));
    if (fabsf(tHue) > 0.001f) {
        float cb = q.g - 0.5f, cr = q.b - 0.5f;
        rot2(&cb, &cr, tHue);
        q.g = cb + 0.5f;
        q.b = cr + 0.5f;
    }
    if (tFr > 0.001f) {
        float fph = (float)t->frameIndex * 0.12f + vy * 8.0f;
        float fa = rc_sinf(fph) * 0.9f;
        float cb = q.g - 0.5f, cr = q.b - 0.5f;
        rot2(&cb, &cr, fa);
        float g = 1.0f + tFr * 1.2f;
        cb *= g;
        cr *= g;
        q.g = mixf(q.g, cb + 0.5f, tFr);
        q.b = mixf(q.b, cr + 0.5f, tFr);
        uint32_t ns = uhash2(((uint32_t)fragY * 1619u) ^ ((uint32_t)t->frameIndex * 2654435761u));
        float fx = (float)x + 0.5f;  /* gl_FragCoord.x */
        float buzz = ((float)uhash2(ns ^ (uint32_t)(fx * 37.0f)) / 4294967295.0f) * 2.0f - 1.0f;
        q.r = q.r + (buzz * 0.15f) * tNz;
    }
    c3 o = rgb5(q);
    return C3(mixf(col.r, o.r, tW), mixf(col.g, o.g, tW), mixf(col.b, o.b, tW));
}

/* ---- p15 (post/p15_edge_enhance.frag:80-120) ----------------------------- */
static c3 p15(img s, const rc_post_params *p, int x, int y) {
    static const float w[5] = {0.0625f, 0.25f, 0.375f, 0.25f, 0.0625f};
    c3 q = ycc5(fetch(s, x, y));
    float Y = q.r, blur = 0.0f;
    for (int dy = -2; dy <= 2; ++dy)
        for (int dx = -2; dx <= 2; ++dx)
            blur += (ycc5(fetch(s, x + dx, y + dy)).r * w[dx + 2]) * w[dy + 2];
    float det = Y - blur;
    float cw = smoothstepf(p->edgeEnhanceCoreThreshold * 0.5f, p->edgeEnhanceCoreThreshold, fabsf(det));
    det *= cw;
    return rgb5(C3(Y + p->edgeEnhanceAmount * det, q.g, q.b));
}

/* ---- p16 (post/p16_interlace.frag:81-141) -------------------------------- */
static c3 p16(img s, const rc_post_params *p, int x, int y) {
    float vx = vu(x, s.w), vy = vu(y, s.h);
    float pxx = 1.0f / (float)s.w, pxy = 1.0f / (float)s.h;
    int even = (y & 1) == 0;
    float vs = even ? ((p->interlaceFieldOffsetPx * 0.5f) * pxy) : ((-p->interlaceFieldOffsetPx * 0.5f) * pxy);
    float yh = luma601(fetch(s, x, y)), ya = luma601(fetch(s, x, y - 1));
    float yb = luma601(fetch(s, x, y + 1)), yl = luma601(fetch(s, x - 1, y));
    float vg = fabsf(ya - yb);
    float hd = signf(yh - yl);
    float comb = 0.0f;
    if (even) {
        float ew = smoothstepf(p->interlaceCombEdgeThresh, p->interlaceCombEdgeThresh * 3.0f, vg);
        comb = (hd * ew) * p->interlaceCombStrength;
    }
    float su = clamp01(vx + comb * pxx), sv = clamp01(vy + vs);
    c3 c = bilinear(s, su, sv);
    float m = even ? (1.0f - p->interlaceLineWeighting) : (1.0f + p->interlaceLineWeighting * 0.5f);
    return C3(c.r * m, c.g * m, c.b * m);
}

/* ---- display (display.frag:11-42), u_totalSamples = 1, u_exposure = 1 ---- */
static inline float aces1(float x) {
    return rc_clampf((x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f), 0.0f, 1.0f);
}
static inline float srgb1(float c) {
    return (c < 0.0031308f) ? 12.92f * c : 1.055f * rc_powf(c, 1.0f / 2.4f) - 0.055f;
}
static inline uint8_t unorm8(float v) {
    float c = rc_clampf(v, 0.0f, 1.0f);
    return (uint8_t)floorf(c * 255.0f + 0.5f);
}

/* Integer hashes of the VHS passes, exported for the known-answer tests. */
uint32_t rco_hash(int32_t kind, uint32_t a, uint32_t b) {
    float f;
    uint32_t u;
    switch (kind) {
        case 0: return pcg_hash(a);
        case 1: return uhash1(a);
        case 2: return uhash2(a);
        case 3: f = fhash1(a, b); memcpy(&u, &f, 4); return u;
        case 4: f = fhash2(a, b); memcpy(&u, &f, 4); return u;
        default: return 0;
    }
}

/* ---- driver ---------------------------------------------------------------- */
/* Every pass is a pure function of its input image per output pixel, so rows
 * run in parallel without changing any value. */
static void run_pass(int pass, img s, float *dst, const rc_post_params *p, const rc_temporal *t, const gbufs *g) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < s.h; ++y) {
        for (int x = 0; x < s.w; ++x) {
            c3 o;
            switch (pass) {
                case 0: o = p00(s, x, y, g->total, p->exposure); break;
                case 1: o = p01(s, p, x, y); break;
                case 2: o = p02(s, p, x, y); break;
                case 3: o = p03(s, p, t, x, y); break;
                case 4: o = p04(s, p, x, y); break;
                case 7: o = p07(s, p, t, x, y); break;
                case 8: o = p08(s, p, x, y); break;
                case 9: o = p09(s, p, x, y); break;
                case 10: o = p10(s, p, t, x, y); break;
                case 11: o = p11(s, p, t, x, y); break;
                case 12: o = p12(s, p, t, x, y); break;
                case 13: o = p13(s, p, t, x, y); break;
                case 14: o = p14(s, p, t, x, y); break;
                case 15: o = p15(s, p, x, y); break;
                case 16: o = p16(s, p, x, y); break;
                case 200: {
                    c3 c = fetch(s, x, y);
                    o = C3(srgb1(aces1(c.r)), srgb1(aces1(c.g)), srgb1(aces1(c.b)));
                    break;
                }
                default:
                    if (pass >= 100 && pass < 105) o = atrous(s, g, p, 1 << (pass - 100), x, y);
                    else o = fetch(s, x, y);
            }
            float *q = dst + 4 * ((size_t)y * (size_t)s.w + (size_t)x);
            q[0] = o.r; q[1] = o.g; q[2] = o.b; q[3] = 1.0f;
        }
    }
}

int32_t rco_post_pass(int32_t pass, const float *src, float *dst, int32_t W, int32_t H, const rc_post_params *p,
                      const rc_temporal *t, const uint16_t *gbuf, const uint16_t *galbedo, const float *variance,
                      int32_t totalSamples) {
    if (!src || !dst || W <= 0 || H <= 0 || !p || !t) return -1;
    if (pass >= 100 && pass < 105 && (!gbuf || !galbedo || !variance)) return -1;
    img s = {src, W, H};
    gbufs g = {gbuf, galbedo, variance, totalSamples, p->exposure};
    run_pass(pass, s, dst, p, t, &g);
    return 0;
}

int32_t rco_post_chain(const float *accum, const uint16_t *gbuf, const uint16_t *galbedo, const float *variance,
                       int32_t W, int32_t H, int32_t totalSamples, const rc_post_params *p, const rc_temporal *t,
                       float *outHdr, uint8_t *outRgba8) {
    if (!accum || W <= 0 || H <= 0 || !p || !t) return -1;
    size_t n = (size_t)W * (size_t)H * 4;
    float *a = (float *)malloc(n * sizeof(float)), *b = (float *)malloc(n * sizeof(float));
    if (!a || !b) { free(a); free(b); return -1; }
    gbufs g = {gbuf, galbedo, variance, totalSamples, p->exposure};
    img s = {accum, W, H};
    run_pass(0, s, a, p, t, &g);  /* PostProcessStack.runOnAccum (PostProcessStack.java:109-118) */
    int exceeded = p->atrousMaxSpp > 0 && totalSamples > p->atrousMaxSpp;
    if (p->atrousEnabled && gbuf && galbedo && variance && !exceeded) {
        int np = p->atrousIterations < 1 ? 1 : (p->atrousIterations > 5 ? 5 : p->atrousIterations);
        for (int i = 0; i < np; ++i) {
            img cur = {a, W, H};
            run_pass(100 + i, cur, b, p, t, &g);
            float *tmp = a; a = b; b = tmp;
        }
    }
    for (int k = 1; k <= 16; ++k) {
        if (k == 5 || k == 6 || !p->enable[k]) continue;
        img cur = {a, W, H};
        run_pass(k, cur, b, p, t, &g);
        float *tmp = a; a = b; b = tmp;
    }
    if (outHdr) memcpy(outHdr, a, n * sizeof(float));
    if (outRgba8) {
        img cur = {a, W, H};
        run_pass(200, cur, b, p, t, &g);
        for (size_t i = 0; i < n; ++i) outRgba8[i] = unorm8(b[i]);
    }
    free(a);
    free(b);
    return 0;
}
