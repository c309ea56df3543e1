// rc_post.hip — post chain for gfx950: p00 normalize, a-trous denoise, VHS
// passes p01..p16 and the display encode (ACES + sRGB -> RGBA8).
//
// Each pass is one kernel over the frame (16x16 pixel tiles, 256 threads),
// ping-ponging between two RGBA32F buffers like PostProcessStack's FBO chain
// (src/main/java/com/retrocam/post/PostProcessStack.java:140-295). Shader
// paths below are relative to src/main/resources/shaders/.
//
// Sampling: fragment-centre UVs ((x + 0.5)/W, (y + 0.5)/H) (fullscreen.vert:9-16).
// Whole-texel offsets from the centre are texel fetches — clamped to the edge
// for the RGBA32F chain images (GL_CLAMP_TO_EDGE, gl/Framebuffer.java:74-77),
// wrapped for the Renderer's G-buffer textures (default GL_REPEAT,
// core/Renderer.java:163-183). Fractional positions (p10, p12, p13, p14, p16)
// are bilinear with edge clamping, in float32.
#include "rc_device.h"
#include "rc_internal.h"

namespace rc {

struct c3 { float r, g, b; };
__device__ __forceinline__ c3 C3(float r, float g, float b) { c3 c; c.r = r; c.g = g; c.b = b; return c; }

struct Img {
    const float4 *__restrict__ p;
    int w, h;
    __device__ __forceinline__ c3 at(int x, int y) const {
        x = min(max(x, 0), w - 1);
        y = min(max(y, 0), h - 1);
        float4 v = p[(size_t)y * (size_t)w + (size_t)x];
        return C3(v.x, v.y, v.z);
    }
    __device__ c3 bilinear(float u, float v) const {
        float x = u * (float)w - 0.5f, y = v * (float)h - 0.5f;
        float fx = floorf(x), fy = floorf(y);
        float ax = x - fx, ay = y - fy;
        int x0 = (int)fx, y0 = (int)fy;
        c3 t00 = at(x0, y0), t10 = at(x0 + 1, y0), t01 = at(x0, y0 + 1), t11 = at(x0 + 1, y0 + 1);
        float bx = 1.0f - ax, by = 1.0f - ay;
        c3 r0 = C3(t00.r * bx + t10.r * ax, t00.g * bx + t10.g * ax, t00.b * bx + t10.b * ax);
        c3 r1 = C3(t01.r * bx + t11.r * ax, t01.g * bx + t11.g * ax, t01.b * bx + t11.b * ax);
        return C3(r0.r * by + r1.r * ay, r0.g * by + r1.g * ay, r0.b * by + r1.b * ay);
    }
};

struct GBufs {
    const ushort4 *gbuf, *galbedo;
    const float *variance;
    int total;
};

struct PostArgs {
    Img src;
    float4 *dst;
    rc_post_params p;
    rc_temporal t;
    const rc_temporal *dt;  // non-null: read t from this device slot (captured frames)
    GBufs g;
    int step;        // a-trous step width
    uchar4 *rgba8;   // display output
};

__device__ __forceinline__ float vu(int x, int w) { return ((float)x + 0.5f) / (float)w; }
__device__ __forceinline__ float clamp01(float v) { return rc_clampf(v, 0.0f, 1.0f); }
__device__ __forceinline__ float sstep(float e0, float e1, float x) {
    float t = rc_clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return (t * t) * (3.0f - 2.0f * t);
}
__device__ __forceinline__ float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ float fractf(float x) { return x - floorf(x); }
__device__ __forceinline__ float luma709(c3 c) { return (c.r * 0.2126f + c.g * 0.7152f) + c.b * 0.0722f; }
__device__ __forceinline__ float luma601(c3 c) { return (0.299f * c.r + 0.587f * c.g) + 0.114f * c.b; }

__device__ __forceinline__ c3 ycc5(c3 c) {
    return C3((0.29900f * c.r + 0.58700f * c.g) + 0.11400f * c.b,
              ((-0.16874f * c.r - 0.33126f * c.g) + 0.50000f * c.b) + 0.5f,
              ((0.50000f * c.r - 0.41869f * c.g) - 0.08131f * c.b) + 0.5f);
}
__device__ __forceinline__ c3 rgb5(c3 y) {
    float cb = y.g - 0.5f, cr = y.b - 0.5f;
    return C3(y.r + 1.40200f * cr, (y.r - 0.34414f * cb) - 0.71414f * cr, y.r + 1.77200f * cb);
}
__device__ __forceinline__ c3 ycc3(c3 c) {
    return C3((0.299f * c.r + 0.587f * c.g) + 0.114f * c.b,
              ((-0.169f * c.r - 0.331f * c.g) + 0.500f * c.b) + 0.5f,
              ((0.500f * c.r - 0.419f * c.g) - 0.081f * c.b) + 0.5f);
}
__device__ __forceinline__ c3 rgb3(c3 y) {
    float cb = y.g - 0.5f, cr = y.b - 0.5f;
    return C3(y.r + 1.402f * cr, (y.r - 0.344f * cb) - 0.714f * cr, y.r + 1.772f * cb);
}

__device__ __forceinline__ uint32_t pcg_hash(uint32_t v) {
    v = v * 747796405u + 2891336453u;
    uint32_t w = ((v >> ((v >> 28u) + 4u)) ^ v) * 277803737u;
    return (w >> 22u) ^ w;
}
// p12_dropout.frag:32-37 (one round) and p13/p14 (two rounds)
__device__ __forceinline__ uint32_t uhash1(uint32_t x) { x ^= x >> 16u; x *= 0x45d9f3bu; x ^= x >> 16u; return x; }
__device__ __forceinline__ uint32_t uhash2(uint32_t x) {
    x = ((x >> 16u) ^ x) * 0x45d9f3bu;
    x = ((x >> 16u) ^ x) * 0x45d9f3bu;
    return (x >> 16u) ^ x;
}
__device__ __forceinline__ float fhash1(uint32_t a, uint32_t b) { return (float)uhash1(a ^ uhash1(b) * 2654435761u) / 4294967295.0f; }
__device__ __forceinline__ float fhash2(uint32_t a, uint32_t b) { return (float)uhash2(a ^ uhash2(b) * 2654435761u) / 4294967295.0f; }
__device__ __forceinline__ float gaussf(float x, float s) { return rc_expf(-(x * x) / ((2.0f * s) * s)); }
__device__ __forceinline__ void rot2(float &x, float &y, float th) {
    float c = rc_cosf(th), s = rc_sinf(th);
    float nx = x * c - y * s, ny = x * s + y * c;
    x = nx; y = ny;
}

// p00_normalize.frag:22-27
__device__ c3 pass_p00(const PostArgs &A, int x, int y) {
    float n = (float)(A.g.total > 1 ? A.g.total : 1);
    c3 c = A.src.at(x, y);
    return C3((c.r / n) * A.p.exposure, (c.g / n) * A.p.exposure, (c.b / n) * A.p.exposure);
}

__device__ __forceinline__ void gb_at(const GBufs &g, int w, int h, int x, int y, float *n, float &d, c3 &alb) {
    x %= w; if (x < 0) x += w;
    y %= h; if (y < 0) y += h;
    size_t i = (size_t)y * (size_t)w + (size_t)x;
    ushort4 q = g.gbuf[i], a = g.galbedo[i];
    n[0] = rc_f16_to_f32(q.x); n[1] = rc_f16_to_f32(q.y); n[2] = rc_f16_to_f32(q.z);
    d = rc_f16_to_f32(q.w);
    alb = C3(rc_f16_to_f32(a.x), rc_f16_to_f32(a.y), rc_f16_to_f32(a.z));
}
__device__ __forceinline__ c3 demod(c3 col, c3 alb) {
    if (luma709(alb) > 0.02f)
        return C3(col.r / fmaxf(alb.r, 0.01f), col.g / fmaxf(alb.g, 0.01f), col.b / fmaxf(alb.b, 0.01f));
    return col;
}

// atrous_denoise.frag:50-122
__device__ c3 pass_atrous(const PostArgs &A, int x, int y) {
    const float hk[5] = {1.0f / 16.0f, 1.0f / 4.0f, 3.0f / 8.0f, 1.0f / 4.0f, 1.0f / 16.0f};
    const Img &s = A.src;
    float cn[3], cd;
    c3 calb;
    gb_at(A.g, s.w, s.h, x, y, cn, cd, calb);
    c3 col = s.at(x, y);
    if (cd > 9000.0f) return col;
    bool hasAlb = luma709(calb) > 0.02f;
    c3 cirr = demod(col, calb);
    float N = (float)(A.g.total > 1 ? A.g.total : 1);
    float sumL2 = A.g.variance[(size_t)y * (size_t)s.w + (size_t)x];
    float ex = fmaxf(A.p.exposure, 1e-4f);
    float meanL = luma709(C3(col.r / ex, col.g / ex, col.b / ex));
    float var = fmaxf(0.0f, sumL2 / N - meanL * meanL);
    float varDenom = (A.p.atrousSigmaColor * A.p.atrousSigmaColor) * var + 1e-6f;
    float cLog = rc_logf(luma709(cirr) + 1e-3f);
    float sd2 = A.p.atrousSigmaDepth * A.p.atrousSigmaDepth + 1e-6f;
    c3 sum = C3(0, 0, 0);
    float wsum = 0.0f;
    for (int j = -2; j <= 2; ++j) {
        for (int i = -2; i <= 2; ++i) {
            int x2 = x + i * A.step, y2 = y + j * A.step;
            float n2[3], d2;
            c3 a2;
            gb_at(A.g, s.w, s.h, x2, y2, n2, d2, a2);
            c3 irr2 = demod(s.at(x2, y2), a2);
            float hw = hk[i + 2] * hk[j + 2];
            float nd = (cn[0] * n2[0] + cn[1] * n2[1]) + cn[2] * n2[2];
            float wn = rc_powf(fmaxf(0.0f, nd), A.p.atrousSigmaNormal);
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

// p01_downsample.frag:36-53
__device__ c3 pass_p01(const PostArgs &A, int x, int y) {
    float sigma = fmaxf(A.p.lumaBlurSigma, 0.01f);
    c3 sum = C3(0, 0, 0);
    float wt = 0.0f;
    for (int i = -4; i <= 4; ++i) {
        float w = gaussf((float)i, sigma);
        c3 c = A.src.at(x + i, y);
        sum = C3(sum.r + c.r * w, sum.g + c.g * w, sum.b + c.b * w);
        wt += w;
    }
    return C3(sum.r / wt, sum.g / wt, sum.b / wt);
}

// p02_chroma_res.frag:55-85
__device__ c3 pass_p02(const PostArgs &A, int x, int y) {
    float sigma = fmaxf(A.p.chromaResSigma, 0.5f);
    int radius = min(max((int)(sigma * 2.5f + 0.5f), 1), 32);
    float Y = ycc5(A.src.at(x, y)).r;
    float cb = 0.0f, cr = 0.0f, wt = 0.0f;
    for (int i = -radius; i <= radius; ++i) {
        float w = gaussf((float)i, sigma);
        c3 q = ycc5(A.src.at(x + i, y));
        cb += q.g * w;
        cr += q.b * w;
        wt += w;
    }
    return rgb5(C3(Y, cb / wt, cr / wt));
}

// p03_color_matrix.frag:56-90
__device__ c3 pass_p03(const PostArgs &A, int x, int y) {
    const rc_post_params &p = A.p;
    c3 q = ycc3(A.src.at(x, y));
    float cx = (q.g - 0.5f) * p.colorSaturation, cy = (q.b - 0.5f) * p.colorSaturation;
    float rad = p.colorHueRot * 0.017453292519943295f;
    float ch = rc_cosf(rad), sh = rc_sinf(rad);
    float nx = ch * cx - sh * cy, ny = sh * cx + ch * cy;
    float drift = rc_sinf(A.t.time * p.colorDriftSpeed) * 0.015f;
    float shift = p.colorTempBias * 0.05f + drift;
    ny = ny + shift;
    nx = nx - shift * 0.75f;
    c3 o = rgb3(C3(q.r, nx + 0.5f, ny + 0.5f));
    return C3(clamp01(o.r), clamp01(o.g), clamp01(o.b));
}

// p04_dynamic_range.frag:52-98
__device__ c3 pass_p04(const PostArgs &A, int x, int y) {
    const rc_post_params &p = A.p;
    c3 q = ycc3(A.src.at(x, y));
    float yin = q.r, Y = yin;
    if (Y < p.drShadowCrush) {
        float t = Y / fmaxf(p.drShadowCrush, 0.0001f);
        Y = ((t * t) * (3.0f - 2.0f * t)) * p.drShadowCrush;
    }
    Y = p.drBlackLift + (1.0f - p.drBlackLift) * Y;
    if (Y > p.drHighlightKnee) {
        float t = (Y - p.drHighlightKnee) / fmaxf(1.0f - p.drHighlightKnee, 0.0001f);
        float rolled = p.drHighlightKnee + (p.drHighlightClip - p.drHighlightKnee) * (1.0f - rc_expf(-3.5f * t));
        Y = fminf(rolled, p.drHighlightClip);
    }
    float sc = (yin > 0.001f) ? (Y / yin) : 1.0f;
    c3 o = rgb3(C3(Y, (q.g - 0.5f) * sc + 0.5f, (q.b - 0.5f) * sc + 0.5f));
    return C3(clamp01(o.r), clamp01(o.g), clamp01(o.b));
}

// p07_ccd_noise.frag:45-115
__device__ __forceinline__ float p07_rand(int px, int py, int seed) {
    uint32_t h = pcg_hash((uint32_t)px ^ ((uint32_t)py * 1619u) ^ ((uint32_t)seed * 2654435761u));
    return ((float)h / 4294967295.0f) * 2.0f - 1.0f;
}
__device__ c3 pass_p07(const PostArgs &A, int x, int y) {
    const rc_post_params &p = A.p;
    c3 q = ycc5(A.src.at(x, y));
    float lum = clamp01(q.r);
    float la = (p.ccdNoiseLumaBase + 0.04f * (1.0f - lum)) * A.t.agcGain;
    float ca = la * p.ccdNoiseChromaScale;
    int f = A.t.frameIndex;
    float nY = p07_rand(x, y, f) * la;
    float nCb = p07_rand(x, y, f + 137) * ca;
    float nCr = p07_rand(x, y, f + 271) * ca;
    int rowSeed = y * 7 + f / 8;
    float band = (p07_rand(0, rowSeed, 0) * 0.005f) * A.t.agcGain;
    return rgb5(C3(q.r + (nY + band), q.g + nCb, q.b + nCr));
}

// p08_ccd_smear.frag:63-112
__device__ c3 pass_p08(const PostArgs &A, int x, int y) {
    const rc_post_params &p = A.p;
    const Img &s = A.src;
    c3 base = s.at(x, y);
    float pa = 0.0f;
    int pt = 0;
    for (int i = 1; i <= 150; ++i) {
        if (i > p.ccdSmearLength) break;
        if (y + i > s.h - 1) break;  // sUv.y > 1.0
        pa += fmaxf(0.0f, luma601(s.at(x, y + i)) - p.ccdSmearThreshold);
        ++pt;
    }
    float prim = pt > 0 ? (pa / (float)pt) * p.ccdSmearIntensity : 0.0f;
    int slen = max(1, (int)((float)p.ccdSmearLength * 0.30f));
    float sa = 0.0f;
    int st = 0;
    for (int i = 1; i <= 150; ++i) {
        if (i > slen) break;
        if (y - i < 0) break;  // sUv.y < 0.0
        sa += fmaxf(0.0f, luma601(s.at(x, y - i)) - p.ccdSmearThreshold) * rc_powf(0.70f, (float)i);
        ++st;
    }
    float sec = st > 0 ? ((sa / (float)st) * p.ccdSmearIntensity) * 0.25f : 0.0f;
    float tot = prim + sec;
    return C3(clamp01(base.r + tot), clamp01(base.g + tot), clamp01(base.b + tot));
}

// p09_chroma_bleed.frag:62-94
__device__ c3 pass_p09(const PostArgs &A, int x, int y) {
    float k = rc_clampf(A.p.chromaBleedFactor, 0.0f, 0.98f);
    float Y = ycc5(A.src.at(x, y)).r;
    float cb = 0.0f, cr = 0.0f, wt = 0.0f, kp = 1.0f, sc = 1.0f - k;
    for (int i = 0; i < 24; ++i) {
        c3 q = ycc5(A.src.at(x - i, y));
        float w = sc * kp;
        cb += q.g * w;
        cr += q.b * w;
        wt += w;
        kp *= k;
    }
    return rgb5(C3(Y, cb / wt, cr / wt));
}

// p10_timebase.frag:44-79
__device__ __forceinline__ float p10_hash(float x) { return fractf(rc_sinf(x * 127.1f + 311.7f) * 43758.5453f); }
__device__ c3 pass_p10(const PostArgs &A, int x, int y) {
    const Img &s = A.src;
    float vx = vu(x, s.w), vy = vu(y, s.h);
    float sy = vy * (float)s.h;
    float fb = (float)(A.t.frameIndex / 4) * 0.137f;
    float sp = rc_sinf(sy * A.p.timebaseFreq + A.t.time * A.p.timebaseSpeed);
    float nx = sy * 0.05f + fb;
    float i = floorf(nx), f = fractf(nx);
    float u = (f * f) * (3.0f - 2.0f * f);
    float np = mixf(p10_hash(i), p10_hash(i + 1.0f), u) * 2.0f - 1.0f;
    float off = A.p.timebaseAmplitudePx * (sp * 0.5f + np * 0.5f);
    return s.bilinear(clamp01(vx + off / (float)s.w), vy);
}

// p11_dot_crawl.frag:48-105
__device__ c3 pass_p11(const PostArgs &A, int x, int y) {
    const Img &s = A.src;
    const rc_post_params &p = A.p;
    float vx = vu(x, s.w), vy = vu(y, s.h);
    c3 col = s.at(x, y);
    float pixX = vx * (float)s.w;
    float lineY = floorf(vy * (float)s.h);
    float lc = luma709(col), ll = luma709(s.at(x - 1, y)), lr = luma709(s.at(x + 1, y));
    float grad = fmaxf(fabsf(lc - ll), fabsf(lr - lc));
    float gate = sstep(p.dotCrawlEdgeThresh, p.dotCrawlEdgeThresh * 3.0f, grad);
    float lp = lineY * 3.14159265f;
    float fp = (float)A.t.frameIndex * (3.14159265f * 0.5f);
    float hp = ((pixX * p.dotCrawlSubcarrier) * 2.0f) * 3.14159265f;
    float ph = (hp + lp) + fp;
    float I = (rc_sinf(ph) * p.dotCrawlIntensity) * gate;
    float Q = ((rc_cosf(ph) * p.dotCrawlIntensity) * gate) * 0.5f;
    return C3(clamp01(col.r + (1.0000f * I + 0.9562f * Q)), clamp01(col.g + (-0.2721f * I - 0.6474f * Q)),
              clamp01(col.b + (-1.1069f * I + 1.7046f * Q)));
}

// p12_dropout.frag:59-200
__device__ c3 pass_p12(const PostArgs &A, int x, int y) {
    const Img &s = A.src;
    const rc_post_params &p = A.p;
    float W = (float)s.w, H = (float)s.h;
    float vx = vu(x, s.w), vy = vu(y, s.h);
    c3 col = s.at(x, y);
    float fragX = vx * W, fragY = floorf(vy * H);
    uint32_t vf = (uint32_t)(A.t.frameIndex / 2);
    uint32_t win = vf / 30u, prev = win - 1u;
    float thr = 1.0f - p.dropoutBurstRate;
    float thisA = fhash1(win, 0xDEADBEEFu) < thr ? 0.0f : 1.0f;
    float prevA = fhash1(prev, 0xDEADBEEFu) < thr ? 0.0f : 1.0f;
    float wpos = (float)(vf % 30u) / 30.0f;
    float act = fmaxf(thisA, prevA * sstep(0.7f, 1.0f, 1.0f - wpos));
    if (act < 0.001f) return col;
    float age = 0.05f + p.tapeAge * 0.95f;
    int nEv = 1 + (int)roundf((act * age) * 13.0f);
    float bw = 0.0f, bu = vx, bv = vy, bb = 0.0f;
    for (int i = 0; i < nEv; ++i) {
        uint32_t si = (uint32_t)i;
        float evY = floorf(fhash1(vf * 31u + si, si * 13u + 1u) * H);
        int evH = (fhash1(vf + si * 3u, si + 77u) > 0.88f) ? 2 : 1;
        if (fragY < evY || fragY >= evY + (float)evH) continue;
        float wr = fhash1(vf * 7u + si, si * 7u + 2u);
        float evW = (fhash1(vf * 2u + si, si + 55u) > 0.88f) ? 80.0f + wr * 140.0f : 8.0f + (wr * wr) * 32.0f;
        float maxStart = fmaxf(0.0f, W - evW);
        float x0 = fhash1(vf * 5u + si, si * 7u + 3u) * maxStart;
        if (fragX < x0 || fragX >= x0 + evW) continue;
        float dL = fragX - x0, dR = (x0 + evW) - fragX;
        float w = sstep(0.0f, 1.5f, dL) * sstep(0.0f, 5.0f, dR);
        float dxn = fhash1(vf * 11u + si, si * 7u + 4u) * 2.0f - 1.0f;
        float dx = dxn * (4.0f + p.tapeAge * 16.0f);
        float dyr = fhash1(vf * 13u + si, si * 7u + 5u);
        float dy = (dyr > 0.6f) ? (dyr > 0.8f ? -1.0f : 1.0f) : 0.0f;
        float str = fhash1(vf * 17u + si, si * 7u + 6u);
        float stretch = (str > 0.7f) ? (1.0f + (str * p.tapeAge) * 0.5f) : 1.0f;
        float cx = x0 + evW * 0.5f;
        float sx = (cx + (fragX - cx) * stretch) + dx;
        float burst = sstep(3.0f, 0.0f, dL) * p.dropoutBright;
        if (w > bw) { bw = w; bu = clamp01(sx / W); bv = clamp01((fragY + dy) / H); bb = burst; }
        break;
    }
    if (bw < 0.001f) return col;
    c3 r = s.bilinear(bu, bv);
    float g = 1.0f + bb * 0.3f;
    r = C3(r.r * g, r.g * g, r.b * g);
    float lr = luma709(r), m = bb * 0.4f;
    r = C3(mixf(r.r, lr, m), mixf(r.g, lr, m), mixf(r.b, lr, m));
    return C3(mixf(col.r, r.r, bw), mixf(col.g, r.g, bw), mixf(col.b, r.b, bw));
}

// p13_headswitch.frag:78-191
__device__ __forceinline__ float p13_noise(float x, uint32_t seed) {
    float i = floorf(x), f = fractf(x);
    float u = (f * f) * (3.0f - 2.0f * f);
    return mixf(fhash2((uint32_t)i, seed), fhash2((uint32_t)i + 1u, seed), u) * 2.0f - 1.0f;
}
__device__ c3 pass_p13(const PostArgs &A, int x, int y) {
    const Img &s = A.src;
    const rc_post_params &p = A.p;
    float vx = vu(x, s.w), vy = vu(y, s.h);
    float dist = floorf(vy * (float)s.h);
    float zone = (float)p.headSwitchLines;
    if (dist >= zone) return s.at(x, y);
    float prox = 1.0f - dist / zone;
    float settle = prox * prox;
    uint32_t fs = (uint32_t)A.t.frameIndex;
    float sn = p13_noise(dist * 0.8f, fs), fn = p13_noise(dist * 3.0f, fs + 7u);
    float jit = (p.headSwitchJitter * settle) * (sn * 0.7f + fn * 0.3f);
    float pr = fhash2(fs, 0xCAFEBABEu) * 6.283185f;
    float ripple = (p.headSwitchRipple * settle) * rc_sinf(dist * 1.570796f + pr);
    float sgn = (fhash2(fs, 0xBEEFu) > 0.5f) ? 1.0f : -1.0f;
    float ang = (p.headSwitchChroma * settle) * sgn;
    c3 q = ycc5(s.bilinear(clamp01(vx + jit / (float)s.w), vy));
    float cb = q.g - 0.5f, cr = q.b - 0.5f;
    rot2(cb, cr, ang);
    return rgb5(C3(q.r + ripple, cb + 0.5f, cr + 0.5f));
}

// p14_tracking.frag:83-269
__device__ __forceinline__ float p14_noise(float x, uint32_t seed) {
    float i = floorf(x), f = fractf(x);
    float u = (f * f) * (3.0f - 2.0f * f);
    return mixf(fhash2((uint32_t)i, seed), fhash2((uint32_t)i + 1u, seed), u);
}
__device__ c3 pass_p14(const PostArgs &A, int x, int y) {
    const Img &s = A.src;
    const rc_post_params &p = A.p;
    float W = (float)s.w, H = (float)s.h;
    float vx = vu(x, s.w), vy = vu(y, s.h);
    float fragY = floorf(vy * H);
    c3 col = s.at(x, y);
    if (p.trackingSeverity < 0.001f) return col;
    float sev = p.trackingSeverity;
    uint32_t frame = (uint32_t)A.t.frameIndex;
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
            float base = (((ds * p.trackingMaxDisplacePx) * sev) * (bt - 0.5f)) * 2.0f;
            float fph = (float)frame * 0.37f + (float)ei * 1.57f;
            float wob = ((rc_sinf(bt * 6.28318f + fph) * p.trackingMaxDisplacePx) * 0.20f) * sev;
            float mj = ((p14_noise(fragY * 0.25f, frame + ei * 17u) * 2.0f - 1.0f) * 2.5f) * sev;
            float tot = (base + wob) + mj;
            float dT = fragY - y0, dB = y1 - fragY;
            float ew = sstep(0.0f, 1.5f, dT) * sstep(0.0f, 4.0f, dB);
            float fr = (1.0f - sstep(0.0f, p.trackingFringeWidthPx, fminf(dT, dB))) * ew;
            float hue = ((ds * 0.436f) * sev) * (1.0f - fr);
            if (ew > tW) { tW = ew; tDisp = tot; tFr = fr; tNz = fr * 0.7f; tHue = hue; }
        }
    }
    if (tW < 0.001f) return col;
    c3 q = ycc5(s.bilinear(clamp01(vx + (tDisp / W) * tW), vy));
    if (fabsf(tHue) > 0.001f) {
        float cb = q.g - 0.5f, cr = q.b - 0.5f;
        rot2(cb, cr, tHue);
        q.g = cb + 0.5f;
        q.b = cr + 0.5f;
    }
    if (tFr > 0.001f) {
        float fa = rc_sinf((float)A.t.frameIndex * 0.12f + vy * 8.0f) * 0.9f;
        float cb = q.g - 0.5f, cr = q.b - 0.5f;
        rot2(cb, cr, fa);
        float g = 1.0f + tFr * 1.2f;
        cb *= g;
        cr *= g;
        q.g = mixf(q.g, cb + 0.5f, tFr);
        q.b = mixf(q.b, cr + 0.5f, tFr);
        uint32_t ns = uhash2(((uint32_t)fragY * 1619u) ^ ((uint32_t)A.t.frameIndex * 2654435761u));
        float fx = (float)x + 0.5f;  // gl_FragCoord.x
        float buzz = ((float)uhash2(ns ^ (uint32_t)(fx * 37.0f)) / 4294967295.0f) * 2.0f - 1.0f;
        q.r = q.r + (buzz * 0.15f) * tNz;
    }
    c3 o = rgb5(q);
    return C3(mixf(col.r, o.r, tW), mixf(col.g, o.g, tW), mixf(col.b, o.b, tW));
}

// p15_edge_enhance.frag:80-120
__device__ c3 pass_p15(const PostArgs &A, int x, int y) {
    const float w[5] = {0.0625f, 0.25f, 0.375f, 0.25f, 0.0625f};
    c3 q = ycc5(A.src.at(x, y));
    float Y = q.r, blur = 0.0f;
    for (int dy = -2; dy <= 2; ++dy)
        for (int dx = -2; dx <= 2; ++dx)
            blur += (ycc5(A.src.at(x + dx, y + dy)).r * w[dx + 2]) * w[dy + 2];
    float det = Y - blur;
    det *= sstep(A.p.edgeEnhanceCoreThreshold * 0.5f, A.p.edgeEnhanceCoreThreshold, fabsf(det));
    return rgb5(C3(Y + A.p.edgeEnhanceAmount * det, q.g, q.b));
}

// p16_interlace.frag:81-141
__device__ c3 pass_p16(const PostArgs &A, int x, int y) {
    const Img &s = A.src;
    const rc_post_params &p = A.p;
    float vx = vu(x, s.w), vy = vu(y, s.h);
    float pxx = 1.0f / (float)s.w, pxy = 1.0f / (float)s.h;
    bool even = (y & 1) == 0;
    float vs = even ? ((p.interlaceFieldOffsetPx * 0.5f) * pxy) : ((-p.interlaceFieldOffsetPx * 0.5f) * pxy);
    float yh = luma601(s.at(x, y)), ya = luma601(s.at(x, y - 1));
    float yb = luma601(s.at(x, y + 1)), yl = luma601(s.at(x - 1, y));
    float dlt = yh - yl;
    float hd = dlt > 0.0f ? 1.0f : (dlt < 0.0f ? -1.0f : 0.0f);
    float comb = 0.0f;
    if (even)
        comb = (hd * sstep(p.interlaceCombEdgeThresh, p.interlaceCombEdgeThresh * 3.0f, fabsf(ya - yb))) *
               p.interlaceCombStrength;
    c3 c = s.bilinear(clamp01(vx + comb * pxx), clamp01(vy + vs));
    float m = even ? (1.0f - p.interlaceLineWeighting) : (1.0f + p.interlaceLineWeighting * 0.5f);
    return C3(c.r * m, c.g * m, c.b * m);
}

// display.frag:11-42 with u_totalSamples = 1, u_exposure = 1 (RenderPipeline.java:243-261)
__device__ __forceinline__ float aces1(float x) {
    return rc_clampf((x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f), 0.0f, 1.0f);
}
__device__ __forceinline__ float srgb1(float c) {
    return (c < 0.0031308f) ? 12.92f * c : 1.055f * rc_powf(c, 1.0f / 2.4f) - 0.055f;
}
__device__ __forceinline__ unsigned char unorm8(float v) {
    return (unsigned char)floorf(rc_clampf(v, 0.0f, 1.0f) * 255.0f + 0.5f);
}

enum { P_P00 = 0, P_ATROUS = 100, P_DISPLAY = 200 };

template <int PASS>
__global__ void __launch_bounds__(256) post_kernel(PostArgs A) {
    if (A.dt) A.t = *A.dt;
    int x = blockIdx.x * 16 + (threadIdx.x & 15);
    int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= A.src.w || y >= A.src.h) return;
    size_t i = (size_t)y * (size_t)A.src.w + (size_t)x;
    if (PASS == P_DISPLAY) {
        c3 c = A.src.at(x, y);
        A.rgba8[i] = make_uchar4(unorm8(srgb1(aces1(c.r))), unorm8(srgb1(aces1(c.g))), unorm8(srgb1(aces1(c.b))),
                                 255);
        return;
    }
    c3 o;
    switch (PASS) {
        case P_P00: o = pass_p00(A, x, y); break;
        case P_ATROUS: o = pass_atrous(A, x, y); break;
        case 1: o = pass_p01(A, x, y); break;
        case 2: o = pass_p02(A, x, y); break;
        case 3: o = pass_p03(A, x, y); break;
        case 4: o = pass_p04(A, x, y); break;
        case 7: o = pass_p07(A, x, y); break;
        case 8: o = pass_p08(A, x, y); break;
        case 9: o = pass_p09(A, x, y); break;
        case 10: o = pass_p10(A, x, y); break;
        case 11: o = pass_p11(A, x, y); break;
        case 12: o = pass_p12(A, x, y); break;
        case 13: o = pass_p13(A, x, y); break;
        case 14: o = pass_p14(A, x, y); break;
        case 15: o = pass_p15(A, x, y); break;
        case 16: o = pass_p16(A, x, y); break;
        default: o = A.src.at(x, y); break;
    }
    A.dst[i] = make_float4(o.r, o.g, o.b, 1.0f);
}

typedef void (*post_fn)(PostArgs);

static post_fn kernel_for(int pass) {
    switch (pass) {
        case P_P00: return post_kernel<P_P00>;
        case P_ATROUS: return post_kernel<P_ATROUS>;
        case 1: return post_kernel<1>;
        case 2: return post_kernel<2>;
        case 3: return post_kernel<3>;
        case 4: return post_kernel<4>;
        case 7: return post_kernel<7>;
        case 8: return post_kernel<8>;
        case 9: return post_kernel<9>;
        case 10: return post_kernel<10>;
        case 11: return post_kernel<11>;
        case 12: return post_kernel<12>;
        case 13: return post_kernel<13>;
        case 14: return post_kernel<14>;
        case 15: return post_kernel<15>;
        case 16: return post_kernel<16>;
        case P_DISPLAY: return post_kernel<P_DISPLAY>;
        default: return nullptr;
    }
}

hipError_t launch_post_pass(hipStream_t stream, int pass, const float4 *src, float4 *dst, int w, int h,
                            const rc_post_params &p, const rc_temporal &t, const ushort4 *gbuf, const ushort4 *galbedo,
                            const float *variance, int total, int step, uchar4 *rgba8, const rc_temporal *dt) {
    post_fn k = kernel_for(pass);
    if (!k) return hipErrorInvalidValue;
    PostArgs A;
    A.dt = dt;
    A.src.p = src;
    A.src.w = w;
    A.src.h = h;
    A.dst = dst;
    A.p = p;
    A.t = t;
    A.g.gbuf = gbuf;
    A.g.galbedo = galbedo;
    A.g.variance = variance;
    A.g.total = total;
    A.step = step;
    A.rgba8 = rgba8;
    dim3 grid((w + 15) / 16, (h + 15) / 16);
    hipLaunchKernelGGL(k, grid, dim3(256), 0, stream, A);
    return hipGetLastError();
}

}  // namespace rc
