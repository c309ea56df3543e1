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
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "rc_device.h"
#include "rc_internal.h"

namespace rc {

struct c3 { float r, g, b; };
__device__ __forceinline__ c3 C3(float r, float g, float b) { c3 c; c.r = r; c.g = g; c.b = b; return c; }

// Bilinear sample at (u, v) of any image accessor (at(x, y) with edge clamping, w, h).
template <class S>
__device__ c3 bilinear_of(const S &s, float u, float v) {
    float x = u * (float)s.w - 0.5f, y = v * (float)s.h - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    float ax = x - fx, ay = y - fy;
    int x0 = (int)fx, y0 = (int)fy;
    c3 t00 = s.at(x0, y0), t10 = s.at(x0 + 1, y0), t01 = s.at(x0, y0 + 1), t11 = s.at(x0 + 1, y0 + 1);
    float bx = 1.0f - ax, by = 1.0f - ay;
    c3 r0 = C3(t00.r * bx + t10.r * ax, t00.g * bx + t10.g * ax, t00.b * bx + t10.b * ax);
    c3 r1 = C3(t01.r * bx + t11.r * ax, t01.g * bx + t11.g * ax, t01.b * bx + t11.b * ax);
    return C3(r0.r * by + r1.r * ay, r0.g * by + r1.g * ay, r0.b * by + r1.b * ay);
}

// An RGBA32F image in global memory.
struct Img {
    const float4 *__restrict__ p;
    int w, h;
    __device__ __forceinline__ c3 at(int x, int y) const {
        x = min(max(x, 0), w - 1);
        y = min(max(y, 0), h - 1);
        float4 v = p[(size_t)y * (size_t)w + (size_t)x];
        return C3(v.x, v.y, v.z);
    }
    __device__ c3 bilinear(float u, float v) const { return bilinear_of(*this, u, v); }
};

// The p00 output read on the fly from the accumulation sums: the same
// operations as pass_p00 (p00_normalize.frag:22-27), so a pass reading through
// it sees exactly the values a stored p00 image would hold.
struct P00Img {
    const float4 *__restrict__ p;
    int w, h;
    float n, exposure;
    __device__ __forceinline__ c3 at(int x, int y) const {
        x = min(max(x, 0), w - 1);
        y = min(max(y, 0), h - 1);
        float4 c = p[(size_t)y * (size_t)w + (size_t)x];
        return C3((c.x / n) * exposure, (c.y / n) * exposure, (c.z / n) * exposure);
    }
    __device__ c3 bilinear(float u, float v) const { return bilinear_of(*this, u, v); }
};

// A previous stage's output held in LDS over the region [x0, x0 + rw) x
// [y0, y0 + rh) of a w x h image, one float4 (rgb, unused) per pixel, row stride
// rw. Reads clamp to the image edge first (the chain images' GL_CLAMP_TO_EDGE);
// the region covers every clamped coordinate the reading pass can reach.
struct LdsImg {
    const float4 *q;
    int x0, y0, rw, rh;
    int w, h;
    __device__ __forceinline__ c3 at(int x, int y) const {
        x = min(max(x, 0), w - 1);
        y = min(max(y, 0), h - 1);
        x = min(max(x - x0, 0), rw - 1);
        y = min(max(y - y0, 0), rh - 1);
        const float4 v = q[y * rw + x];
        return C3(v.x, v.y, v.z);
    }
    __device__ c3 bilinear(float u, float v) const { return bilinear_of(*this, u, v); }
};

struct GBufs {
    const ushort4 *gbuf, *galbedo;
    const float *variance;
    int total;
};

// Per-tap weights that depend only on the settings, evaluated once on the host
// with the same routines and operation order as the passes' loops (so each
// entry and each running sum is the value the loop would compute).
struct PostWeights {
    float w01[9];    // p01: gauss(i, max(lumaBlurSigma, 0.01)), i = -4..4
    float wt01;      //      their sum in loop order
    float w02[65];   // p02: gauss(i, max(chromaResSigma, 0.5)), i = -r..r
    float wt02;
    float w08[151];  // p08: pow(0.70, i), i = 0..150 (secondary taps)
    float w09[24];   // p09: (1 - k) k^i, k^i as the running product
    float wt09;
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
    PostWeights wts;
};

// Settings and inputs a pass reads (PostArgs' fields) with the temporal inputs
// resolved once (a device slot for captured frames). Passes read the kernel
// arguments through it; writing to the arguments would move them to scratch.
struct SweepView {
    const rc_post_params &p;
    rc_temporal t;
    const GBufs &g;
    int step;
    const PostWeights &wts;
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
__host__ __device__ __forceinline__ float gaussf(float x, float s) { return rc_expf(-(x * x) / ((2.0f * s) * s)); }
__device__ __forceinline__ void rot2(float &x, float &y, float th) {
    float c = rc_cosf(th), s = rc_sinf(th);
    float nx = x * c - y * s, ny = x * s + y * c;
    x = nx; y = ny;
}

// p00_normalize.frag:22-27
template <class S, class AT>
__device__ c3 pass_p00(const S &src, const AT &A, int x, int y) {
    float n = (float)(A.g.total > 1 ? A.g.total : 1);
    c3 c = src.at(x, y);
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
template <class S, class AT>
__device__ c3 pass_atrous(const S &src, const AT &A, int x, int y) {
    const float hk[5] = {1.0f / 16.0f, 1.0f / 4.0f, 3.0f / 8.0f, 1.0f / 4.0f, 1.0f / 16.0f};
    const S &s = src;
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
template <class S, class AT>
__device__ c3 pass_p01(const S &src, const AT &A, int x, int y) {
    c3 sum = C3(0, 0, 0);
    for (int i = -4; i <= 4; ++i) {
        float w = A.wts.w01[i + 4];  // gauss(i, max(sigma, 0.01))
        c3 c = src.at(x + i, y);
        sum = C3(sum.r + c.r * w, sum.g + c.g * w, sum.b + c.b * w);
    }
    float wt = A.wts.wt01;
    return C3(sum.r / wt, sum.g / wt, sum.b / wt);
}

// p02_chroma_res.frag:55-85
template <class S, class AT>
__device__ c3 pass_p02(const S &src, const AT &A, int x, int y) {
    float sigma = fmaxf(A.p.chromaResSigma, 0.5f);
    int radius = min(max((int)(sigma * 2.5f + 0.5f), 1), 32);
    float Y = ycc5(src.at(x, y)).r;
    float cb = 0.0f, cr = 0.0f;
    for (int i = -radius; i <= radius; ++i) {
        float w = A.wts.w02[i + radius];  // gauss(i, sigma)
        c3 q = ycc5(src.at(x + i, y));
        cb += q.g * w;
        cr += q.b * w;
    }
    float wt = A.wts.wt02;
    return rgb5(C3(Y, cb / wt, cr / wt));
}

// p03_color_matrix.frag:56-90
template <class S, class AT>
__device__ c3 pass_p03(const S &src, const AT &A, int x, int y) {
    const rc_post_params &p = A.p;
    c3 q = ycc3(src.at(x, y));
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
template <class S, class AT>
__device__ c3 pass_p04(const S &src, const AT &A, int x, int y) {
    const rc_post_params &p = A.p;
    c3 q = ycc3(src.at(x, y));
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
template <class S, class AT>
__device__ c3 pass_p07(const S &src, const AT &A, int x, int y) {
    const rc_post_params &p = A.p;
    c3 q = ycc5(src.at(x, y));
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
template <class S, class AT>
__device__ c3 pass_p08(const S &src, const AT &A, int x, int y) {
    const rc_post_params &p = A.p;
    const S &s = src;
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
        sa += fmaxf(0.0f, luma601(s.at(x, y - i)) - p.ccdSmearThreshold) * A.wts.w08[i];  // pow(0.70, i)
        ++st;
    }
    float sec = st > 0 ? ((sa / (float)st) * p.ccdSmearIntensity) * 0.25f : 0.0f;
    float tot = prim + sec;
    return C3(clamp01(base.r + tot), clamp01(base.g + tot), clamp01(base.b + tot));
}

// p09_chroma_bleed.frag:62-94
template <class S, class AT>
__device__ c3 pass_p09(const S &src, const AT &A, int x, int y) {
    float Y = ycc5(src.at(x, y)).r;
    float cb = 0.0f, cr = 0.0f;
    for (int i = 0; i < 24; ++i) {
        c3 q = ycc5(src.at(x - i, y));
        float w = A.wts.w09[i];  // (1 - k) k^i
        cb += q.g * w;
        cr += q.b * w;
    }
    float wt = A.wts.wt09;
    return rgb5(C3(Y, cb / wt, cr / wt));
}

// p10_timebase.frag:44-79
__device__ __forceinline__ float p10_hash(float x) { return fractf(rc_sinf(x * 127.1f + 311.7f) * 43758.5453f); }
template <class S, class AT>
__device__ c3 pass_p10(const S &src, const AT &A, int x, int y) {
    const S &s = src;
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
template <class S, class AT>
__device__ c3 pass_p11(const S &src, const AT &A, int x, int y) {
    const S &s = src;
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
template <class S, class AT>
__device__ c3 pass_p12(const S &src, const AT &A, int x, int y) {
    const S &s = src;
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
template <class S, class AT>
__device__ c3 pass_p13(const S &src, const AT &A, int x, int y) {
    const S &s = src;
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
template <class S, class AT>
__device__ c3 pass_p14(const S &src, const AT &A, int x, int y) {
    const S &s = src;
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
template <class S, class AT>
__device__ c3 pass_p15(const S &src, const AT &A, int x, int y) {
    const float w[5] = {0.0625f, 0.25f, 0.375f, 0.25f, 0.0625f};
    c3 q = ycc5(src.at(x, y));
    float Y = q.r, blur = 0.0f;
    for (int dy = -2; dy <= 2; ++dy)
        for (int dx = -2; dx <= 2; ++dx)
            blur += (ycc5(src.at(x + dx, y + dy)).r * w[dx + 2]) * w[dy + 2];
    float det = Y - blur;
    det *= sstep(A.p.edgeEnhanceCoreThreshold * 0.5f, A.p.edgeEnhanceCoreThreshold, fabsf(det));
    return rgb5(C3(Y + A.p.edgeEnhanceAmount * det, q.g, q.b));
}

// p16_interlace.frag:81-141
template <class S, class AT>
__device__ c3 pass_p16(const S &src, const AT &A, int x, int y) {
    const S &s = src;
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

// The per-tap weights of p01, p02, p08 and p09 for settings p (PostWeights).
static void fill_weights(PostWeights &W, const rc_post_params &p) {
    W = PostWeights{};
    const float s1 = fmaxf(p.lumaBlurSigma, 0.01f);
    float wt = 0.0f;
    for (int i = -4; i <= 4; ++i) {
        W.w01[i + 4] = gaussf((float)i, s1);
        wt += W.w01[i + 4];
    }
    W.wt01 = wt;
    const float s2 = fmaxf(p.chromaResSigma, 0.5f);
    const int r = std::min(std::max((int)(s2 * 2.5f + 0.5f), 1), 32);
    wt = 0.0f;
    for (int i = -r; i <= r; ++i) {
        W.w02[i + r] = gaussf((float)i, s2);
        wt += W.w02[i + r];
    }
    W.wt02 = wt;
    for (int i = 0; i <= 150; ++i) W.w08[i] = rc_powf(0.70f, (float)i);
    const float k = rc_clampf(p.chromaBleedFactor, 0.0f, 0.98f);
    float kp = 1.0f, sc = 1.0f - k;
    wt = 0.0f;
    for (int i = 0; i < 24; ++i) {
        W.w09[i] = sc * kp;
        wt += W.w09[i];
        kp *= k;
    }
    W.wt09 = wt;
}

enum { P_P00 = 0, P_ATROUS = 100, P_DISPLAY = 200 };

// One pass over the image in 16 x 16 tiles. The tiles (row-major) are handed to
// the 8 XCDs in contiguous runs (dispatch slot b goes to XCD b % 8), so each
// XCD works through a band of rows and the neighbourhoods its taps read (the
// a-trous 5 x 5 at step up to 16, smears, blurs) stay in that XCD's L2.
template <int PASS>
__global__ void __launch_bounds__(256) post_kernel(PostArgs A) {
    const SweepView V{A.p, A.dt ? *A.dt : A.t, A.g, A.step, A.wts};
    const int tilesX = (A.src.w + 15) >> 4, tiles = tilesX * ((A.src.h + 15) >> 4);
    const int tile = (int)(blockIdx.x & 7) * ((tiles + 7) >> 3) + (int)(blockIdx.x >> 3);
    if (tile >= tiles) return;
    int x = (tile % tilesX) * 16 + (threadIdx.x & 15);
    int y = (tile / tilesX) * 16 + (threadIdx.x >> 4);
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
        case P_P00: o = pass_p00(A.src, V, x, y); break;
        case P_ATROUS: o = pass_atrous(A.src, V, x, y); break;
        case 1: o = pass_p01(A.src, V, x, y); break;
        case 2: o = pass_p02(A.src, V, x, y); break;
        case 3: o = pass_p03(A.src, V, x, y); break;
        case 4: o = pass_p04(A.src, V, x, y); break;
        case 7: o = pass_p07(A.src, V, x, y); break;
        case 8: o = pass_p08(A.src, V, x, y); break;
        case 9: o = pass_p09(A.src, V, x, y); break;
        case 10: o = pass_p10(A.src, V, x, y); break;
        case 11: o = pass_p11(A.src, V, x, y); break;
        case 12: o = pass_p12(A.src, V, x, y); break;
        case 13: o = pass_p13(A.src, V, x, y); break;
        case 14: o = pass_p14(A.src, V, x, y); break;
        case 15: o = pass_p15(A.src, V, x, y); break;
        case 16: o = pass_p16(A.src, V, x, y); break;
        default: o = A.src.at(x, y); break;
    }
    A.dst[i] = make_float4(o.r, o.g, o.b, 1.0f);
    if (A.rgba8)  // the chain's last pass: the display encode of its output too
        A.rgba8[i] = make_uchar4(unorm8(srgb1(aces1(o.r))), unorm8(srgb1(aces1(o.g))), unorm8(srgb1(aces1(o.b))), 255);
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
    fill_weights(A.wts, p);
    const unsigned tiles = (unsigned)((w + 15) / 16) * (unsigned)((h + 15) / 16);
    hipLaunchKernelGGL(k, dim3(8 * ((tiles + 7) / 8)), dim3(256), 0, stream, A);
    return hipGetLastError();
}

// ---- fused sweeps -----------------------------------------------------------
// The chain of PostProcessStack.runChain (PostProcessStack.java:140-280) runs
// as a few full-frame sweeps instead of one launch per pass. A sweep covers the
// frame in tiles; for each tile its first stage reads a global image (the
// accumulation, through P00Img, for p00; the denoiser's output or the previous
// sweep's otherwise), and each later stage reads the previous stage's output
// from LDS over the region that stage's reads can reach (the tile grown by the
// reach of every later stage). Each stage runs the same per-pixel function as
// the per-pass kernel on the same values, so the output is the per-pass chain's
// bit for bit. The last stage writes the tile to global memory (and the display
// encode when the sweep ends the chain).

constexpr int kMaxSweepStages = 20;
constexpr int kSweepThreads = 512;
constexpr size_t kSweepLdsBudget = 80 * 1024;  // two 512-thread workgroups per CU (160 KB)

// Pass ids of a sweep stage: 0 = p00 (a copy of the first stage's source), 100 =
// a-trous, 1..16 = VHS pN. Bit of a pass in a kernel's pass mask:
__host__ __device__ constexpr uint32_t pass_bit(int pass) { return pass == P_ATROUS ? (1u << 5) : (1u << pass); }
// Kernels are compiled for a few pass sets, so each holds only the registers its
// passes need: p00 + a-trous + the row-local passes, p08..p11, p12..p16, all.
constexpr uint32_t kMaskRow = pass_bit(0) | pass_bit(P_ATROUS) | pass_bit(1) | pass_bit(2) | pass_bit(3) |
                              pass_bit(4) | pass_bit(7);
constexpr uint32_t kMaskMid = pass_bit(0) | pass_bit(8) | pass_bit(9) | pass_bit(10) | pass_bit(11);
constexpr uint32_t kMaskTail = pass_bit(0) | pass_bit(12) | pass_bit(13) | pass_bit(14) | pass_bit(15) | pass_bit(16);
constexpr uint32_t kMaskAll = 0x1FFBFu;  // 0..16 without 6 (p05 and p06 never run)

struct SweepStage {
    int pass;
    int l, r, d, u;  // growth of this stage's output region beyond the tile (left, right, down, up)
};

struct SweepArgs {
    PostArgs A;      // settings, temporal inputs, G-buffers, a-trous step
    int n;
    SweepStage st[kMaxSweepStages];
    int steps[kMaxSweepStages];  // a-trous step of each stage
    int tw, th;      // tile size
    int tilesX, tiles, tilesPerXcd;
    int plane;       // float4 per LDS buffer: the largest region of stages 0..n-2
    float4 *out;     // the last stage's output
    uchar4 *rgba8;   // display encode of the last stage's output, or null
};

template <uint32_t MASK, class S>
__device__ __forceinline__ c3 run_pass(int pass, const S &s, const SweepView &A, int x, int y) {
#define RC_SWEEP_CASE(k, call) \
    case k:                    \
        if constexpr ((MASK & pass_bit(k)) != 0) return call; else break;
    switch (pass) {
        RC_SWEEP_CASE(P_ATROUS, pass_atrous(s, A, x, y))
        RC_SWEEP_CASE(1, pass_p01(s, A, x, y))
        RC_SWEEP_CASE(2, pass_p02(s, A, x, y))
        RC_SWEEP_CASE(3, pass_p03(s, A, x, y))
        RC_SWEEP_CASE(4, pass_p04(s, A, x, y))
        RC_SWEEP_CASE(7, pass_p07(s, A, x, y))
        RC_SWEEP_CASE(8, pass_p08(s, A, x, y))
        RC_SWEEP_CASE(9, pass_p09(s, A, x, y))
        RC_SWEEP_CASE(10, pass_p10(s, A, x, y))
        RC_SWEEP_CASE(11, pass_p11(s, A, x, y))
        RC_SWEEP_CASE(12, pass_p12(s, A, x, y))
        RC_SWEEP_CASE(13, pass_p13(s, A, x, y))
        RC_SWEEP_CASE(14, pass_p14(s, A, x, y))
        RC_SWEEP_CASE(15, pass_p15(s, A, x, y))
        RC_SWEEP_CASE(16, pass_p16(s, A, x, y))
        default: break;
    }
#undef RC_SWEEP_CASE
    return s.at(x, y);  // p00: the first stage's source is read through P00Img
}

template <class S0, uint32_t MASK>
__global__ void __launch_bounds__(kSweepThreads) sweep_kernel(SweepArgs P, S0 src0) {
    extern __shared__ float4 lds[];
    // Workgroups go round-robin to the 8 XCDs; give each XCD a contiguous run of
    // tiles so that neighbouring tiles' halo reads share its L2.
    const int b = blockIdx.x;
    const int tile = (b & 7) * P.tilesPerXcd + (b >> 3);
    if (tile >= P.tiles) return;
    const int W = src0.w, H = src0.h;
    const int ox = (tile % P.tilesX) * P.tw, oy = (tile / P.tilesX) * P.th;
    LdsImg prev{nullptr, 0, 0, 1, 1, W, H};
    for (int k = 0; k < P.n; ++k) {
        const SweepStage S = P.st[k];
        const SweepView V{P.A.p, P.A.dt ? *P.A.dt : P.A.t, P.A.g, P.steps[k], P.A.wts};
        const int rx0 = max(ox - S.l, 0), rx1 = min(ox + P.tw + S.r, W);
        const int ry0 = max(oy - S.d, 0), ry1 = min(oy + P.th + S.u, H);
        const int rw = rx1 - rx0, cnt = rw * (ry1 - ry0);
        const float inv = 1.0f / (float)rw;
        const bool last = k == P.n - 1;
        float4 *o = lds + (k & 1) * P.plane;
        for (int i = threadIdx.x; i < cnt; i += kSweepThreads) {
            int yy = (int)(((float)i + 0.5f) * inv);  // i / rw, corrected below
            if (yy * rw > i) --yy;
            else if ((yy + 1) * rw <= i) ++yy;
            const int x = rx0 + (i - yy * rw), y = ry0 + yy;
            const c3 c = k == 0 ? run_pass<MASK>(S.pass, src0, V, x, y) : run_pass<MASK>(S.pass, prev, V, x, y);
            if (last) {
                const size_t gi = (size_t)y * (size_t)W + (size_t)x;
                P.out[gi] = make_float4(c.r, c.g, c.b, 1.0f);
                if (P.rgba8)
                    P.rgba8[gi] = make_uchar4(unorm8(srgb1(aces1(c.r))), unorm8(srgb1(aces1(c.g))),
                                              unorm8(srgb1(aces1(c.b))), 255);
            } else {
                o[i] = make_float4(c.r, c.g, c.b, 0.0f);
            }
        }
        if (last) break;
        __syncthreads();
        prev = LdsImg{o, rx0, ry0, rw, ry1 - ry0, W, H};
    }
}

// Pixels a pass's output at (x, y) reads from its input, beyond (x, y): left,
// right, down (y-) and up (y+), for these settings. Bilinear reads add one
// texel each way (the sample position is computed in float and can round to
// either side of a texel centre).
struct Reach { int l = 0, r = 0, d = 0, u = 0; };

static int ceil_abs(float v) { return (int)std::ceil(std::fabs((double)v)); }

static int p02_radius(const rc_post_params &p) {
    float sigma = std::fmax(p.chromaResSigma, 0.5f);
    return std::min(std::max((int)(sigma * 2.5f + 0.5f), 1), 32);
}

static Reach pass_reach(int pass, const rc_post_params &p, int step) {
    Reach R;
    auto both = [&](int h, int v) { R.l = R.r = h; R.d = R.u = v; };
    switch (pass) {
        case P_ATROUS: both(2 * step, 2 * step); break;
        case 1: both(4, 0); break;
        case 2: both(p02_radius(p), 0); break;
        case 8:
            R.u = std::min(std::max(p.ccdSmearLength, 0), 150);
            R.d = std::min(std::max(1, (int)((float)p.ccdSmearLength * 0.30f)), 150);
            break;
        case 9: R.l = 23; break;
        case 10: both(ceil_abs(p.timebaseAmplitudePx) + 2, 1); break;
        case 11: both(1, 0); break;
        case 12: {
            // dx up to 4 + 16 age; the stretch moves a pixel of a <= 220 px event by up to
            // 110 * 0.5 |age| from the event centre; the sample row is fragY + dy - 1 or
            // fragY + dy with dy in {-1, 0, 1}
            float age = std::fabs(p.tapeAge);
            both(ceil_abs(4.0f + 16.0f * age + 55.0f * age) + 3, 3);
            break;
        }
        case 13: both(ceil_abs(p.headSwitchJitter) + 2, 1); break;
        case 14: {
            float sev = std::fabs(p.trackingSeverity), m = std::fabs(p.trackingMaxDisplacePx);
            both(ceil_abs(m * sev * 1.2f + 2.5f * sev) + 2, 1);
            break;
        }
        case 15: both(2, 2); break;
        case 16:
            R.l = R.r = std::max(ceil_abs(p.interlaceCombStrength) + 2, 1);
            R.d = R.u = ceil_abs(p.interlaceFieldOffsetPx * 0.5f) + 2;
            break;
        default: break;  // p00, p03, p04, p07: pointwise
    }
    return R;
}

// Relative per-pixel work of a pass (measured per-pass kernel times), for choosing
// tile shapes and groupings.
static float pass_cost(int pass, const rc_post_params &p) {
    switch (pass) {
        case P_ATROUS: return 60.0f;
        case 1: return 3.0f;
        case 2: return 1.0f + 0.5f * (float)p02_radius(p);
        case 8: return 2.0f + 0.6f * (float)std::min(std::max(p.ccdSmearLength, 0), 150);
        case 9: return 6.0f;
        case 12: return 3.0f;
        case 15: return 5.0f;
        case 0: case 3: case 4: case 7: case 13: return 1.0f;
        default: return 2.0f;
    }
}
// Work of reading one pixel of the first stage's input from L2/HBM, in the same units.
constexpr float kReadCost = 1.0f;

struct SweepPlan {
    int first = 0, count = 0;  // ops [first, first + count)
    int tw = 0, th = 0;
    size_t ldsBytes = 0;
    float work = 0.0f;         // modelled work per output pixel
    float ratio = 0.0f;        // that over the passes' own work when run alone
};

// Best tile shape for ops [first, first + count): the shape within the LDS budget
// with the least modelled work (each stage's region times its cost, plus the
// first stage's input footprint). false when no shape fits.
static bool plan_shape(const int *ops, const int *steps, int first, int count, int w, int h, const rc_post_params &p,
                       SweepPlan &out) {
    std::vector<Reach> grow(count);
    Reach g;
    for (int k = count - 1; k >= 0; --k) {
        grow[k] = g;
        Reach r = pass_reach(ops[first + k], p, steps[first + k]);
        if (k == 0) {  // the first stage's reads reach its input in global memory
            g.l += r.l; g.r += r.r; g.d += r.d; g.u += r.u;
            break;
        }
        g.l += r.l; g.r += r.r; g.d += r.d; g.u += r.u;
    }
    const Reach in = g;  // the first stage's input footprint beyond the tile
    const int shapes[][2] = {{w, 1}, {w, 2}, {512, 1}, {256, 1}, {256, 2}, {128, 4}, {128, 8}, {64, 8},
                             {64, 16}, {32, 16}, {32, 32}, {64, 32}, {16, 16}, {16, 32}};
    bool found = false;
    double own = 0.0;
    for (int k = 0; k < count; ++k) own += pass_cost(ops[first + k], p);
    own += kReadCost;
    for (const auto &sh : shapes) {
        const int tw = std::min(sh[0], w), th = std::min(sh[1], h);
        if (tw <= 0 || th <= 0) continue;
        const double tile = (double)tw * (double)th;
        auto region = [&](const Reach &e) {
            const int rw = tw >= w ? w : std::min(tw + e.l + e.r, w);
            const int rh = th >= h ? h : std::min(th + e.d + e.u, h);
            return (size_t)rw * (size_t)rh;
        };
        size_t maxRegion = 0;
        double work = kReadCost * (double)region(in);
        for (int k = 0; k < count; ++k) {
            const size_t rg = region(grow[k]);
            if (k < count - 1) maxRegion = std::max(maxRegion, rg);
            work += pass_cost(ops[first + k], p) * (double)rg;
        }
        const size_t lds = maxRegion * 2 * sizeof(float4);
        if (lds > kSweepLdsBudget) continue;
        const float per = (float)(work / tile);
        if (!found || per < out.work - 1e-6f) {
            out.first = first;
            out.count = count;
            out.tw = tw;
            out.th = th;
            out.ldsBytes = lds;
            out.work = per;
            out.ratio = (float)(work / (tile * own));
            found = true;
        }
    }
    return found;
}

// Cost of a sweep launch (ramp-up, drain, one full-frame write and read) in
// work units per pixel: a pass joins the running sweep when that saves work.
constexpr float kLaunchCost = 6.0f;

// Groups the chain's ops (p00 and VHS passes) into sweeps: each next op joins
// the running sweep if some tile shape fits the LDS budget and the joint
// sweep's modelled work is below the two sweeps' work plus a launch.
std::vector<SweepPlan> plan_sweeps(const int *ops, const int *steps, int nOps, int w, int h,
                                   const rc_post_params &p) {
    std::vector<SweepPlan> plans;
    int i = 0;
    while (i < nOps) {
        SweepPlan cur;
        if (!plan_shape(ops, steps, i, 1, w, h, p, cur)) {  // a single stage needs no LDS
            cur.first = i; cur.count = 1; cur.tw = std::min(64, w); cur.th = std::min(16, h);
        }
        int j = i + 1;
        while (j < nOps && j - i < kMaxSweepStages) {
            SweepPlan joint, alone;
            if (!plan_shape(ops, steps, i, j - i + 1, w, h, p, joint)) break;
            if (!plan_shape(ops, steps, j, 1, w, h, p, alone)) break;
            if (joint.work > cur.work + alone.work + kLaunchCost) break;
            cur = joint;
            ++j;
        }
        plans.push_back(cur);
        i = j;
    }
    return plans;
}

static uint32_t smallest_mask(uint32_t need) {
    for (uint32_t m : {kMaskRow, kMaskMid, kMaskTail})
        if ((need & ~m) == 0) return m;
    return kMaskAll;
}

template <class S0>
static void launch_sweep_kernel(uint32_t mask, unsigned grid, size_t lds, hipStream_t stream, const SweepArgs &P,
                                const S0 &s0) {
    switch (mask) {
        case kMaskRow: hipLaunchKernelGGL((sweep_kernel<S0, kMaskRow>), dim3(grid), dim3(kSweepThreads), lds, stream, P, s0); break;
        case kMaskMid: hipLaunchKernelGGL((sweep_kernel<S0, kMaskMid>), dim3(grid), dim3(kSweepThreads), lds, stream, P, s0); break;
        case kMaskTail: hipLaunchKernelGGL((sweep_kernel<S0, kMaskTail>), dim3(grid), dim3(kSweepThreads), lds, stream, P, s0); break;
        default: hipLaunchKernelGGL((sweep_kernel<S0, kMaskAll>), dim3(grid), dim3(kSweepThreads), lds, stream, P, s0); break;
    }
}

hipError_t launch_post_sweeps(hipStream_t stream, const SweepSource &src, const int *opsIn, const int *stepsIn,
                              int nOpsIn, int w, int h, float4 *bufA, float4 *bufB, uchar4 *rgba8,
                              const rc_post_params &p, const rc_temporal &t, const rc_temporal *dt,
                              const ushort4 *gbuf, const ushort4 *galbedo, const float *variance, int total,
                              const float4 **outFinal, int *launches) {
    const float4 *cur = src.p;
    bool p00 = src.p00;
    int nl = 0, k0 = 0;
    hipError_t e;
    auto next = [&]() { return (cur == bufA) ? bufB : bufA; };
    // a-trous passes (they follow p00 directly): each reads 25 taps up to 32 pixels
    // away in both directions and G-buffer taps per tap, so each runs alone on its
    // 16 x 16-tile kernel, which keeps more waves in flight than a sweep kernel
    if (nOpsIn > 0 && opsIn[0] == P_ATROUS) {
        if (p00) {
            float4 *dst = next();
            if ((e = launch_post_pass(stream, 0, cur, dst, w, h, p, t, gbuf, galbedo, variance, total, 0, nullptr, dt)))
                return e;
            ++nl;
            cur = dst;
            p00 = false;
        }
        while (k0 < nOpsIn && opsIn[k0] == P_ATROUS) {
            float4 *dst = next();
            uchar4 *enc = (k0 + 1 == nOpsIn) ? rgba8 : nullptr;  // the chain's last pass encodes the display too
            if ((e = launch_post_pass(stream, P_ATROUS, cur, dst, w, h, p, t, gbuf, galbedo, variance, total,
                                      stepsIn[k0], enc, dt)))
                return e;
            ++nl;
            cur = dst;
            ++k0;
        }
    }
    // the rest in sweeps (p00 first when it is still pending)
    int ops[kMaxSweepStages + 8], steps[kMaxSweepStages + 8], nOps = 0;
    if (p00) { ops[nOps] = 0; steps[nOps++] = 0; }
    for (int k = k0; k < nOpsIn && nOps < kMaxSweepStages + 8; ++k) { ops[nOps] = opsIn[k]; steps[nOps++] = stepsIn[k]; }
    if (nOps == 0) {  // the chain ended with a-trous, whose kernel encoded the display
        if (outFinal) *outFinal = cur;
        if (launches) *launches = nl;
        return hipSuccess;
    }
    std::vector<SweepPlan> plans = plan_sweeps(ops, steps, nOps, w, h, p);
    PostWeights wts;
    fill_weights(wts, p);
    for (size_t s = 0; s < plans.size(); ++s) {
        const SweepPlan &pl = plans[s];
        float4 *dst = next();
        const bool lastSweep = s + 1 == plans.size();
        if (pl.count == 1) {  // one pass: its own kernel (with the display encode when it ends the chain)
            if ((e = launch_post_pass(stream, ops[pl.first], cur, dst, w, h, p, t, gbuf, galbedo, variance, total,
                                      steps[pl.first], lastSweep ? rgba8 : nullptr, dt)))
                return e;
            ++nl;
            cur = dst;
            p00 = false;
            continue;
        }
        SweepArgs P;
        P.A.dt = dt;
        P.A.src.p = nullptr;
        P.A.src.w = w;
        P.A.src.h = h;
        P.A.dst = nullptr;
        P.A.p = p;
        P.A.t = t;
        P.A.g.gbuf = gbuf;
        P.A.g.galbedo = galbedo;
        P.A.g.variance = variance;
        P.A.g.total = total;
        P.A.step = 0;
        P.A.rgba8 = nullptr;
        P.A.wts = wts;
        P.n = pl.count;
        uint32_t need = 0;
        Reach g;  // each stage's region growth: the reach of every later stage of the sweep
        for (int k = pl.count - 1; k >= 0; --k) {
            const int op = ops[pl.first + k];
            P.st[k].pass = op;
            P.steps[k] = steps[pl.first + k];
            P.st[k].l = g.l; P.st[k].r = g.r; P.st[k].d = g.d; P.st[k].u = g.u;
            need |= pass_bit(op);
            if (k > 0) {
                Reach r = pass_reach(op, p, steps[pl.first + k]);
                g.l += r.l; g.r += r.r; g.d += r.d; g.u += r.u;
            }
        }
        P.tw = pl.tw;
        P.th = pl.th;
        P.tilesX = (w + pl.tw - 1) / pl.tw;
        P.tiles = P.tilesX * ((h + pl.th - 1) / pl.th);
        P.tilesPerXcd = (P.tiles + 7) / 8;
        P.plane = (int)(pl.ldsBytes / (2 * sizeof(float4)));
        P.out = dst;
        P.rgba8 = lastSweep ? rgba8 : nullptr;
        const unsigned grid = 8u * (unsigned)P.tilesPerXcd;
        const uint32_t mask = smallest_mask(need);
        if (p00) {
            P00Img s0{cur, w, h, (float)(total > 1 ? total : 1), p.exposure};
            launch_sweep_kernel(mask, grid, pl.ldsBytes, stream, P, s0);
        } else {
            Img s0{cur, w, h};
            launch_sweep_kernel(mask, grid, pl.ldsBytes, stream, P, s0);
        }
        if ((e = hipGetLastError()) != hipSuccess) return e;
        ++nl;
        cur = dst;
        p00 = false;
    }
    if (outFinal) *outFinal = cur;
    if (launches) *launches = nl;
    return hipSuccess;
}

}  // namespace rc
