// rc_net.hip — learned denoiser on gfx950: a 5-level U-Net (3x3 convolutions,
// 32..128 channels, 482,787 parameters) with hand-written forward and backward
// kernels. Project-new (the reference has only the a-trous filter,
// AtrousDenoiser.java); its inputs are the reference's render outputs
// (pathtrace.comp:6-7, 644-647): radiance, first-hit albedo, shading normal and
// hit distance.
//
//   e1 = relu(conv(f))            12 -> 32    level 0 (full resolution)
//   e2 = relu(conv(pool(e1)))     32 -> 64    level 1
//   e3 = relu(conv(pool(e2)))     64 -> 64    level 2
//   e4 = relu(conv(pool(e3)))     64 -> 96    level 3
//   e5 = relu(conv(pool(e4)))     96 -> 128   level 4
//   d4 = relu(conv([up(e5), e4])) 224 -> 64   level 3
//   d3 = relu(conv([up(d4), e3])) 128 -> 64   level 2
//   d2 = relu(conv([up(d3), e2])) 128 -> 32   level 1
//   d1 = relu(conv([up(d2), e1]))  64 -> 32   level 0
//   x  = m * expm1(min(conv1x1(d1) + f[0:3], 20)),  m = max(albedo, 0.01)
//        (the loss sees x as is; the denoised image is max(x, 0))
// pool = 2x2 mean, up = nearest, zero 'same' padding. Features f: log1p of the
// albedo-demodulated radiance (radiance / m), albedo, normal, log1p(depth)/4,
// log1p of the standard error of the mean luma, log2(spp)/10 (channels 12..31
// are zero padding). Raw inputs per pixel (12 floats): radiance rgb, albedo
// rgb, normal xyz, depth, mean of luma^2 over the samples, sample count.
// Loss = mean (x - r)^2 / (r^2 + 0.01) against a high-spp render r.
//
// Precision: activations, features and weights are bfloat16 in HBM (NHWC,
// channels innermost); every 3x3 convolution is an implicit GEMM on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation; bias, ReLU, pooling,
// gradients of activations (FP32 in HBM, rounded to bfloat16 only as MFMA
// operands), the head and the optimizer (Adam on FP32 master weights) are FP32.
//
// Kernels:
//   conv3x3_fwd<NT,MT,G> out = act(W * [a | b] + bias); the input is the channel
//                       concatenation of a (optionally nearest-upsampled) and b,
//                       read in place — no concat or upsample buffers. Also the
//                       input gradient: dX = W' * dZ with W' the tap-flipped,
//                       transposed weights.
//   conv3x3_tile<CIN,COUT,TH,G,WV>  the forward convolution of the layers of at
//                       most 128 input channels: weight fragments in registers,
//                       each block's input tile staged in LDS (same K order and
//                       operands: bit-identical outputs).
//   conv3x3_dw          dW[co][tap][ci] = sum_p dZ[p][co] X[p+tap][ci]: a GEMM
//                       with K = pixels, operands transposed through LDS,
//                       per-chunk partials reduced in a fixed order
//                       (deterministic).
//   head                1x1 conv + residual + expm1, loss and its gradient;
//                       head_ensemble averages the passes of a flip ensemble.
#include <cmath>
#include <cstring>
#include <vector>

#include "rc_internal.h"

namespace rc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kNumLayers = 10;
constexpr int kFeatC = 32;  // padded feature channels (12 used)
constexpr int kRawC = 12;   // raw input floats per pixel

// name, cout, cin (real), cin (padded), kernel, level, source a (layer or -1
// for the features), a upsampled?, a pooled?, source b (skip layer or -1)
struct LayerSpec {
    int cout, cinReal, cinPad, k, level;
    int srcA, aUp, aPool, ca;
    int srcB, cb;
};
// layer ids: 0 e1, 1 e2, 2 e3, 3 e4, 4 e5, 5 d4, 6 d3, 7 d2, 8 d1, 9 out
constexpr LayerSpec kLayers[kNumLayers] = {
    {32, 12, 32, 3, 0, -1, 0, 0, 32, -1, 0},
    {64, 32, 32, 3, 1, 0, 0, 1, 32, -1, 0},
    {64, 64, 64, 3, 2, 1, 0, 1, 64, -1, 0},
    {96, 64, 64, 3, 3, 2, 0, 1, 64, -1, 0},
    {128, 96, 96, 3, 4, 3, 0, 1, 96, -1, 0},
    {64, 224, 224, 3, 3, 4, 1, 0, 128, 3, 96},
    {64, 128, 128, 3, 2, 5, 1, 0, 64, 2, 64},
    {32, 128, 128, 3, 1, 6, 1, 0, 64, 1, 64},
    {32, 64, 64, 3, 0, 7, 1, 0, 32, 0, 32},
    {3, 32, 32, 1, 0, 8, 0, 0, 32, -1, 0},
};

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf16x8 load8(const uint16_t *p) {
    return __builtin_bit_cast(bf16x8, *(const u32x4 *)p);
}
__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, u32x4{0u, 0u, 0u, 0u}); }

// ---------------------------------------------------------------------------
// 3x3 convolution, implicit GEMM: M = output pixels, N = output channels,
// K = 9 taps x input channels (multiple of 32). A wave computes 16*MT pixels x
// 16*NT channels; the 4 waves of a block cover 4/G pixel tiles x G channel
// groups (cout = 16 NT G), so a small level can run G times as many waves.
// Each output's K order (tap, then 32-channel chunk) does not depend on NT, MT
// or G. MFMA 16x16x32 operand layout (lane l): A row l%16, k 8*(l/16)+j;
// B column l%16, k 8*(l/16)+j; D column l%16, rows 4*(l/16)+r.
struct ConvArgs {
    const uint16_t *a;  // [pixels of a][ca]   (a at half resolution when aUp)
    const uint16_t *b;  // [N*H*W][cb] or null
    int ca, cb, aUp;
    int N, H, W;
    const uint16_t *w;  // [cout][9][ca+cb]
    const float *bias;  // [cout] or null
    int cout, relu;
    uint16_t *outB;     // bf16 output, or
    float *outF;        // f32 output
    const uint16_t *zero;  // >= 8 zero bf16 (conv3x3_tile's padding source)
    uint16_t *pool;        // conv3x3_tile: also the 2 x 2 mean of the output (pool2), or null
    // mirW > 0: a holds ONE image (N = 1 layout) of a w x h frame padded to W x H
    // by edge replication; image n of the batch is that frame mirrored by n (bit
    // 0 left-right, bit 1 top-bottom) and, for bit 2, moved by kEnsembleShift
    // pixels right and down, read in place
    int mirW, mirH;
};

// The diagonal move of the ensemble's images 4..7: half the 16-pixel period of
// the network's coarsest level, so that their pooling windows straddle those of
// images 0..3.
constexpr int kEnsembleShift = 8;

// Pixel of the single stored image that image n's pixel (y, x) reads when the
// batch is the ensemble of one padded w x h frame (ConvArgs::mirW): the frame's
// pixel (x, y), or (x - s, y - s) for images 4..7, clamped to the frame (edge
// replication), then mirrored.
__device__ __forceinline__ size_t mirror_src(int n, int y, int x, int W, int w, int h) {
    if (n & 4) {
        x = max(x - kEnsembleShift, 0);
        y = max(y - kEnsembleShift, 0);
    }
    int xb = min(x, w - 1), yb = min(y, h - 1);
    if (n & 1) xb = w - 1 - xb;
    if (n & 2) yb = h - 1 - yb;
    return (size_t)yb * W + xb;
}

template <int NT, int MT, int G>
__global__ void __launch_bounds__(256) conv3x3_fwd(ConvArgs A) {
    static_assert(G == 1 || G == 2 || G == 4, "channel groups per block");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int npix = A.N * A.H * A.W;
    const int base = (blockIdx.x * (4 / G) + wave / G) * 16 * MT;
    const int co0 = (wave % G) * 16 * NT;  // this wave's first output channel
    if (base >= npix) return;
    const int cin = A.ca + A.cb;
    const int kq = 8 * (lane >> 4);  // this lane's k offset inside a 32-wide chunk
    f32x4 acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    int pn[MT], py[MT], px[MT];
    bool pv[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) {
        int p = base + m * 16 + (lane & 15);
        pv[m] = p < npix;
        int pp = pv[m] ? p : 0;
        px[m] = pp % A.W;
        int q = pp / A.W;
        py[m] = q % A.H;
        pn[m] = q / A.H;
    }
    const int Hh = A.H >> 1, Wh = A.W >> 1;
    for (int tap = 0; tap < 9; ++tap) {
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
        const uint16_t *srcA[MT];
        const uint16_t *srcB[MT];
        bool inb[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            int ys = py[m] + dy, xs = px[m] + dx;
            inb[m] = pv[m] && ys >= 0 && ys < A.H && xs >= 0 && xs < A.W;
            ys = inb[m] ? ys : 0;
            xs = inb[m] ? xs : 0;
            size_t pa = A.aUp    ? ((size_t)pn[m] * Hh + (ys >> 1)) * Wh + (xs >> 1)
                        : A.mirW ? mirror_src(pn[m], ys, xs, A.W, A.mirW, A.mirH)
                                 : ((size_t)pn[m] * A.H + ys) * A.W + xs;
            srcA[m] = A.a + pa * A.ca + kq;
            srcB[m] = A.b ? A.b + (((size_t)pn[m] * A.H + ys) * A.W + xs) * A.cb + kq : nullptr;
        }
        const uint16_t *wt = A.w + ((size_t)(co0 + (lane & 15)) * 9 + tap) * cin + kq;
        for (int c0 = 0; c0 < cin; c0 += 32) {
            bf16x8 av[MT];
#pragma unroll
            for (int m = 0; m < MT; ++m) {
                const uint16_t *s = c0 < A.ca ? srcA[m] + c0 : srcB[m] + (c0 - A.ca);
                av[m] = inb[m] ? load8(s) : zero8();
            }
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                bf16x8 bv = load8(wt + (size_t)t * 16 * 9 * cin + c0);
#pragma unroll
                for (int m = 0; m < MT; ++m)
                    acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[m], bv, acc[m][t], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int co = co0 + t * 16 + (lane & 15);
        const float bb = A.bias ? A.bias[co] : 0.0f;
#pragma unroll
        for (int m = 0; m < MT; ++m) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int p = base + m * 16 + 4 * (lane >> 4) + r;
                if (p >= npix) continue;
                float v = acc[m][t][r] + bb;
                if (A.relu) v = fmaxf(v, 0.0f);
                if (A.outB) A.outB[(size_t)p * A.cout + co] = f2bf(v);
                else A.outF[(size_t)p * A.cout + co] = v;
            }
        }
    }
}

// The same convolution for every layer of at most 128 input channels, tiled
// and persistent. A tile is TH x 16 output pixels of one image; a block's WV
// waves are G output-channel groups x WV/G pixel groups: wave w takes channels
// [(w % G) COUT/G, +COUT/G) and tile rows (w / G) MT .. + MT - 1 (MT = TH G / WV),
// 16 pixels each.
//
// Each wave loads its weight fragments (9 taps x CIN/32 chunks x COUT/(16 G)
// channel tiles) into registers once; the block then walks its tiles. A tile's
// (TH + 2) x 18 input halo, every input channel (zero outside the image; the a
// source upsampled in place), is copied from global memory straight into LDS
// (global_load_lds_dwordx4: lane i of a wave writes LDS bytes base + 16 i), into
// one of two buffers, while the MFMAs run on the other: the next tile's halo is
// in flight during the current tile's compute. Halo pixels are CIN + 8 channels
// apart in LDS (the 8 padding channels against bank conflicts of the 16 pixel
// lanes are filled from a zero buffer, like the padding outside the image).
// Each step's A fragments are read from LDS one step ahead. With the MFMA's
// operands exchanged (weights as A, pixels as B) a lane ends with 4 consecutive
// channels of one pixel, stored as one 8-byte write.
//
// The K order per output (tap, then 32-channel chunk) and the operands are those
// of conv3x3_fwd (with the MFMA's A and B operands exchanged, which yields the
// transposed tile with every element bit for bit), so the outputs are
// bit-identical. Tiles are handed to the 8 XCDs in contiguous runs (block b runs
// on XCD b % 8 and takes tiles of that XCD's run), so neighbouring tiles share
// their halo rows in one XCD's L2.
constexpr int kTileX = 16, kHaloX = kTileX + 2;

template <int CIN, int COUT, int TH, int THREADS>
struct TileShape {
    static constexpr int HY = TH + 2;
    static constexpr int PERP = CIN / 8 + 1;  // 16-byte items per halo pixel, padding included
    static constexpr int ITEMS = HY * kHaloX * PERP;
    static constexpr int ROUNDS = (ITEMS + THREADS - 1) / THREADS;
    static constexpr int HALO_BYTES = ROUNDS * THREADS * 16;  // one halo buffer
    static constexpr int LDS_BYTES = 2 * HALO_BYTES;
};

// s_waitcnt with vmcnt = n (0..63) and the other counters left alone.
#define RC_WAIT_VMCNT(n) __builtin_amdgcn_s_waitcnt(((n) & 15) | (((n) >> 4) << 14) | (7 << 4) | (15 << 8))
// s_waitcnt lgkmcnt(0), the other counters left alone.
#define RC_WAIT_LGKMCNT0() __builtin_amdgcn_s_waitcnt(15 | (3 << 14) | (7 << 4))

template <int CIN, int COUT, int TH, int G, int WV>
__global__ void __launch_bounds__(64 * WV) conv3x3_tile(ConvArgs A) {
    static_assert(CIN % 32 == 0 && CIN <= 128 && COUT % (16 * G) == 0, "channels");
    static_assert(WV % G == 0 && (TH * G) % WV == 0, "waves and rows per wave");
    constexpr int THREADS = 64 * WV;
    using TS = TileShape<CIN, COUT, TH, THREADS>;
    constexpr int NTW = COUT / (16 * G);  // channel tiles per wave
    constexpr int MT = TH * G / WV;       // pixel rows per wave
    constexpr int RS = CIN + 8;           // LDS stride (bf16) of a halo pixel
    constexpr int PERP = TS::PERP;
    constexpr int KC = CIN / 32, STEPS = 9 * KC;
    static_assert(2 * TS::ROUNDS <= 63, "vmcnt range");
    // the two halo buffers, as distinct arrays so that the compiler can tell the
    // copies into one buffer from the reads of the other
    __shared__ __align__(16) uint16_t halo0[TS::HALO_BYTES / 2];
    __shared__ __align__(16) uint16_t halo1[TS::HALO_BYTES / 2];
    const int tilesX = (A.W + kTileX - 1) / kTileX, tilesY = (A.H + TH - 1) / TH;
    const int tiles = tilesX * tilesY * A.N;
    const int per = (tiles + 7) >> 3;                    // tiles of one XCD's run
    const int runBegin = (blockIdx.x & 7) * per;
    const int runEnd = min(runBegin + per, tiles);
    const int stride = gridDim.x >> 3;                  // blocks on this XCD
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int cg = wave % G, pg = wave / G;
    const int kq = 8 * (lane >> 4);  // this lane's k offset inside a 32-wide chunk
    const int Hh = A.H >> 1, Wh = A.W >> 1;

    // starts the copies of tile `tile`'s halo into LDS buffer `dst`
    auto fetch = [&](int tile, uint16_t *dst) __attribute__((always_inline)) {
        const int tx0 = (tile % tilesX) * kTileX;
        const int q = tile / tilesX;
        const int ty0 = (q % tilesY) * TH, n = q / tilesY;
#pragma unroll
        for (int r = 0; r < TS::ROUNDS; ++r) {
            const int it = r * THREADS + (int)threadIdx.x;
            const int hp = it / PERP, c = (it - hp * PERP) * 8;
            const int hy = hp / kHaloX, hx = hp - hy * kHaloX;
            const int ys = ty0 + hy - 1, xs = tx0 + hx - 1;
            const uint16_t *src = A.zero;
            if (it < TS::ITEMS && c < CIN && ys >= 0 && ys < A.H && xs >= 0 && xs < A.W) {
                if (c < A.ca) {
                    const size_t pa = A.aUp    ? ((size_t)n * Hh + (ys >> 1)) * Wh + (xs >> 1)
                                      : A.mirW ? mirror_src(n, ys, xs, A.W, A.mirW, A.mirH)
                                               : ((size_t)n * A.H + ys) * A.W + xs;
                    src = A.a + pa * A.ca + c;
                } else {
                    src = A.b + (((size_t)n * A.H + ys) * A.W + xs) * A.cb + (c - A.ca);
                }
            }
            __builtin_amdgcn_global_load_lds(
                (const __attribute__((address_space(1))) void *)src,
                (__attribute__((address_space(3))) void *)(dst + (size_t)(r * THREADS + wave * 64) * 8),
                16, 0, 0);
        }
    };

    // weight fragments of this wave's channels, all steps
    bf16x8 wf[STEPS][NTW];
    const uint16_t *wl = A.w + (size_t)(cg * NTW * 16 + (lane & 15)) * 9 * CIN + kq;
#pragma unroll
    for (int st = 0; st < STEPS; ++st)
#pragma unroll
        for (int t = 0; t < NTW; ++t) wf[st][t] = load8(wl + (size_t)t * 16 * 9 * CIN + (st / KC) * CIN + (st % KC) * 32);
    // this lane's biases: channels co .. co + 3 of each channel tile
    float bias[NTW][4];
#pragma unroll
    for (int t = 0; t < NTW; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) bias[t][r] = A.bias ? A.bias[(cg * NTW + t) * 16 + 4 * (lane >> 4) + r] : 0.0f;
    int tile = runBegin + (int)(blockIdx.x >> 3);
    if (tile < runEnd) fetch(tile, halo0);
    if (tile + stride < runEnd) {
        fetch(tile + stride, halo1);
        RC_WAIT_VMCNT(TS::ROUNDS);  // weights and the first halo
    } else {
        RC_WAIT_VMCNT(0);
    }

    // one tile: compute from `halo` (then refilled with the tile two on)
    auto step = [&](int tile, uint16_t *halo) __attribute__((always_inline)) {
        // this wave's copies of `tile` have landed (those of the next tile may
        // be in flight); the barrier makes every wave's visible
        if (tile + stride < runEnd) RC_WAIT_VMCNT(TS::ROUNDS);
        else RC_WAIT_VMCNT(0);
        __builtin_amdgcn_s_barrier();

        f32x4 acc[MT][NTW];
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int t = 0; t < NTW; ++t) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
        const uint16_t *hl = halo + ((size_t)(pg * MT) * kHaloX + (lane & 15)) * RS + kq;
        auto asrc = [&](int st, int m) {
            const int tap = st / KC;
            return hl + ((size_t)(m + tap / 3) * kHaloX + tap % 3) * RS + (st % KC) * 32;
        };
        bf16x8 av[2][MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) av[0][m] = load8(asrc(0, m));
#pragma unroll
        for (int st = 0; st < STEPS; ++st) {
            if (st + 1 < STEPS) {
#pragma unroll
                for (int m = 0; m < MT; ++m) av[(st + 1) & 1][m] = load8(asrc(st + 1, m));
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int t = 0; t < NTW; ++t)
#pragma unroll
                for (int m = 0; m < MT; ++m)
                    acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf[st][t], av[st & 1][m], acc[m][t], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }

        {
            const int tx0 = (tile % tilesX) * kTileX;
            const int q = tile / tilesX;
            const int ty0 = (q % tilesY) * TH, n = q / tilesY;
            const int x = tx0 + (lane & 15);
#pragma unroll
            for (int t = 0; t < NTW; ++t) {
                const int co = (cg * NTW + t) * 16 + 4 * (lane >> 4);
                uint16_t b[MT][4];  // the stored bfloat16 outputs
#pragma unroll
                for (int m = 0; m < MT; ++m) {
                    const int y = ty0 + pg * MT + m;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        float v = acc[m][t][r] + bias[t][r];
                        if (A.relu) v = fmaxf(v, 0.0f);
                        b[m][r] = f2bf(v);
                    }
                    uint2 qq;
                    qq.x = (uint32_t)b[m][0] | ((uint32_t)b[m][1] << 16);
                    qq.y = (uint32_t)b[m][2] | ((uint32_t)b[m][3] << 16);
                    if (x < A.W && y < A.H) *(uint2 *)(A.outB + (((size_t)n * A.H + y) * A.W + x) * COUT + co) = qq;
                }
                // 2 x 2 mean of the stored values for the next level, as pool2
                // computes it: ((y, x) + (y, x + 1)) + ((y + 1, x) + (y + 1, x + 1)),
                // the x + 1 column taken from the neighbouring lane
                if (A.pool && (MT % 2) == 0) {
                    const int Ho = A.H >> 1, Wo = A.W >> 1;
#pragma unroll
                    for (int m = 0; m + 1 < MT; m += 2) {
                        const int yo = (ty0 + pg * MT + m) >> 1, xo = x >> 1;
                        float s[4];
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const float a0 = bf2f(b[m][r]), b0 = bf2f(b[m + 1][r]);
                            const float a1 = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(a0), 0xB1, 0xF, 0xF, false));
                            const float b1 = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(b0), 0xB1, 0xF, 0xF, false));
                            s[r] = ((a0 + a1) + (b0 + b1)) * 0.25f;
                        }
                        uint2 qq;
                        qq.x = (uint32_t)f2bf(s[0]) | ((uint32_t)f2bf(s[1]) << 16);
                        qq.y = (uint32_t)f2bf(s[2]) | ((uint32_t)f2bf(s[3]) << 16);
                        if (!(lane & 1) && xo < Wo && yo < Ho)
                            *(uint2 *)(A.pool + (((size_t)n * Ho + yo) * Wo + xo) * COUT + co) = qq;
                    }
                }
            }
        }
        RC_WAIT_LGKMCNT0();            // this wave's halo reads are done
        __builtin_amdgcn_s_barrier();  // ... and every wave's
        // the buffer just consumed takes the halo two tiles on
        if (tile + 2 * stride < runEnd) fetch(tile + 2 * stride, halo);
    };
    while (tile < runEnd) {
        step(tile, halo0);
        tile += stride;
        if (tile >= runEnd) break;
        step(tile, halo1);
        tile += stride;
    }
}

// ---------------------------------------------------------------------------
// Weight gradient of a 3x3 convolution. Block (chunk, tap): pixels
// [chunk*pixPerChunk, +pixPerChunk) staged kDwStep at a time and fed to the
// MFMAs as 32-pixel K chunks in pixel order; dZ and the tap-shifted input are
// staged in LDS transposed ([channel][pixel]) so that each lane's 8
// consecutive K values are one 16-byte LDS read.
struct DwArgs {
    const uint16_t *dz;  // [npix][cout]
    const uint16_t *a;
    const uint16_t *b;
    int ca, cb, aUp;
    int N, H, W;
    int cout;
    int pixPerChunk;
    float *part;         // [chunks][cout][9][cin]
};
constexpr int kDwStep = 64;    // pixels staged per LDS round: two 32-pixel MFMA K chunks
constexpr int kDwPad = kDwStep + 8;  // LDS row length in bf16 (144-byte rows, 16-byte aligned)
constexpr int kDwMaxTiles = 16;

__global__ void __launch_bounds__(256) conv3x3_dw(DwArgs A) {
    __shared__ __attribute__((aligned(16))) uint16_t dzs[128 * kDwPad];
    __shared__ __attribute__((aligned(16))) uint16_t xs[224 * kDwPad];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tap = blockIdx.y, dy = tap / 3 - 1, dx = tap % 3 - 1;
    const int cin = A.ca + A.cb;
    const int npix = A.N * A.H * A.W;
    const int p0 = blockIdx.x * A.pixPerChunk;
    const int p1 = min(npix, p0 + A.pixPerChunk);
    const int ntc = cin / 16, nmt = A.cout / 16;
    const int ntiles = ntc * nmt;
    const int kq = 8 * (lane >> 4);
    const int Hh = A.H >> 1, Wh = A.W >> 1;
    f32x4 acc[kDwMaxTiles];
#pragma unroll
    for (int i = 0; i < kDwMaxTiles; ++i) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int ps = p0; ps < p1; ps += kDwStep) {
        // stage dZ[ps..ps+kDwStep-1][cout] -> dzs[co][px]
        const int gz = A.cout / 8;
        for (int i = threadIdx.x; i < kDwStep * gz; i += 256) {
            int pxl = i / gz, cg = i % gz;
            int p = ps + pxl;
            u32x4 v = u32x4{0u, 0u, 0u, 0u};
            if (p < p1) v = *(const u32x4 *)(A.dz + (size_t)p * A.cout + cg * 8);
            const uint16_t *h = (const uint16_t *)&v;
#pragma unroll
            for (int j = 0; j < 8; ++j) dzs[(cg * 8 + j) * kDwPad + pxl] = h[j];
        }
        // stage X[p + tap][cin] -> xs[ci][px]
        const int gx = cin / 8;
        for (int i = threadIdx.x; i < kDwStep * gx; i += 256) {
            int pxl = i / gx, cg = i % gx;
            int p = ps + pxl;
            u32x4 v = u32x4{0u, 0u, 0u, 0u};
            if (p < p1) {
                int x = p % A.W, q = p / A.W, y = q % A.H, n = q / A.H;
                int ys = y + dy, xs2 = x + dx;
                if (ys >= 0 && ys < A.H && xs2 >= 0 && xs2 < A.W) {
                    int c = cg * 8;
                    if (c < A.ca) {
                        size_t pa = A.aUp ? ((size_t)n * Hh + (ys >> 1)) * Wh + (xs2 >> 1)
                                          : ((size_t)n * A.H + ys) * A.W + xs2;
                        v = *(const u32x4 *)(A.a + pa * A.ca + c);
                    } else {
                        v = *(const u32x4 *)(A.b + (((size_t)n * A.H + ys) * A.W + xs2) * A.cb + (c - A.ca));
                    }
                }
            }
            const uint16_t *h = (const uint16_t *)&v;
#pragma unroll
            for (int j = 0; j < 8; ++j) xs[(cg * 8 + j) * kDwPad + pxl] = h[j];
        }
        __syncthreads();
        // per tile, the 32-pixel K chunks in pixel order (a chunk wholly past p1 is skipped)
#pragma unroll
        for (int i = 0; i < kDwMaxTiles; ++i) {
            const int tile = wave + 4 * i;
            if (tile < ntiles) {
                const int mt = tile / ntc, nt = tile % ntc;
#pragma unroll
                for (int k0 = 0; k0 < kDwStep; k0 += 32) {
                    if (ps + k0 >= p1) break;
                    bf16x8 av = __builtin_bit_cast(
                        bf16x8, *(const u32x4 *)&dzs[(mt * 16 + (lane & 15)) * kDwPad + k0 + kq]);
                    bf16x8 bv = __builtin_bit_cast(
                        bf16x8, *(const u32x4 *)&xs[(nt * 16 + (lane & 15)) * kDwPad + k0 + kq]);
                    acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[i], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    float *out = A.part + (size_t)blockIdx.x * A.cout * 9 * cin;
#pragma unroll
    for (int i = 0; i < kDwMaxTiles; ++i) {
        const int tile = wave + 4 * i;
        if (tile < ntiles) {
            const int mt = tile / ntc, nt = tile % ntc;
            const int ci = nt * 16 + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int co = mt * 16 + 4 * (lane >> 4) + r;
                out[((size_t)co * 9 + tap) * cin + ci] = acc[i][r];
            }
        }
    }
}

// Sums weight-gradient partials over chunks (fixed order) into the flat
// gradient W[co][tap][ciReal].
__global__ void reduce_dw(const float *part, int chunks, int cout, int cinPad, int cinReal, float *g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = cout * 9 * cinReal;
    if (i >= n) return;
    const int ci = i % cinReal, t = i / cinReal, tap = t % 9, co = t / 9;
    const size_t j = ((size_t)co * 9 + tap) * cinPad + ci;
    const size_t stride = (size_t)cout * 9 * cinPad;
    float s = 0.0f;
    for (int c = 0; c < chunks; ++c) s += part[c * stride + j];
    g[i] = s;
}

// Bias gradient: per-chunk column sums of dZ, then a fixed-order reduction.
// Bias gradient, stage 1: column sums of a bf16 [npix][c] tensor over chunks
// of pixPerChunk pixels (c <= 256) into part[chunk][c]. Thread t of the block
// takes channel t % c of pixel phase t / c (256 / c phases, so a wave reads
// whole pixel rows), sums its phase's pixels in order, and the phases are
// added in order through LDS.
__global__ void __launch_bounds__(256) colsum_part(const uint16_t *dz, int npix, int c, int pixPerChunk,
                                                   float *part) {
    __shared__ float red[256];
    const int t = threadIdx.x, phases = 256 / c;
    const int ch = t % c, j = t / c;
    const int p0 = blockIdx.x * pixPerChunk, p1 = min(npix, p0 + pixPerChunk);
    float s = 0.0f;
    if (j < phases)
        for (int p = p0 + j; p < p1; p += phases) s += bf2f(dz[(size_t)p * c + ch]);
    red[t] = s;
    __syncthreads();
    if (t < c) {
        float r = red[t];
        for (int k = 1; k < phases; ++k) r += red[k * c + t];
        part[(size_t)blockIdx.x * c + t] = r;
    }
}
// Stage 2: one block per channel; thread t sums chunks t, t + 256, ... in
// order, then a fixed LDS tree adds the 256 partial sums.
__global__ void __launch_bounds__(256) reduce_cols(const float *part, int chunks, int c, float *g) {
    __shared__ float red[256];
    const int t = threadIdx.x, ch = blockIdx.x;
    float s = 0.0f;
    for (int k = t; k < chunks; k += 256) s += part[(size_t)k * c + ch];
    red[t] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    if (t == 0) g[ch] = red[0];
}

// 2x2 mean pooling of a bf16 NHWC tensor (H, W of the input, even).
__global__ void pool2(const uint16_t *x, int N, int H, int W, int C, uint16_t *y) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int Ho = H >> 1, Wo = W >> 1;
    const size_t n = (size_t)N * Ho * Wo * C;
    if (i >= n) return;
    const int c = (int)(i % C);
    size_t q = i / C;
    const int xo = (int)(q % Wo);
    q /= Wo;
    const int yo = (int)(q % Ho);
    const int nn = (int)(q / Ho);
    const size_t r0 = (((size_t)nn * H + 2 * yo) * W + 2 * xo) * C + c;
    const size_t r1 = r0 + (size_t)W * C;
    float s = (bf2f(x[r0]) + bf2f(x[r0 + C])) + (bf2f(x[r1]) + bf2f(x[r1 + C]));
    y[i] = f2bf(s * 0.25f);
}

// Gradient of the up-sampled half of a decoder input: dZlow[q][c] =
// (sum over the 2x2 children of dXcat[child][c]) * (Ylow[q][c] > 0).
__global__ void up_grad(const float *dxcat, int ccat, int N, int H, int W, int clow, const uint16_t *ylow,
                        uint16_t *dzlow) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int Hh = H >> 1, Wh = W >> 1;
    const size_t n = (size_t)N * Hh * Wh * clow;
    if (i >= n) return;
    const int c = (int)(i % clow);
    size_t q = i / clow;
    const int xo = (int)(q % Wh);
    q /= Wh;
    const int yo = (int)(q % Hh);
    const int nn = (int)(q / Hh);
    const size_t r0 = (((size_t)nn * H + 2 * yo) * W + 2 * xo) * ccat + c;
    const size_t r1 = r0 + (size_t)W * ccat;
    float s = (dxcat[r0] + dxcat[r0 + ccat]) + (dxcat[r1] + dxcat[r1 + ccat]);
    dzlow[i] = f2bf(bf2f(ylow[i]) > 0.0f ? s : 0.0f);
}

// Encoder level k: dZk[p][c] = (dXpool[parent(p)][c] / 4 + dXcat[p][coff + c]) * (Yk[p][c] > 0).
__global__ void enc_grad(const float *dxpool, const float *dxcat, int ccat, int coff, int N, int H, int W, int C,
                         const uint16_t *y, uint16_t *dz) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)N * H * W * C;
    if (i >= n) return;
    const int c = (int)(i % C);
    size_t q = i / C;
    const int x = (int)(q % W);
    q /= W;
    const int yy = (int)(q % H);
    const int nn = (int)(q / H);
    const size_t pp = (((size_t)nn * (H >> 1) + (yy >> 1)) * (W >> 1) + (x >> 1)) * C + c;
    const size_t p = i / C;
    float g = dxpool[pp] * 0.25f + dxcat[p * ccat + coff + c];
    dz[i] = f2bf(bf2f(y[i]) > 0.0f ? g : 0.0f);
}

// Head: o = W d1 + b, y = o + f[0:3], x = m * expm1(min(y, 20)) with
// m = max(albedo feature, 0.01). The loss takes x unclamped (a lower clamp
// would zero the gradient of every pixel it catches); the image is max(x, 0).
// Inference writes x (cropped to the image) into dst; training adds the loss
// and the head's parameter gradient into per-block partials and writes
// dZ(d1) = (W^T dy) * (d1 > 0).
struct HeadArgs {
    const uint16_t *d1;    // [npix][32]
    const uint16_t *feat;  // [npix][32]
    const uint16_t *wb;    // [3][32] bf16
    const float *bias;     // [3]
    int npix, H, W;
    // inference
    float4 *dst;
    int dstW, dstH;
    // training
    const float *target;   // [npix][3]
    float *part;           // [blocks][100]: loss, dW[3][32], db[3]
    uint16_t *dz;          // [npix][32]
    float invCount;
};
// One output channel of the 1x1 head at pixel p: y = bias + W d1 + the
// residual log-radiance feature, x = m * expm1(min(y, 20)) with the albedo
// demodulation factor m (the denoised radiance before the clamp at 0).
__device__ __forceinline__ float head_channel(const uint16_t *wb, const float *bias, const uint16_t *feat, size_t p,
                                              int j, const float *d1v, float &y, float &m) {
    float o = bias[j];
    for (int c = 0; c < 32; ++c) o += bf2f(wb[j * 32 + c]) * d1v[c];
    y = o + bf2f(feat[p * 32 + j]);
    const float yc = fminf(y, 20.0f);
    m = fmaxf(bf2f(feat[p * 32 + 3 + j]), 0.01f);
    return m * expm1f(yc);
}

// Head of a flip ensemble: the batch holds `passes` images of (Hp, Wp), image k
// the frame mirrored by k (bit 0 left-right, bit 1 top-bottom) and moved by
// kEnsembleShift pixels right and down for bit 2; the features are the frame's,
// stored once. Pixel (x, y) of the w x h output is (((v0 + v1) + v2) + ...) *
// (1 / passes) of the clamped head outputs vk at the pixel's positions in the
// images; one pass writes v0 itself.
struct EnsembleArgs {
    const uint16_t *d1, *feat, *wb;
    const float *bias;
    int Wp, Hp, w, h, passes;
    float4 *dst;
};
__global__ void __launch_bounds__(256) head_ensemble_kernel(EnsembleArgs A) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= A.w * A.h) return;
    const int x = i % A.w, y = i / A.w;
    // the features are stored once (image 0); image k's pixel (xk, yk) is the
    // frame's (x, y)
    const size_t pf = (size_t)y * A.Wp + x;
    float acc[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll 1
    for (int k = 0; k < A.passes; ++k) {
        const int sh = (k & 4) ? kEnsembleShift : 0;
        const int xk = ((k & 1) ? A.w - 1 - x : x) + sh, yk = ((k & 2) ? A.h - 1 - y : y) + sh;
        const size_t p = (size_t)k * A.Wp * A.Hp + (size_t)yk * A.Wp + xk;
        float d1v[32];
#pragma unroll
        for (int c = 0; c < 32; ++c) d1v[c] = bf2f(A.d1[p * 32 + c]);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            float yy, m;
            const float v = fmaxf(head_channel(A.wb, A.bias, A.feat, pf, j, d1v, yy, m), 0.0f);
            acc[j] = k == 0 ? v : acc[j] + v;
        }
    }
    const float s = 1.0f / (float)A.passes;
    A.dst[i] = make_float4(acc[0] * s, acc[1] * s, acc[2] * s, 1.0f);
}

__global__ void __launch_bounds__(256) head_kernel(HeadArgs A) {
    __shared__ float red[256];
    const int p = blockIdx.x * 256 + threadIdx.x;
    const bool valid = p < A.npix;
    float d1v[32];
    float dy[3] = {0.f, 0.f, 0.f};
    float lossv = 0.0f;
    if (valid) {
#pragma unroll
        for (int c = 0; c < 32; ++c) d1v[c] = bf2f(A.d1[(size_t)p * 32 + c]);
        float xo[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            float y, m;
            xo[j] = head_channel(A.wb, A.bias, A.feat, p, j, d1v, y, m);
            if (A.target) {
                float r = A.target[(size_t)p * 3 + j];
                float den = r * r + 0.01f;
                float e = xo[j] - r;
                lossv += e * e / den;
                float dx = 2.0f * e / den * A.invCount;
                dy[j] = y < 20.0f ? dx * (xo[j] + m) : 0.0f;
            }
        }
        if (A.dst) {
            int x = p % A.W, q = p / A.W, yy = q % A.H;
            if (x < A.dstW && yy < A.dstH)
                A.dst[(size_t)yy * A.dstW + x] =
                    make_float4(fmaxf(xo[0], 0.0f), fmaxf(xo[1], 0.0f), fmaxf(xo[2], 0.0f), 1.0f);
        }
        if (A.target) {
#pragma unroll
            for (int c = 0; c < 32; ++c) {
                float g = bf2f(A.wb[c]) * dy[0] + bf2f(A.wb[32 + c]) * dy[1] + bf2f(A.wb[64 + c]) * dy[2];
                A.dz[(size_t)p * 32 + c] = f2bf(d1v[c] > 0.0f ? g : 0.0f);
            }
        }
    } else {
#pragma unroll
        for (int c = 0; c < 32; ++c) d1v[c] = 0.0f;
    }
    if (!A.target) return;
    // block reductions: loss, dW[j][c] = sum dy_j d1_c, db[j] = sum dy_j
    float *out = A.part + (size_t)blockIdx.x * 100;
    for (int k = 0; k < 100; ++k) {
        float v;
        if (k == 0) v = lossv;
        else if (k <= 96) v = dy[(k - 1) / 32] * d1v[(k - 1) % 32];
        else v = dy[k - 97];
        red[threadIdx.x] = v;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[k] = red[0];
        __syncthreads();
    }
}
// Sums head partials: loss (scaled by invCount) into *loss, head gradient into g.
// One block per head output k (the loss, 96 weight and 3 bias gradients):
// thread t sums the partials of blocks t, t + 256, ... in order, then a fixed
// LDS tree adds the 256 sums.
__global__ void __launch_bounds__(256) head_reduce(const float *part, int blocks, float invCount, float *loss,
                                                   float *gW, float *gB) {
    __shared__ float red[256];
    const int t = threadIdx.x, k = blockIdx.x;
    float s = 0.0f;
    for (int b = t; b < blocks; b += 256) s += part[(size_t)b * 100 + k];
    red[t] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    if (t == 0) {
        const float v = red[0];
        if (k == 0) *loss = v * invCount;
        else if (k <= 96) gW[k - 1] = v;
        else gB[k - 97] = v;
    }
}

// Features from raw render channels (radiance rgb, albedo rgb, normal xyz,
// depth, mean luma^2, spp): log1p(max(rad, 0) / max(albedo, 0.01)), albedo,
// normal, log1p(max(depth, 0))/4, log1p(standard error of the mean luma),
// log2(spp)/10, zero pad. The albedo is rounded to bfloat16 before it divides,
// so the head's remodulation uses the same value.
__device__ __forceinline__ void write_features(uint16_t *f, const float *raw) {
    uint16_t v[kFeatC];
#pragma unroll
    for (int c = 0; c < kFeatC; ++c) v[c] = 0;
#pragma unroll
    for (int c = 3; c < 9; ++c) v[c] = f2bf(raw[c]);
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] = f2bf(log1pf(fmaxf(raw[c], 0.0f) / fmaxf(bf2f(v[3 + c]), 0.01f)));
    v[9] = f2bf(log1pf(fmaxf(raw[9], 0.0f)) * 0.25f);
    const float n = fmaxf(raw[11], 1.0f);
    const float luma = (raw[0] * 0.2126f + raw[1] * 0.7152f) + raw[2] * 0.0722f;
    const float var = fmaxf(raw[10] - luma * luma, 0.0f) / n;
    v[10] = f2bf(log1pf(sqrtf(var)));
    v[11] = f2bf(log2f(n) * 0.1f);
#pragma unroll
    for (int c = 0; c < kFeatC; c += 8) *(u32x4 *)(f + c) = *(const u32x4 *)&v[c];
}

// Training batch: crop b = hash(seed, b) % nCrops, optional x/y flips.
struct BatchArgs {
    const float *raw;     // [nCrops][h][w][12]
    const float *tgt;     // [nCrops][h][w][3]
    int nCrops, h, w, batch;
    uint32_t seed;
    int augment;
    uint16_t *feat;       // [batch][h][w][32]
    float *target;        // [batch][h][w][3]
};
__device__ __forceinline__ uint32_t hash_u32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}
__global__ void make_batch(BatchArgs A) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)A.batch * A.h * A.w;
    if (i >= n) return;
    const int x = (int)(i % A.w);
    const int y = (int)((i / A.w) % A.h);
    const int b = (int)(i / ((size_t)A.w * A.h));
    const uint32_t hsh = hash_u32(A.seed * 0x9e3779b9u + (uint32_t)b * 0x85ebca6bu + 0x165667b1u);
    const int crop = (int)(hsh % (uint32_t)A.nCrops);
    const int xs = (A.augment && (hsh & 0x10000u)) ? A.w - 1 - x : x;
    const int ys = (A.augment && (hsh & 0x20000u)) ? A.h - 1 - y : y;
    const size_t src = ((size_t)crop * A.h + ys) * A.w + xs;
    write_features(A.feat + i * kFeatC, A.raw + src * kRawC);
#pragma unroll
    for (int c = 0; c < 3; ++c) A.target[i * 3 + c] = A.tgt[src * 3 + c];
}
// Features of explicit inputs (no cropping): raw [n][h][w][12].
__global__ void raw_features(const float *raw, size_t npix, uint16_t *feat) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < npix) write_features(feat + i * kFeatC, raw + i * kRawC);
}
// Features of a rendered frame, padded to (Hp, Wp) by edge replication:
// radiance from the post image (p00 output), albedo/normal/depth from the
// RGBA16F G-buffers, mean luma^2 = variance sum / spp. One image: the mirrored
// images of a flip ensemble read it in place (ConvArgs::mirW).
__global__ void frame_features(const float4 *hdr, const ushort4 *gbuf, const ushort4 *galb, const float *variance,
                               int total, int w, int h, int Wp, int Hp, uint16_t *feat) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Wp * Hp) return;
    const int x = min(i % Wp, w - 1), y = min(i / Wp, h - 1);
    const size_t s = (size_t)y * w + x;
    float4 c = hdr[s];
    ushort4 g = gbuf[s], a = galb[s];
    float raw[kRawC] = {c.x, c.y, c.z, rc_f16_to_f32(a.x), rc_f16_to_f32(a.y), rc_f16_to_f32(a.z),
                        rc_f16_to_f32(g.x), rc_f16_to_f32(g.y), rc_f16_to_f32(g.z), rc_f16_to_f32(g.w),
                        variance[s] / (float)total, (float)total};
    write_features(feat + (size_t)i * kFeatC, raw);
}

// Adam on the FP32 master weights (t = step count after this update).
__global__ void adam_kernel(float *theta, float *m, float *v, const float *g, int n, float lr, float b1, float b2,
                            float eps, float bc1, float bc2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    theta[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// bf16 weight copies: forward layout wb[co][tap][ciPad] (zero padded) and the
// input-gradient layout wt[ci][8 - tap][co].
struct WPrepArgs {
    const float *theta;
    uint16_t *wb, *wt;
    int pOff[kNumLayers], wbOff[kNumLayers], wtOff[kNumLayers], cnt[kNumLayers];
    int cout[kNumLayers], k[kNumLayers], cinPad[kNumLayers], cinReal[kNumLayers];
};
__global__ void prep_weights(WPrepArgs A) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int L = 0;
    while (L < kNumLayers && i >= A.cnt[L]) {
        i -= A.cnt[L];
        ++L;
    }
    if (L >= kNumLayers) return;
    const int kk = A.k[L] * A.k[L], cinPad = A.cinPad[L], cinReal = A.cinReal[L];
    const int ci = i % cinPad, t = i / cinPad, tap = t % kk, co = t / kk;
    float v = ci < cinReal ? A.theta[A.pOff[L] + ((size_t)co * kk + tap) * cinReal + ci] : 0.0f;
    uint16_t h = f2bf(v);
    A.wb[A.wbOff[L] + i] = h;
    if (A.k[L] == 3) A.wt[A.wtOff[L] + ((size_t)ci * 9 + (8 - tap)) * A.cout[L] + co] = h;
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side.
struct Net {
    int nParams = 0;
    int pOff[kNumLayers], bOff[kNumLayers], wbOff[kNumLayers], wtOff[kNumLayers], wbCnt[kNumLayers];
    int wbTotal = 0, wtTotal = 0;
    float *theta = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
    uint16_t *wb = nullptr, *wt = nullptr;
    uint16_t *zero = nullptr;  // 256 zero bf16: the source of the tiled convolution's padding loads
    int step = 0;
    bool ready = false;
    // dataset
    float *raw = nullptr, *tgt = nullptr;  // [nCrops][h][w][12], [nCrops][h][w][3]
    int nCrops = 0, cropH = 0, cropW = 0;
    // workspace: capacity in level-0 pixels (N*H*W); level k holds cap >> 2k
    size_t wsCap = 0;
    bool wsTrain = false;
    uint16_t *feat = nullptr;
    uint16_t *act[kNumLayers] = {};   // e1..e5, d4..d1 (index 9 unused)
    uint16_t *pooled[4] = {};         // pool(e1..e4)
    uint16_t *dz[kNumLayers] = {};    // gradients of pre-activations
    float *dxcat[4] = {};             // decoder input gradients (d1, d2, d3, d4)
    float *dxpool[4] = {};            // encoder input gradients (e2..e5)
    float *target = nullptr;
    float *part = nullptr;            // weight-gradient partials
    size_t partBytes = 0;
    float *hpart = nullptr;           // head partials
    float *lossDev = nullptr;         // per-step losses
    int lossCap = 0;
};

namespace {

void net_layout(Net &n) {
    int off = 0, wb = 0, wt = 0;
    for (int L = 0; L < kNumLayers; ++L) {
        const LayerSpec &s = kLayers[L];
        n.pOff[L] = off;
        off += s.cout * s.k * s.k * s.cinReal;
        n.bOff[L] = off;
        off += s.cout;
        n.wbOff[L] = wb;
        n.wbCnt[L] = s.cout * s.k * s.k * s.cinPad;
        wb += n.wbCnt[L];
        n.wtOff[L] = wt;
        if (s.k == 3) wt += s.cout * 9 * s.cinPad;
    }
    n.nParams = off;
    n.wbTotal = wb;
    n.wtTotal = wt;
}

template <typename T>
hipError_t nalloc(T *&p, size_t count) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T) > 0 ? count * sizeof(T) : 16);
    p = (T *)q;
    return e;
}
template <typename T>
void nfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

void free_ws(Net &n) {
    nfree(n.feat);
    for (auto &p : n.act) nfree(p);
    for (auto &p : n.pooled) nfree(p);
    for (auto &p : n.dz) nfree(p);
    for (auto &p : n.dxcat) nfree(p);
    for (auto &p : n.dxpool) nfree(p);
    nfree(n.target);
    nfree(n.part);
    nfree(n.hpart);
    n.wsCap = 0;
    n.wsTrain = false;
    n.partBytes = 0;
}

size_t level_pix(int N, int H, int W, int level) { return (size_t)N * (H >> level) * (W >> level); }
// Capacity of a level-k buffer for cap level-0 pixels (H, W multiples of 16).
size_t level_cap(size_t cap, int level) { return cap >> (2 * level); }

constexpr int kDwPixPerChunk = 1024;
static_assert(kDwPixPerChunk % kDwStep == 0, "chunks hold whole staging rounds");

// Grows the workspace to hold N x H x W (H, W multiples of 16); buffers are
// kept for smaller shapes, so alternating view sizes do not reallocate.
hipError_t ensure_ws(rc_ctx *ctx, Net &n, int N, int H, int W, bool train) {
    const size_t need = (size_t)N * H * W;
    if (need <= n.wsCap && (n.wsTrain || !train)) return hipSuccess;
    if (ctx->capturing) return hipErrorStreamCaptureUnsupported;  // would allocate inside a captured frame
    const size_t cap = need > n.wsCap ? need : n.wsCap;
    const bool tr = train || n.wsTrain;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return e;
    ++ctx->freeGen;
    free_ws(n);
    auto A = [&](auto *&p, size_t cnt) {
        if (e == hipSuccess) e = nalloc(p, cnt);
    };
    A(n.feat, cap * kFeatC);
    for (int L = 0; L < 9; ++L) A(n.act[L], level_cap(cap, kLayers[L].level) * kLayers[L].cout);
    for (int k = 0; k < 4; ++k) A(n.pooled[k], level_cap(cap, k + 1) * kLayers[k].cout);
    if (tr) {
        for (int L = 0; L < 9; ++L) A(n.dz[L], level_cap(cap, kLayers[L].level) * kLayers[L].cout);
        for (int j = 0; j < 4; ++j) {  // dxcat[j] for decoder d(j+1): layer 8 - j
            const LayerSpec &s = kLayers[8 - j];
            A(n.dxcat[j], level_cap(cap, s.level) * s.cinPad);
        }
        for (int k = 0; k < 4; ++k) {  // dxpool[k] for encoder e(k+2): level k+1, cin = cout of e(k+1)
            A(n.dxpool[k], level_cap(cap, k + 1) * kLayers[k].cout);
        }
        A(n.target, cap * 3);
        size_t maxPart = 0;
        for (int L = 0; L < 9; ++L) {
            const LayerSpec &s = kLayers[L];
            size_t np = level_cap(cap, s.level);
            size_t chunks = (np + kDwPixPerChunk - 1) / kDwPixPerChunk;
            size_t sz = chunks * (size_t)s.cout * 9 * s.cinPad;
            size_t csz = ((np + 255) / 256) * (size_t)s.cout;
            if (sz > maxPart) maxPart = sz;
            if (csz > maxPart) maxPart = csz;
        }
        A(n.part, maxPart);
        n.partBytes = maxPart * 4;
        A(n.hpart, ((cap + 255) / 256) * 100);
    }
    if (e != hipSuccess) {
        free_ws(n);
        return e;
    }
    n.wsCap = cap;
    n.wsTrain = tr;
    return hipSuccess;
}

unsigned blocks_for(size_t n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }

template <int NT, int G>
void launch_conv_nt(hipStream_t s, const ConvArgs &A) {
    const size_t npix = (size_t)A.N * A.H * A.W;
    const size_t perBlock = (size_t)32 * (4 / G);  // 4/G pixel tiles of 32
    const unsigned grid = (unsigned)((npix + perBlock - 1) / perBlock);
    hipLaunchKernelGGL((conv3x3_fwd<NT, 2, G>), dim3(grid), dim3(256), 0, s, A);
}

// Below this many 32-pixel tiles (4 waves per SIMD on 256 CUs) a layer splits
// its output channels over the waves of a block.
constexpr size_t kFewTiles = 4096;

// conv3x3_tile for the (input, output) channel counts of the three finest
// levels: e1 32/32, e2 32/64, e3 64/64, d3 128/64, d2 128/32, d1 64/32; false
// for any other shape.
template <int CIN, int COUT, int TH, int G, int WV = 4>
void launch_tile(hipStream_t s, ConvArgs A, bool &pooled) {
    using TS = TileShape<CIN, COUT, TH, 64 * WV>;
    if ((TH * G / WV) % 2) A.pool = nullptr;  // the pooled output needs row pairs per wave
    pooled = A.pool != nullptr;
    const int tiles = A.N * ((A.H + TH - 1) / TH) * ((A.W + kTileX - 1) / kTileX);
    const int per = (tiles + 7) / 8;
    // blocks per XCD: as many as fit on its 32 CUs at once (LDS-bound), at most
    // one per tile of its run
    const int perCu = std::max(1, (int)(160 * 1024 / TS::LDS_BYTES));
    const int bpx = std::min(per, 32 * perCu);
    hipLaunchKernelGGL((conv3x3_tile<CIN, COUT, TH, G, WV>), dim3(8 * bpx), dim3(64 * WV), 0, s, A);
}

// pooled: whether the launch also wrote A.pool (the 2 x 2 mean of the output;
// H and W even).
bool launch_conv_tiled(hipStream_t s, const ConvArgs &A, bool &pooled) {
    const int cin = A.ca + A.cb;
    pooled = false;
    if ((A.ca & 7) || (A.cb & 7) || !A.outB || !A.zero) return false;
    if (A.pool && ((A.H | A.W) & 1)) return false;
    if (cin == 32 && A.cout == 32) launch_tile<32, 32, 8, 2>(s, A, pooled);
    else if (cin == 32 && A.cout == 64) launch_tile<32, 64, 8, 4>(s, A, pooled);
    else if (cin == 64 && A.cout == 64) launch_tile<64, 64, 8, 4>(s, A, pooled);
    else if (cin == 64 && A.cout == 32) launch_tile<64, 32, 8, 2>(s, A, pooled);
    else if (cin == 128 && A.cout == 32) launch_tile<128, 32, 4, 2>(s, A, pooled);
    else if (cin == 128 && A.cout == 64) launch_tile<128, 64, 4, 4>(s, A, pooled);
    else if (cin == 64 && A.cout == 96) launch_tile<64, 96, 4, 6, 6>(s, A, pooled);    // e4: 6 channel groups
    else if (cin == 96 && A.cout == 128) launch_tile<96, 128, 2, 8, 8>(s, A, pooled);  // e5: 8 channel groups
    else return false;
    return true;
}

hipError_t launch_conv(hipStream_t s, const ConvArgs &A, bool allowGroups = true) {
    if (A.cout % 16) return hipErrorInvalidValue;
    const bool few = allowGroups && ((size_t)A.N * A.H * A.W + 31) / 32 < kFewTiles;
    switch (A.cout / 16) {
        case 2: few ? launch_conv_nt<1, 2>(s, A) : launch_conv_nt<2, 1>(s, A); break;
        case 4: few ? launch_conv_nt<1, 4>(s, A) : launch_conv_nt<4, 1>(s, A); break;
        case 6: few ? launch_conv_nt<3, 2>(s, A) : launch_conv_nt<6, 1>(s, A); break;
        case 8: few ? launch_conv_nt<2, 4>(s, A) : launch_conv_nt<8, 1>(s, A); break;
        case 14: few ? launch_conv_nt<7, 2>(s, A) : launch_conv_nt<14, 1>(s, A); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Input sources of conv layer L (a, b) at the workspace's shape.
void layer_inputs(Net &n, int L, const uint16_t *&a, const uint16_t *&b) {
    const LayerSpec &s = kLayers[L];
    if (s.srcA < 0) a = n.feat;
    else if (s.aPool) a = n.pooled[s.srcA];
    else a = n.act[s.srcA];
    b = s.srcB >= 0 ? n.act[s.srcB] : nullptr;
}

// mirW, mirH > 0: the features hold one padded mirW x mirH frame and image n of
// the batch is that frame mirrored by n (a flip ensemble).
hipError_t forward(rc_ctx *ctx, Net &n, int N, int H, int W, int mirW = 0, int mirH = 0) {
    hipStream_t s = ctx->stream;
    bool pooledBy[kNumLayers] = {};  // the layer's convolution wrote its pooled output
    for (int L = 0; L < 9; ++L) {
        const LayerSpec &sp = kLayers[L];
        if (L >= 1 && L <= 4 && !pooledBy[L - 1]) {  // pool the previous encoder output
            const int lv = kLayers[L - 1].level;
            const size_t cnt = level_pix(N, H, W, lv + 1) * kLayers[L - 1].cout;
            hipLaunchKernelGGL(pool2, dim3(blocks_for(cnt)), dim3(256), 0, s, n.act[L - 1], N, H >> lv, W >> lv,
                               kLayers[L - 1].cout, n.pooled[L - 1]);
        }
        ConvArgs A{};
        layer_inputs(n, L, A.a, A.b);
        A.ca = sp.ca;
        A.cb = sp.cb;
        A.aUp = sp.aUp;
        A.N = N;
        A.H = H >> sp.level;
        A.W = W >> sp.level;
        A.w = n.wb + n.wbOff[L];
        A.bias = n.theta + n.bOff[L];
        A.cout = sp.cout;
        A.relu = 1;
        A.outB = n.act[L];
        A.zero = n.zero;
        if (sp.srcA < 0) {  // the features
            A.mirW = mirW;
            A.mirH = mirH;
        }
        // the encoders write their pooled output (the next level's input) when
        // their kernel can
        A.pool = L <= 3 ? n.pooled[L] : nullptr;
        // every layer of 128 input channels or fewer stages its input tiles in LDS
        if (ctx->netKernels == RC_NET_KERNELS_TILED && launch_conv_tiled(s, A, pooledBy[L])) continue;
        A.pool = nullptr;
        hipError_t e = launch_conv(s, A, ctx->netKernels != RC_NET_KERNELS_SINGLE_GROUP);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

hipError_t weight_grad(rc_ctx *ctx, Net &n, int L, int N, int H, int W) {
    hipStream_t s = ctx->stream;
    const LayerSpec &sp = kLayers[L];
    DwArgs D{};
    D.dz = n.dz[L];
    layer_inputs(n, L, D.a, D.b);
    D.ca = sp.ca;
    D.cb = sp.cb;
    D.aUp = sp.aUp;
    D.N = N;
    D.H = H >> sp.level;
    D.W = W >> sp.level;
    D.cout = sp.cout;
    D.pixPerChunk = kDwPixPerChunk;
    D.part = n.part;
    const size_t np = level_pix(N, H, W, sp.level);
    const unsigned chunks = blocks_for(np, kDwPixPerChunk);
    hipLaunchKernelGGL(conv3x3_dw, dim3(chunks, 9), dim3(256), 0, s, D);
    const int cnt = sp.cout * 9 * sp.cinReal;
    hipLaunchKernelGGL(reduce_dw, dim3(blocks_for(cnt)), dim3(256), 0, s, n.part, (int)chunks, sp.cout, sp.cinPad,
                       sp.cinReal, n.grad + n.pOff[L]);
    const unsigned cchunks = blocks_for(np, 256);
    if (sp.cout > 256) return hipErrorInvalidValue;  // colsum_part holds a channel per thread
    hipLaunchKernelGGL(colsum_part, dim3(cchunks), dim3(256), 0, s, n.dz[L], (int)np, sp.cout, 256, n.part);
    hipLaunchKernelGGL(reduce_cols, dim3(sp.cout), dim3(256), 0, s, n.part, (int)cchunks, sp.cout,
                       n.grad + n.bOff[L]);
    return hipGetLastError();
}

hipError_t input_grad(rc_ctx *ctx, Net &n, int L, int N, int H, int W, float *out) {
    const LayerSpec &sp = kLayers[L];
    ConvArgs A{};
    A.a = n.dz[L];
    A.ca = sp.cout;
    A.b = nullptr;
    A.cb = 0;
    A.aUp = 0;
    A.N = N;
    A.H = H >> sp.level;
    A.W = W >> sp.level;
    A.w = n.wt + n.wtOff[L];
    A.bias = nullptr;
    A.cout = sp.cinPad;
    A.relu = 0;
    A.outF = out;
    return launch_conv(ctx->stream, A);
}

// Forward + loss + backward on the workspace batch (features and target
// already in place); leaves the gradient in n.grad and the loss in *lossDev.
hipError_t forward_backward(rc_ctx *ctx, Net &n, int N, int H, int W, float *lossDev) {
    hipStream_t s = ctx->stream;
    hipError_t e = forward(ctx, n, N, H, W);
    if (e != hipSuccess) return e;
    const size_t np = level_pix(N, H, W, 0);
    HeadArgs h{};
    h.d1 = n.act[8];
    h.feat = n.feat;
    h.wb = n.wb + n.wbOff[9];
    h.bias = n.theta + n.bOff[9];
    h.npix = (int)np;
    h.H = H;
    h.W = W;
    h.target = n.target;
    h.part = n.hpart;
    h.dz = n.dz[8];
    h.invCount = 1.0f / (float)(np * 3);
    const unsigned hb = blocks_for(np);
    hipLaunchKernelGGL(head_kernel, dim3(hb), dim3(256), 0, s, h);
    hipLaunchKernelGGL(head_reduce, dim3(100), dim3(256), 0, s, n.hpart, (int)hb, h.invCount, lossDev,
                       n.grad + n.pOff[9], n.grad + n.bOff[9]);
    // decoders d1 (8), d2 (7), d3 (6), d4 (5): weight gradient, input gradient,
    // then the up-sampled source's dZ.
    for (int L = 8; L >= 5; --L) {
        const LayerSpec &sp = kLayers[L];
        if ((e = weight_grad(ctx, n, L, N, H, W)) != hipSuccess) return e;
        float *dx = n.dxcat[8 - L];
        if ((e = input_grad(ctx, n, L, N, H, W, dx)) != hipSuccess) return e;
        const int lowL = sp.srcA;  // d2, d3, d4 or e5
        const size_t cnt = level_pix(N, H, W, sp.level + 1) * sp.ca;
        hipLaunchKernelGGL(up_grad, dim3(blocks_for(cnt)), dim3(256), 0, s, dx, sp.cinPad, N, H >> sp.level,
                           W >> sp.level, sp.ca, n.act[lowL], n.dz[lowL]);
    }
    // encoders e5 (4) .. e2 (1): weight gradient, gradient of the pooled input,
    // then the previous encoder's dZ (pooled path + skip path).
    for (int L = 4; L >= 1; --L) {
        if ((e = weight_grad(ctx, n, L, N, H, W)) != hipSuccess) return e;
        float *dxp = n.dxpool[L - 1];
        if ((e = input_grad(ctx, n, L, N, H, W, dxp)) != hipSuccess) return e;
        const int k = L - 1;                 // encoder whose output was pooled
        const LayerSpec &dec = kLayers[8 - k];  // decoder that takes e(k+1) as its skip
        const int lv = kLayers[k].level;
        const size_t cnt = level_pix(N, H, W, lv) * kLayers[k].cout;
        hipLaunchKernelGGL(enc_grad, dim3(blocks_for(cnt)), dim3(256), 0, s, dxp, n.dxcat[k], dec.cinPad, dec.ca, N,
                           H >> lv, W >> lv, kLayers[k].cout, n.act[k], n.dz[k]);
    }
    if ((e = weight_grad(ctx, n, 0, N, H, W)) != hipSuccess) return e;
    return hipGetLastError();
}

hipError_t prep(rc_ctx *ctx, Net &n) {
    WPrepArgs A{};
    A.theta = n.theta;
    A.wb = n.wb;
    A.wt = n.wt;
    for (int L = 0; L < kNumLayers; ++L) {
        A.pOff[L] = n.pOff[L];
        A.wbOff[L] = n.wbOff[L];
        A.wtOff[L] = n.wtOff[L];
        A.cnt[L] = n.wbCnt[L];
        A.cout[L] = kLayers[L].cout;
        A.k[L] = kLayers[L].k;
        A.cinPad[L] = kLayers[L].cinPad;
        A.cinReal[L] = kLayers[L].cinReal;
    }
    hipLaunchKernelGGL(prep_weights, dim3(blocks_for(n.wbTotal)), dim3(256), 0, ctx->stream, A);
    return hipGetLastError();
}

int check_shape(rc_ctx *ctx, int N, int H, int W) {
    if (N < 1 || H < 16 || W < 16 || (H % 16) || (W % 16))
        return set_error(ctx, RC_ERR_INVALID_ARG, "net input: n >= 1 and h, w multiples of 16 (>= 16) required");
    if ((size_t)N * H * W > (size_t)64 << 20) return set_error(ctx, RC_ERR_LIMIT, "net input larger than 64 Mpixels");
    return RC_OK;
}

Net *get_net(rc_ctx *ctx) {
    if (!ctx->net) {
        ctx->net = new (std::nothrow) Net();
        if (ctx->net) net_layout(*ctx->net);
    }
    return ctx->net;
}

// Seeded He-normal initialisation on the host (PCG32 + Box-Muller).
void he_init(std::vector<float> &theta, const Net &n, uint32_t seed) {
    uint64_t state = 0x853c49e6748fea9bull ^ ((uint64_t)seed << 1), inc = ((uint64_t)seed << 1) | 1u;
    auto next = [&]() -> uint32_t {
        uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((32 - rot) & 31));
    };
    auto unif = [&]() { return ((double)next() + 0.5) / 4294967296.0; };
    theta.assign(n.nParams, 0.0f);
    for (int L = 0; L < kNumLayers; ++L) {
        const LayerSpec &s = kLayers[L];
        const int cnt = s.cout * s.k * s.k * s.cinReal;
        const double sd = std::sqrt(2.0 / (double)(s.k * s.k * s.cinReal));
        for (int i = 0; i < cnt; ++i) {
            double z = std::sqrt(-2.0 * std::log(unif())) * std::cos(6.283185307179586 * unif());
            theta[n.pOff[L] + i] = (float)(z * sd);
        }
    }
}

hipError_t alloc_params(Net &n) {
    hipError_t e = hipSuccess;
    if (!n.theta) e = nalloc(n.theta, n.nParams);
    if (e == hipSuccess && !n.m) e = nalloc(n.m, n.nParams);
    if (e == hipSuccess && !n.v) e = nalloc(n.v, n.nParams);
    if (e == hipSuccess && !n.grad) e = nalloc(n.grad, n.nParams);
    if (e == hipSuccess && !n.wb) e = nalloc(n.wb, n.wbTotal);
    if (e == hipSuccess && !n.wt) e = nalloc(n.wt, n.wtTotal);
    if (e == hipSuccess && !n.zero) e = nalloc(n.zero, 256);
    return e;
}

int set_params_host(rc_ctx *ctx, Net &n, const float *theta) {
    hipError_t e = alloc_params(n);
    if (e == hipSuccess) e = hipMemcpyAsync(n.theta, theta, (size_t)n.nParams * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(n.m, 0, (size_t)n.nParams * 4, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(n.v, 0, (size_t)n.nParams * 4, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(n.wt, 0, (size_t)n.wtTotal * 2, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(n.zero, 0, 256 * 2, ctx->stream);
    if (e == hipSuccess) e = prep(ctx, n);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "net parameter upload");
    n.step = 0;
    n.ready = true;
    return RC_OK;
}

}  // namespace

void free_net(Net *n) {
    if (!n) return;
    free_ws(*n);
    nfree(n->theta); nfree(n->m); nfree(n->v); nfree(n->grad); nfree(n->wb); nfree(n->wt); nfree(n->zero);
    nfree(n->raw); nfree(n->tgt); nfree(n->lossDev);
    delete n;
}

bool net_ready(const rc_ctx *ctx) { return ctx->net && ctx->net->ready; }

// Denoises the post image src (w x h RGBA32F, p00 output) into dst using the
// frame's G-buffers: the mean of ctx->netFlips network passes over the frame
// as is, mirrored left-right, top-bottom and both (in that order), then (8
// passes) the same four moved kEnsembleShift pixels right and down, each pass's
// output mapped back; (((p0 + p1) + p2) + ...) * (1 / passes) per channel. The
// passes run as one batch of images through the network; with 8 passes the
// padded image has room for the move.
int run_net_denoise(rc_ctx *ctx, const float4 *src, const ushort4 *gbuf, const ushort4 *galbedo,
                    const float *variance, int total, float4 *dst, int w, int h) {
    Net &n = *ctx->net;
    const int room = ctx->netFlips > 4 ? kEnsembleShift : 0;
    const int Wp = (w + room + 15) / 16 * 16, Hp = (h + room + 15) / 16 * 16;
    const int passes = ctx->netFlips;
    if ((size_t)passes * Wp * Hp > (size_t)64 << 20)
        return set_error(ctx, RC_ERR_LIMIT, "net denoise: passes x padded frame above 64 Mpixels (lower rc_set_net_flips)");
    hipError_t e = ensure_ws(ctx, n, passes, Hp, Wp, false);
    if (e != hipSuccess) return hip_fail(ctx, e, "net workspace");
    hipLaunchKernelGGL(frame_features, dim3(blocks_for((size_t)Wp * Hp)), dim3(256), 0, ctx->stream, src, gbuf,
                       galbedo, variance, total, w, h, Wp, Hp, n.feat);
    if ((e = forward(ctx, n, passes, Hp, Wp, w, h)) != hipSuccess) return hip_fail(ctx, e, "net forward");
    EnsembleArgs hd{};
    hd.d1 = n.act[8];
    hd.feat = n.feat;
    hd.wb = n.wb + n.wbOff[9];
    hd.bias = n.theta + n.bOff[9];
    hd.Wp = Wp;
    hd.Hp = Hp;
    hd.w = w;
    hd.h = h;
    hd.passes = passes;
    hd.dst = dst;
    hipLaunchKernelGGL(head_ensemble_kernel, dim3(blocks_for((size_t)w * h)), dim3(256), 0, ctx->stream, hd);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(ctx, e, "net head");
    return RC_OK;
}

}  // namespace rc

using namespace rc;

extern "C" {

int32_t rc_net_param_count(void) {
    Net n;
    net_layout(n);
    return n.nParams;
}

int rc_net_init(rc_ctx *ctx, uint32_t seed) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (int rc_ = capture_block(ctx, "rc_net_init")) return rc_;
    (void)hipSetDevice(ctx->device);
    Net *n = get_net(ctx);
    if (!n) return set_error(ctx, RC_ERR_OOM, "host allocation failed");
    std::vector<float> theta;
    he_init(theta, *n, seed);
    return set_params_host(ctx, *n, theta.data());
}

int rc_net_set_params(rc_ctx *ctx, const float *theta, int32_t count) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (int rc_ = capture_block(ctx, "rc_net_set_params")) return rc_;
    (void)hipSetDevice(ctx->device);
    Net *n = get_net(ctx);
    if (!n) return set_error(ctx, RC_ERR_OOM, "host allocation failed");
    if (!theta || count != n->nParams) return set_error(ctx, RC_ERR_INVALID_ARG, "parameter count mismatch");
    return set_params_host(ctx, *n, theta);
}

int rc_net_get_params(rc_ctx *ctx, float *theta, int32_t count) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (int rc_ = capture_block(ctx, "rc_net_get_params")) return rc_;
    if (!net_ready(ctx)) return set_error(ctx, RC_ERR_STATE, "net not initialised (rc_net_init / rc_net_set_params)");
    Net &n = *ctx->net;
    if (!theta || count != n.nParams) return set_error(ctx, RC_ERR_INVALID_ARG, "parameter count mismatch");
    hipError_t e = hipMemcpyAsync(theta, n.theta, (size_t)n.nParams * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "net parameter read");
    return RC_OK;
}

int rc_net_set_dataset(rc_ctx *ctx, const float *raw, const float *target, int32_t n, int32_t h, int32_t w) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (int rc_ = capture_block(ctx, "rc_net_set_dataset")) return rc_;
    (void)hipSetDevice(ctx->device);
    Net *net = get_net(ctx);
    if (!net) return set_error(ctx, RC_ERR_OOM, "host allocation failed");
    // The dataset holds 60 B per pixel (raw + target) and is indexed with size_t by
    // make_batch; it is not bound by the forward workspace's 64 Mpixel limit.
    if (n < 1) return set_error(ctx, RC_ERR_INVALID_ARG, "training set: n >= 1 required");
    if ((size_t)n * h * w > ((size_t)512 << 20))
        return set_error(ctx, RC_ERR_LIMIT, "training set larger than 512 Mpixels");
    int rc = check_shape(ctx, 1, h, w);
    if (rc) return rc;
    if (!raw || !target) return set_error(ctx, RC_ERR_INVALID_ARG, "NULL dataset");
    hipError_t e = hipStreamSynchronize(ctx->stream);
    nfree(net->raw);
    nfree(net->tgt);
    const size_t px = (size_t)n * h * w;
    if (e == hipSuccess) e = nalloc(net->raw, px * kRawC);
    if (e == hipSuccess) e = nalloc(net->tgt, px * 3);
    if (e == hipSuccess) e = hipMemcpyAsync(net->raw, raw, px * kRawC * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(net->tgt, target, px * 12, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        nfree(net->raw);
        nfree(net->tgt);
        net->nCrops = 0;
        return hip_fail(ctx, e, "net dataset upload");
    }
    net->nCrops = n;
    net->cropH = h;
    net->cropW = w;
    return RC_OK;
}

int rc_net_train(rc_ctx *ctx, int32_t steps, int32_t batch, float lr, uint32_t seed, float *losses) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (int rc_ = capture_block(ctx, "rc_net_train")) return rc_;
    if (!net_ready(ctx)) return set_error(ctx, RC_ERR_STATE, "net not initialised (rc_net_init / rc_net_set_params)");
    Net &n = *ctx->net;
    if (n.nCrops <= 0) return set_error(ctx, RC_ERR_STATE, "no dataset (rc_net_set_dataset)");
    if (steps < 1 || batch < 1 || !(lr > 0.0f)) return set_error(ctx, RC_ERR_INVALID_ARG, "steps, batch >= 1, lr > 0");
    int rc = check_shape(ctx, batch, n.cropH, n.cropW);
    if (rc) return rc;
    (void)hipSetDevice(ctx->device);
    hipError_t e = ensure_ws(ctx, n, batch, n.cropH, n.cropW, true);
    if (e == hipSuccess && n.lossCap < steps) {
        nfree(n.lossDev);
        e = nalloc(n.lossDev, (size_t)steps);
        n.lossCap = e == hipSuccess ? steps : 0;
    }
    if (e != hipSuccess) return hip_fail(ctx, e, "net training buffers");
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const size_t npx = (size_t)batch * n.cropH * n.cropW;
    for (int it = 0; it < steps; ++it) {
        BatchArgs B{};
        B.raw = n.raw;
        B.tgt = n.tgt;
        B.nCrops = n.nCrops;
        B.h = n.cropH;
        B.w = n.cropW;
        B.batch = batch;
        B.seed = seed * 2654435761u + (uint32_t)(n.step + it);
        B.augment = 1;
        B.feat = n.feat;
        B.target = n.target;
        hipLaunchKernelGGL(make_batch, dim3(blocks_for(npx)), dim3(256), 0, ctx->stream, B);
        if ((e = forward_backward(ctx, n, batch, n.cropH, n.cropW, n.lossDev + it)) != hipSuccess)
            return hip_fail(ctx, e, "net forward/backward");
        const int t = n.step + it + 1;
        const float bc1 = 1.0f - (float)std::pow(0.9, t), bc2 = 1.0f - (float)std::pow(0.999, t);
        hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(n.nParams)), dim3(256), 0, ctx->stream, n.theta, n.m, n.v,
                           n.grad, n.nParams, lr, b1, b2, eps, bc1, bc2);
        if ((e = prep(ctx, n)) != hipSuccess) return hip_fail(ctx, e, "net weight update");
    }
    n.step += steps;
    if (losses) {
        e = hipMemcpyAsync(losses, n.lossDev, (size_t)steps * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return hip_fail(ctx, e, "net loss read");
    }
    return RC_OK;
}

int rc_net_grad(rc_ctx *ctx, const float *raw, const float *target, int32_t n, int32_t h, int32_t w, float *loss,
                float *grad) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (int rc_ = capture_block(ctx, "rc_net_grad")) return rc_;
    if (!net_ready(ctx)) return set_error(ctx, RC_ERR_STATE, "net not initialised (rc_net_init / rc_net_set_params)");
    int rc = check_shape(ctx, n, h, w);
    if (rc) return rc;
    if (!raw || !target || !loss || !grad) return set_error(ctx, RC_ERR_INVALID_ARG, "NULL argument");
    (void)hipSetDevice(ctx->device);
    Net &net = *ctx->net;
    hipError_t e = ensure_ws(ctx, net, n, h, w, true);
    if (e == hipSuccess && net.lossCap < 1) {
        e = nalloc(net.lossDev, 1);
        net.lossCap = e == hipSuccess ? 1 : 0;
    }
    const size_t px = (size_t)n * h * w;
    float *rawDev = nullptr;
    if (e == hipSuccess) e = nalloc(rawDev, px * kRawC);
    if (e == hipSuccess) e = hipMemcpyAsync(rawDev, raw, px * kRawC * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(net.target, target, px * 12, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(raw_features, dim3(blocks_for(px)), dim3(256), 0, ctx->stream, rawDev, px, net.feat);
        e = forward_backward(ctx, net, n, h, w, net.lossDev);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(grad, net.grad, (size_t)net.nParams * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(loss, net.lossDev, 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (rawDev) (void)hipFree(rawDev);
    if (e != hipSuccess) return hip_fail(ctx, e, "net gradient");
    return RC_OK;
}

int rc_net_forward(rc_ctx *ctx, const float *raw, int32_t n, int32_t h, int32_t w, float *out) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (int rc_ = capture_block(ctx, "rc_net_forward")) return rc_;
    if (!net_ready(ctx)) return set_error(ctx, RC_ERR_STATE, "net not initialised (rc_net_init / rc_net_set_params)");
    int rc = check_shape(ctx, n, h, w);
    if (rc) return rc;
    if (!raw || !out) return set_error(ctx, RC_ERR_INVALID_ARG, "NULL argument");
    (void)hipSetDevice(ctx->device);
    Net &net = *ctx->net;
    hipError_t e = ensure_ws(ctx, net, n, h, w, false);
    const size_t px = (size_t)n * h * w;
    float *rawDev = nullptr;
    float4 *outDev = nullptr;
    if (e == hipSuccess) e = nalloc(rawDev, px * kRawC);
    if (e == hipSuccess) e = nalloc(outDev, px);
    if (e == hipSuccess) e = hipMemcpyAsync(rawDev, raw, px * kRawC * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(raw_features, dim3(blocks_for(px)), dim3(256), 0, ctx->stream, rawDev, px, net.feat);
        e = forward(ctx, net, n, h, w);
    }
    if (e == hipSuccess) {
        HeadArgs hd{};
        hd.d1 = net.act[8];
        hd.feat = net.feat;
        hd.wb = net.wb + net.wbOff[9];
        hd.bias = net.theta + net.bOff[9];
        hd.npix = (int)px;
        hd.H = (int)(px / w);  // rows of all images stacked
        hd.W = w;
        hd.dst = outDev;
        hd.dstW = w;
        hd.dstH = (int)(px / w);
        hipLaunchKernelGGL(head_kernel, dim3(blocks_for(px)), dim3(256), 0, ctx->stream, hd);
        e = hipGetLastError();
    }
    std::vector<float4> tmp;
    if (e == hipSuccess) {
        tmp.resize(px);
        e = hipMemcpyAsync(tmp.data(), outDev, px * 16, hipMemcpyDeviceToHost, ctx->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (rawDev) (void)hipFree(rawDev);
    if (outDev) (void)hipFree(outDev);
    if (e != hipSuccess) return hip_fail(ctx, e, "net forward");
    for (size_t i = 0; i < px; ++i) {
        out[3 * i + 0] = tmp[i].x;
        out[3 * i + 1] = tmp[i].y;
        out[3 * i + 2] = tmp[i].z;
    }
    return RC_OK;
}

int rc_set_net_kernels(rc_ctx *ctx, int32_t mode) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (mode != RC_NET_KERNELS_TILED && mode != RC_NET_KERNELS_DIRECT && mode != RC_NET_KERNELS_SINGLE_GROUP)
        return set_error(ctx, RC_ERR_INVALID_ARG, "bad network kernel mode");
    if (int rc_ = capture_block(ctx, "rc_set_net_kernels")) return rc_;
    ctx->netKernels = mode;
    return RC_OK;
}

int rc_set_net_flips(rc_ctx *ctx, int32_t passes) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (passes != 1 && passes != 2 && passes != 4 && passes != 8)
        return set_error(ctx, RC_ERR_INVALID_ARG, "passes must be 1, 2, 4 or 8");
    ctx->netFlips = passes;
    return RC_OK;
}

int rc_set_denoiser(rc_ctx *ctx, int32_t mode) {
    if (!ctx) return RC_ERR_INVALID_ARG;
    if (mode != RC_DENOISE_ATROUS && mode != RC_DENOISE_NET)
        return set_error(ctx, RC_ERR_INVALID_ARG, "mode must be RC_DENOISE_ATROUS or RC_DENOISE_NET");
    if (mode == RC_DENOISE_NET && !net_ready(ctx))
        return set_error(ctx, RC_ERR_STATE, "net not initialised (rc_net_init / rc_net_set_params)");
    ctx->denoiser = mode;
    return RC_OK;
}

}  // extern "C"
