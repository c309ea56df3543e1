// rc_pathtrace.hip — progressive path tracing for gfx950, megakernel form.
//
// One work-item per pixel; a launch runs nSamples consecutive progressive
// passes (frame indices f0 .. f0+nSamples-1) and keeps the running sums in
// registers, so the accumulation buffers are read and written once per launch.
// The per-sample arithmetic restates pathtrace.comp:401-655 (raygen,
// traceRadiance, NEE, accumulate) with the same operation order.
//
// Scheduling: each lane owns one pixel and regenerates its path state as soon
// as a path ends (next LCA channel, else next sample), so the lanes of a wave
// stay busy until every lane has finished all its samples. Samples of a pixel
// are accumulated in frame order, so every sum is the reference's sequential
// one. Traversal stacks live in LDS (rc_device.h). The wavefront form
// (rc_wavefront.hip) computes the same results with separate extend / shadow /
// shade kernels.
//
// Launch shape: 256-thread workgroups covering 16x16 pixel tiles, each wave64
// an 8x8 sub-tile so camera rays of one wave are spatially coherent.
#include "rc_device.h"
#include "rc_internal.h"
#include "rc_shade.h"

namespace rc {

// All samples of one pixel (px, py) of the pass described by A.
template <bool STATS, bool FP64>
__device__ __forceinline__ void trace_pixel(const PassArgs &A, int px, int py, const StackRef &st, Counters &C) {
    const SceneView &S = A.S;
    const CamArgs &cam = launch_camera(A);
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
            bool end;
            Hit h = closest_hit<STATS>(S, ro, rd, st, C);
            const bool firstRec = bounce == 0 && (!lca || ch == 1);
            if (h.tri < 0) {
                end = true;  // black sky (pathtrace.comp:470-473)
                if (firstRec && A.frame0 + k == 0) {
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
                if (firstRec) {
                    // first-hit albedo (pathtrace.comp:481-483); the G channel's under LCA
                    firstTri = h.tri;
                    if (A.frame0 + k == 0) {
                        v3 fa = (m2.z > 0.5f) ? mk(1.0f, 1.0f, 1.0f) : xyz(m0);
                        A.galbedo[pi] = make_ushort4(rc_f32_to_f16(fa.x), rc_f32_to_f16(fa.y),
                                                     rc_f32_to_f16(fa.z), rc_f32_to_f16(1.0f));
                    }
                }
                const bool last = bounce + 1 >= A.maxBounces;
                NeeSample ns;
                end = shade_vertex<STATS>(S, A.nee, A.neeClamp, last, bounce, m0, m1, m2, hn, hp, ro, rd, thr, rad,
                                          seed, lastDiffuse, ns, C);
                if (ns.shoot && unoccluded<STATS>(S, ns.o, ns.d, ns.maxDist, st, C)) rad = rad + ns.weight;
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
}

// Occupancy: 7 waves per SIMD (<= 72 VGPRs) beats the unconstrained 4 waves
// (111 VGPRs) by ~20% on the full frame; the divergent traversal is latency bound.
template <bool STATS, bool FP64>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(7, 8))) pathtrace_kernel(PassArgs A) {
    __shared__ int2 stk[kShortStack][kBlock];
    const StackRef st{stk, (int)threadIdx.x};
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int px = blockIdx.x * 16 + (wave & 1) * 8 + (lane & 7);
    const int py = blockIdx.y * 16 + (wave >> 1) * 8 + (lane >> 3);
    Counters C;
    if (STATS) C = Counters{};
    trace_pixel<STATS, FP64>(A, px, py, st, C);
    if (STATS) wave_add_counters(A.counters, C);
}

// Many views (own scene, camera, sample count and images) in one persistent
// launch: each workgroup takes 16x16 tiles from a global counter until all
// tiles of all views are done. tileOff[v] is the first tile of view v (views
// are laid out heaviest first, so long tiles start early).
template <bool STATS, bool FP64>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(7, 8))) pathtrace_views_kernel(const PassArgs *views, const int *tileOff,
                                                                 int nViews, int *counter,
                                                                 unsigned long long *counters) {
    __shared__ int2 stk[kShortStack][kBlock];
    __shared__ int tileSh;
    const StackRef st{stk, (int)threadIdx.x};
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int totalTiles = tileOff[nViews];
    Counters C;
    if (STATS) C = Counters{};
    while (true) {
        if (threadIdx.x == 0) tileSh = atomicAdd(counter, 1);
        __syncthreads();
        const int t = tileSh;
        __syncthreads();
        if (t >= totalTiles) break;
        int lo = 0, hi = nViews;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (tileOff[mid] <= t) lo = mid;
            else hi = mid;
        }
        const PassArgs &A = views[lo];
        const int lt = t - tileOff[lo];
        const int tilesX = (A.cam.width + 15) >> 4;
        const int px = (lt % tilesX) * 16 + (wave & 1) * 8 + (lane & 7);
        const int py = (lt / tilesX) * 16 + (wave >> 1) * 8 + (lane >> 3);
        trace_pixel<STATS, FP64>(A, px, py, st, C);
    }
    if (STATS) wave_add_counters(counters, C);
}

__global__ void zero_counter(int *c) { *c = 0; }

CamArgs cam_args(const rc_camera &cam) {
    CamArgs c;
    for (int k = 0; k < 3; ++k) {
        c.pos[k] = cam.camPos[k]; c.right[k] = cam.camRight[k];
        c.up[k] = cam.camUp[k]; c.fwd[k] = cam.camForward[k];
        c.lca[k] = cam.lcaDelta[k];
    }
    c.focalDist = cam.focalDist;
    c.lensRadius = cam.lensRadius;
    c.tanHalf = cam.fovTanHalf;
    c.saStrength = cam.saStrength;
    c.blades = cam.apertureBlades;
    c.width = cam.width;
    c.height = cam.height;
    return c;
}

PassArgs pass_args(rc_ctx *ctx, const rc_camera &cam, const rc_pt_params &p, int nSamples) {
    PassArgs A;
    A.S = scene_view(ctx->scene);
    A.cam = cam_args(cam);
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
    A.tileShift = ctx->laneTileShift;
    A.dcam = nullptr;  // the caller sets it (camera_slot)
    return A;
}

hipError_t launch_pathtrace(rc_ctx *ctx, const rc_camera &cam, const rc_pt_params &p, int nSamples,
                            const CamArgs *dcam) {
    PassArgs A = pass_args(ctx, cam, p, nSamples);
    A.dcam = dcam;
    dim3 grid((cam.width + 15) / 16, (cam.height + 15) / 16);
    const bool stats = ctx->countersEnabled && ctx->counters;
    const bool fp64 = ctx->accumMode == RC_ACCUM_FP64;
    if (stats && fp64) hipLaunchKernelGGL((pathtrace_kernel<true, true>), grid, dim3(kBlock), 0, ctx->stream, A);
    else if (stats) hipLaunchKernelGGL((pathtrace_kernel<true, false>), grid, dim3(kBlock), 0, ctx->stream, A);
    else if (fp64) hipLaunchKernelGGL((pathtrace_kernel<false, true>), grid, dim3(kBlock), 0, ctx->stream, A);
    else hipLaunchKernelGGL((pathtrace_kernel<false, false>), grid, dim3(kBlock), 0, ctx->stream, A);
    return hipGetLastError();
}

hipError_t launch_pathtrace_views(rc_ctx *ctx, const PassArgs *views, const int *tileOff, int nViews, int *counter,
                                  int maxBlocks, bool fp64) {
    const bool stats = ctx->countersEnabled && ctx->counters;
    static int occ[4] = {0, 0, 0, 0};
    const int variant = (stats ? 2 : 0) + (fp64 ? 1 : 0);
    auto kernel = stats ? (fp64 ? pathtrace_views_kernel<true, true> : pathtrace_views_kernel<true, false>)
                        : (fp64 ? pathtrace_views_kernel<false, true> : pathtrace_views_kernel<false, false>);
    if (occ[variant] == 0) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, kBlock, 0) != hipSuccess || nb < 1) nb = 1;
        occ[variant] = nb;
    }
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
    int grid = occ[variant] * cus;
    if (grid > maxBlocks) grid = maxBlocks;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(zero_counter, dim3(1), dim3(1), 0, ctx->stream, counter);
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(kBlock), 0, ctx->stream, views, tileOff, nViews, counter,
                       ctx->counters);
    return hipGetLastError();
}

}  // namespace rc
