// rc_wavefront.hip — progressive path tracing for gfx950, wavefront form.
//
// The per-pixel path loop of pathtrace.comp:454-655 is split into kernels that
// each do one kind of work over a compacted queue, so the lanes of a wave run
// the same code:
//   init    per pixel: G-buffer ray (frame 0), first camera ray
//   extend  closest hit of every queued ray (pathtrace.comp:142-203)
//   shadow  any hit of every queued NEE ray (pathtrace.comp:211-242)
//   shade   per queued path: add the resolved NEE term, shade the hit
//           (emission, roulette, BSDF, NEE setup), or finish the sample,
//           accumulate it in frame order and start the pixel's next sample
// Each pixel owns one path slot and works through its samples in order, so the
// accumulated sums are the reference's sequential ones (bit-identical to the
// megakernel form and to the CPU oracle). A shadow ray issued at step t is
// traced before step t+1's shade, which adds its term before anything else —
// the reference's order of additions into the path radiance.
//
// Queues are appended with one atomic per wave (ballot + prefix popcount).
// Path state lives in HBM as structure-of-arrays float4 records.
#include "rc_device.h"
#include "rc_internal.h"
#include "rc_shade.h"

namespace rc {

namespace {

// flag bits of the slot state (bounce counts up to maxBounces = 16)
constexpr int kBounceMask = 31;
constexpr int kChShift = 5;
constexpr int kLastDiffuse = 1 << 7;
constexpr int kPendingNee = 1 << 8;
constexpr int kEnding = 1 << 9;
constexpr int kNoRay = (int)0x80000000;

struct WfArgs {
    SceneView S;
    CamArgs cam;
    int nee;
    float neeClamp;
    int maxBounces;
    int frame0;
    int nSamples;
    int tilesX;
    int fp64;
    // images
    float4 *accum;
    double4 *accum64;
    ushort4 *gbuf, *galbedo;
    float *variance;
    int *primId, *gbufId;
    // slot state
    float4 *stThr, *stRad, *stNee, *stSmp;
    int *vis;
    // queues (cur -> next)
    const int *listCur;
    const float4 *rayOCur, *rayDCur;
    const float4 *hits;
    int *listNext;
    float4 *rayONext, *rayDNext;
    int *shSlot;
    float4 *shO, *shD;
    int *cnt;          // [0] list count cur, [1] list count next, [2] shadow count
    unsigned long long *counters;
};

__device__ __forceinline__ void slot_pixel(int s, int tilesX, int &px, int &py) {
    int tile = s >> 8, r = s & 255;
    int wave = r >> 6, lane = r & 63;
    px = (tile % tilesX) * 16 + (wave & 1) * 8 + (lane & 7);
    py = (tile / tilesX) * 16 + (wave >> 1) * 8 + (lane >> 3);
}

// Appends one entry per lane with want = true; returns the entry index (or -1).
// Every lane of the wave must call it.
__device__ __forceinline__ int wave_append(int *counter, bool want) {
    unsigned long long mask = __ballot(want);
    if (mask == 0) return -1;
    int lane = __lane_id();
    int leader = __ffsll((long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(mask));
    base = __shfl(base, leader, 64);
    unsigned long long below = lane == 0 ? 0ull : (mask & ((~0ull) >> (64 - lane)));
    return want ? base + __popcll(below) : -1;
}

__device__ __forceinline__ void start_path(const WfArgs &A, int px, int py, int k, int ch, bool lca,
                                           uint32_t &seed, v3 &ro, v3 &rd) {
    const int frame = A.frame0 + k;
    float fOff = 0.0f;
    if (lca) {
        seed = init_seed(px, py, frame * 3 + ch);
        fOff = A.cam.lca[ch];
    } else {
        seed = init_seed(px, py, frame);
    }
    gen_ray(A.cam, px, py, seed, fOff, ro, rd);
}

template <bool STATS>
__global__ void __launch_bounds__(kBlock) wf_init(WfArgs A) {
    __shared__ int2 stk[kShortStack][kBlock];
    const StackRef st{stk, (int)threadIdx.x};
    Counters C;
    if (STATS) C = Counters{};
    const int s = blockIdx.x * kBlock + threadIdx.x;
    int px, py;
    slot_pixel(s, A.tilesX, px, py);
    const bool inside = px < A.cam.width && py < A.cam.height;
    v3 ro = mk(0, 0, 0), rd = mk(0, 0, 1);
    uint32_t seed = 0;
    if (inside) {
        const size_t pi = (size_t)py * (size_t)A.cam.width + (size_t)px;
        if (A.frame0 == 0) {  // G-buffer (pathtrace.comp:590-595, 643-646)
            uint32_t gs = init_seed(px, py, 99991);
            gen_ray(A.cam, px, py, gs, 0.0f, ro, rd);
            Hit h = closest_hit<STATS>(A.S, ro, rd, st, C);
            float g0 = 0.0f, g1 = 1.0f, g2 = 0.0f, g3 = 1e4f;
            if (h.tri >= 0) {
                int m;
                v3 nn = hit_normal(A.S, h, m);
                v3 n = dot(rd, nn) < 0.0f ? nn : neg(nn);
                g0 = n.x; g1 = n.y; g2 = n.z; g3 = h.t;
            }
            A.gbuf[pi] = make_ushort4(rc_f32_to_f16(g0), rc_f32_to_f16(g1), rc_f32_to_f16(g2), rc_f32_to_f16(g3));
            A.gbufId[pi] = h.tri;
        }
        const bool lca = (fabsf(A.cam.lca[0]) + fabsf(A.cam.lca[2])) > 1e-4f;
        start_path(A, px, py, 0, 0, lca, seed, ro, rd);
        if (STATS) C.paths++;
        A.stThr[s] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(seed));
        A.stRad[s] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(0));
        A.stNee[s] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(0));
        A.stSmp[s] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1));
    }
    int at = wave_append(A.cnt, inside);
    if (at >= 0) {
        A.listNext[at] = s;
        A.rayONext[at] = make_float4(ro.x, ro.y, ro.z, 0.0f);
        A.rayDNext[at] = make_float4(rd.x, rd.y, rd.z, 0.0f);
    }
    if (STATS) wave_add_counters(A.counters, C);
}

template <bool STATS>
__global__ void __launch_bounds__(kBlock) wf_extend(WfArgs A, float4 *hits) {
    __shared__ int2 stk[kShortStack][kBlock];
    const StackRef st{stk, (int)threadIdx.x};
    Counters C;
    if (STATS) C = Counters{};
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < A.cnt[0] && A.listCur[i] >= 0) {
        float4 o = A.rayOCur[i], d = A.rayDCur[i];
        Hit h = closest_hit<STATS>(A.S, mk(o.x, o.y, o.z), mk(d.x, d.y, d.z), st, C);
        hits[i] = make_float4(h.t, h.bu, h.bv, __int_as_float(h.tri));
    }
    if (STATS) wave_add_counters(A.counters, C);
}

template <bool STATS>
__global__ void __launch_bounds__(kBlock) wf_shadow(WfArgs A) {
    __shared__ int2 stk[kShortStack][kBlock];
    const StackRef st{stk, (int)threadIdx.x};
    Counters C;
    if (STATS) C = Counters{};
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j < A.cnt[2]) {
        float4 o = A.shO[j], d = A.shD[j];
        A.vis[A.shSlot[j]] = unoccluded<STATS>(A.S, mk(o.x, o.y, o.z), mk(d.x, d.y, d.z), o.w, st, C) ? 1 : 0;
    }
    if (STATS) wave_add_counters(A.counters, C);
}

// Resets the shadow-queue count before the shade kernel refills it.
__global__ void wf_clear_shadow(int *cnt) { cnt[2] = 0; }
// After shade: the next list becomes current.
__global__ void wf_advance(int *cnt) {
    cnt[0] = cnt[1];
    cnt[1] = 0;
}

template <bool STATS>
__global__ void __launch_bounds__(kBlock) wf_shade(WfArgs A) {
    Counters C;
    if (STATS) C = Counters{};
    const int i = blockIdx.x * kBlock + threadIdx.x;
    bool emitRay = false, emitShadow = false, emitNoRay = false;
    v3 ro = mk(0, 0, 0), rd = mk(0, 0, 1);
    NeeSample ns;
    ns.shoot = false;
    int s = 0;
    if (i < A.cnt[0]) {
        const int e = A.listCur[i];
        s = e & 0x7fffffff;
        const bool hasRay = e >= 0;
        int px, py;
        slot_pixel(s, A.tilesX, px, py);
        const size_t pi = (size_t)py * (size_t)A.cam.width + (size_t)px;
        const bool lca = (fabsf(A.cam.lca[0]) + fabsf(A.cam.lca[2])) > 1e-4f;
        float4 T = A.stThr[s], R = A.stRad[s], Nn = A.stNee[s], Sm = A.stSmp[s];
        v3 thr = mk(T.x, T.y, T.z), rad = mk(R.x, R.y, R.z);
        uint32_t seed = __float_as_uint(T.w);
        int flags = __float_as_int(R.w);
        int k = __float_as_int(Nn.w);
        int firstTri = __float_as_int(Sm.w);
        int bounce = flags & kBounceMask, ch = (flags >> kChShift) & 3;
        bool lastDiffuse = (flags & kLastDiffuse) != 0;
        if (flags & kPendingNee) {
            if (A.vis[s]) rad = rad + mk(Nn.x, Nn.y, Nn.z);
            flags &= ~kPendingNee;
        }
        bool finish = (flags & kEnding) != 0;
        if (!finish && hasRay) {
            float4 hv = A.hits[i];
            Hit h;
            h.t = hv.x; h.bu = hv.y; h.bv = hv.z; h.tri = __float_as_int(hv.w);
            float4 o = A.rayOCur[i], d = A.rayDCur[i];
            ro = mk(o.x, o.y, o.z);
            rd = mk(d.x, d.y, d.z);
            bool end;
            const bool firstRec = bounce == 0 && (!lca || ch == 1);
            if (h.tri < 0) {
                end = true;
                if (firstRec && A.frame0 + k == 0) {
                    const uint16_t one = rc_f32_to_f16(1.0f);
                    A.galbedo[pi] = make_ushort4(one, one, one, one);
                }
            } else {
                if (STATS) C.shadeHits++;
                int mi;
                v3 hn = hit_normal(A.S, h, mi);
                v3 hp = ro + rd * h.t;
                const float4 *M = A.S.mats + 4 * (size_t)mi;
                float4 m0 = M[0], m1 = M[1], m2 = M[2];
                if (firstRec) {
                    firstTri = h.tri;
                    if (A.frame0 + k == 0) {
                        v3 fa = (m2.z > 0.5f) ? mk(1.0f, 1.0f, 1.0f) : xyz(m0);
                        A.galbedo[pi] = make_ushort4(rc_f32_to_f16(fa.x), rc_f32_to_f16(fa.y),
                                                     rc_f32_to_f16(fa.z), rc_f32_to_f16(1.0f));
                    }
                }
                const bool last = bounce + 1 >= A.maxBounces;
                end = shade_vertex<STATS>(A.S, A.nee, A.neeClamp, last, bounce, m0, m1, m2, hn, hp, ro, rd, thr,
                                          rad, seed, lastDiffuse, ns, C);
                ++bounce;
                if (last) end = true;
                if (ns.shoot) {
                    flags |= kPendingNee;
                    Nn.x = ns.weight.x; Nn.y = ns.weight.y; Nn.z = ns.weight.z;
                    emitShadow = true;
                }
            }
            if (!end) {
                emitRay = true;
            } else if (ns.shoot) {
                flags |= kEnding;  // finish after the shadow ray has been traced
                emitNoRay = true;
            } else {
                finish = true;
            }
        }
        if (finish) {
            flags &= ~kEnding;
            float sR = Sm.x, sG = Sm.y, sB = Sm.z;
            if (!lca) {
                sR = rad.x; sG = rad.y; sB = rad.z;
            } else if (ch == 0) {
                sR = rad.x;
            } else if (ch == 1) {
                sG = rad.y;
            } else {
                sB = rad.z;
            }
            bool next = true;
            if (lca && ch < 2) {
                ++ch;
            } else {
                // sample complete (pathtrace.comp:633-654)
                ch = 0;
                if (isnan(sR) || isnan(sG) || isnan(sB) || isinf(sR) || isinf(sG) || isinf(sB)) sR = sG = sB = 0.0f;
                float luma = (sR * 0.2126f + sG * 0.7152f) + sB * 0.0722f;
                float l2 = luma * luma;
                if (A.frame0 + k == 0) {
                    A.accum[pi] = make_float4(sR, sG, sB, 1.0f);
                    if (A.fp64) A.accum64[pi] = make_double4(sR, sG, sB, 1.0);
                    A.variance[pi] = l2;
                } else if (A.fp64) {
                    double4 a = A.accum64[pi];
                    a.x += (double)sR; a.y += (double)sG; a.z += (double)sB; a.w += 1.0;
                    A.accum64[pi] = a;
                    A.accum[pi] = make_float4((float)a.x, (float)a.y, (float)a.z, (float)a.w);
                    A.variance[pi] = A.variance[pi] + l2;
                } else {
                    float4 a = A.accum[pi];
                    A.accum[pi] = make_float4(a.x + sR, a.y + sG, a.z + sB, a.w + 1.0f);
                    A.variance[pi] = A.variance[pi] + l2;
                }
                A.primId[pi] = firstTri;
                firstTri = -1;
                ++k;
                next = k < A.nSamples;
            }
            Sm.x = sR; Sm.y = sG; Sm.z = sB;
            if (next) {
                start_path(A, px, py, k, ch, lca, seed, ro, rd);
                if (STATS) C.paths++;
                thr = mk(1.0f, 1.0f, 1.0f);
                rad = mk(0.0f, 0.0f, 0.0f);
                bounce = 0;
                lastDiffuse = false;
                emitRay = true;
            }
        }
        flags = (flags & ~(kBounceMask | (3 << kChShift) | kLastDiffuse)) | bounce | (ch << kChShift) |
                (lastDiffuse ? kLastDiffuse : 0);
        A.stThr[s] = make_float4(thr.x, thr.y, thr.z, __uint_as_float(seed));
        A.stRad[s] = make_float4(rad.x, rad.y, rad.z, __int_as_float(flags));
        A.stNee[s] = make_float4(Nn.x, Nn.y, Nn.z, __int_as_float(k));
        A.stSmp[s] = make_float4(Sm.x, Sm.y, Sm.z, __int_as_float(firstTri));
    }
    int at = wave_append(&A.cnt[1], emitRay || emitNoRay);
    if (at >= 0) {
        A.listNext[at] = emitNoRay ? (s | kNoRay) : s;
        if (emitRay) {
            A.rayONext[at] = make_float4(ro.x, ro.y, ro.z, 0.0f);
            A.rayDNext[at] = make_float4(rd.x, rd.y, rd.z, 0.0f);
        }
    }
    int sj = wave_append(&A.cnt[2], emitShadow);
    if (sj >= 0) {
        A.shSlot[sj] = s;
        A.shO[sj] = make_float4(ns.o.x, ns.o.y, ns.o.z, ns.maxDist);
        A.shD[sj] = make_float4(ns.d.x, ns.d.y, ns.d.z, 0.0f);
    }
    if (STATS) wave_add_counters(A.counters, C);
}

}  // namespace

void free_wavefront(Wavefront &w) {
    void *ptrs[] = {w.stThr, w.stRad, w.stNee, w.stSmp, w.vis, w.list[0], w.list[1], w.rayO[0], w.rayO[1],
                    w.rayD[0], w.rayD[1], w.hits, w.shSlot, w.shO, w.shD, w.cnt};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (w.hostCnt) (void)hipHostFree(w.hostCnt);
    w = Wavefront{};
}

static hipError_t ensure_wavefront(rc_ctx *ctx, int slots) {
    Wavefront &w = ctx->wf;
    if (w.slots >= slots) return hipSuccess;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return e;
    free_wavefront(w);
    size_t n = (size_t)slots;
    auto A = [&](void **p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes);
    };
    A((void **)&w.stThr, n * 16); A((void **)&w.stRad, n * 16); A((void **)&w.stNee, n * 16);
    A((void **)&w.stSmp, n * 16); A((void **)&w.vis, n * 4);
    for (int b = 0; b < 2; ++b) {
        A((void **)&w.list[b], n * 4); A((void **)&w.rayO[b], n * 16); A((void **)&w.rayD[b], n * 16);
    }
    A((void **)&w.hits, n * 16); A((void **)&w.shSlot, n * 4); A((void **)&w.shO, n * 16);
    A((void **)&w.shD, n * 16); A((void **)&w.cnt, 64);
    if (e == hipSuccess) e = hipHostMalloc((void **)&w.hostCnt, 64, hipHostMallocDefault);
    if (e != hipSuccess) {
        free_wavefront(w);
        return e;
    }
    w.slots = slots;
    return hipSuccess;
}

hipError_t launch_wavefront(rc_ctx *ctx, const rc_camera &cam, const rc_pt_params &p, int nSamples) {
    const int tilesX = (cam.width + 15) / 16, tilesY = (cam.height + 15) / 16;
    const int slots = tilesX * tilesY * 256;
    hipError_t e = ensure_wavefront(ctx, slots);
    if (e != hipSuccess) return e;
    Wavefront &w = ctx->wf;
    WfArgs A;
    A.S = scene_view(ctx->scene);
    A.cam = cam_args(cam);
    A.nee = p.neeEnabled;
    A.neeClamp = p.neeFireflyClamp;
    A.maxBounces = p.maxBounces;
    A.frame0 = ctx->totalSamples;
    A.nSamples = nSamples;
    A.tilesX = tilesX;
    A.fp64 = ctx->accumMode == RC_ACCUM_FP64 ? 1 : 0;
    A.accum = ctx->img.accum;
    A.accum64 = ctx->img.accum64;
    A.gbuf = ctx->img.gbuf;
    A.galbedo = ctx->img.galbedo;
    A.variance = ctx->img.variance;
    A.primId = ctx->img.primId;
    A.gbufId = ctx->img.gbufId;
    A.stThr = w.stThr; A.stRad = w.stRad; A.stNee = w.stNee; A.stSmp = w.stSmp;
    A.vis = w.vis;
    A.shSlot = w.shSlot; A.shO = w.shO; A.shD = w.shD;
    A.cnt = w.cnt;
    A.hits = w.hits;
    A.counters = ctx->counters;
    const bool stats = ctx->countersEnabled && ctx->counters;
    const dim3 grid(slots / kBlock), block(kBlock);
    hipStream_t sm = ctx->stream;

    e = hipMemsetAsync(w.cnt, 0, 64, sm);
    if (e != hipSuccess) return e;
    int cur = 0;
    A.listCur = w.list[1]; A.rayOCur = w.rayO[1]; A.rayDCur = w.rayD[1];
    A.listNext = w.list[cur]; A.rayONext = w.rayO[cur]; A.rayDNext = w.rayD[cur];
    // init appends into cnt[0] (the "next" counter slot is cnt[1] later; init writes the current list)
    if (stats) hipLaunchKernelGGL(wf_init<true>, grid, block, 0, sm, A);
    else hipLaunchKernelGGL(wf_init<false>, grid, block, 0, sm, A);
    if ((e = hipGetLastError()) != hipSuccess) return e;

    const bool lca = (fabsf(cam.lcaDelta[0]) + fabsf(cam.lcaDelta[2])) > 1e-4f;
    const long long maxIter = (long long)nSamples * (lca ? 3 : 1) * (p.maxBounces + 2) + 4;
    bool drained = false;
    for (long long it = 0; it < maxIter && !drained; ++it) {
        const int nxt = cur ^ 1;
        A.listCur = w.list[cur]; A.rayOCur = w.rayO[cur]; A.rayDCur = w.rayD[cur];
        A.listNext = w.list[nxt]; A.rayONext = w.rayO[nxt]; A.rayDNext = w.rayD[nxt];
        if (stats) {
            hipLaunchKernelGGL(wf_extend<true>, grid, block, 0, sm, A, w.hits);
            hipLaunchKernelGGL(wf_shadow<true>, grid, block, 0, sm, A);
            hipLaunchKernelGGL(wf_clear_shadow, dim3(1), dim3(1), 0, sm, w.cnt);
            hipLaunchKernelGGL(wf_shade<true>, grid, block, 0, sm, A);
        } else {
            hipLaunchKernelGGL(wf_extend<false>, grid, block, 0, sm, A, w.hits);
            hipLaunchKernelGGL(wf_shadow<false>, grid, block, 0, sm, A);
            hipLaunchKernelGGL(wf_clear_shadow, dim3(1), dim3(1), 0, sm, w.cnt);
            hipLaunchKernelGGL(wf_shade<false>, grid, block, 0, sm, A);
        }
        hipLaunchKernelGGL(wf_advance, dim3(1), dim3(1), 0, sm, w.cnt);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        cur = nxt;
        if ((it & 15) == 15) {
            e = hipMemcpyAsync(w.hostCnt, w.cnt, 16, hipMemcpyDeviceToHost, sm);
            if (e == hipSuccess) e = hipStreamSynchronize(sm);
            if (e != hipSuccess) return e;
            drained = w.hostCnt[0] == 0 && w.hostCnt[2] == 0;
        }
    }
    if (!drained) {  // maxIter bounds every path; check that the queues are empty
        e = hipMemcpyAsync(w.hostCnt, w.cnt, 16, hipMemcpyDeviceToHost, sm);
        if (e == hipSuccess) e = hipStreamSynchronize(sm);
        if (e != hipSuccess) return e;
        if (w.hostCnt[0] != 0 || w.hostCnt[2] != 0) return hipErrorUnknown;
    }
    return hipSuccess;
}

}  // namespace rc
