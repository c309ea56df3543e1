// rc_pathtrace_lanes.hip — progressive path tracing for gfx950 with one state
// machine per lane and per-sample work units ("lane scheduler" form).
//
// Why: in the megakernel (rc_pathtrace.hip) a wave walks the BVH until its
// slowest lane is done, then shades, then walks shadow rays until the slowest
// is done, and a lane whose pixel is cheap idles until the wave's most
// expensive pixel has finished all its samples. Here:
//   * a work unit is ONE sample (all LCA channels) of one pixel; lanes pull
//     units from a per-view atomic counter (one atomic per wave and round), so
//     no lane waits for another lane's pixel;
//   * each lane is a state machine: WALK (one BVH step per round), SHADE (hit
//     processing, NEE, next ray), FETCH (take the next unit); the wave keeps
//     stepping walks until `thresh` lanes are waiting, then shades them
//     together;
//   * the sample's radiance goes to an HBM sample buffer [sample][tiled pixel];
//     lanes_reduce then adds the samples of every pixel in frame order, which
//     is exactly the reference's sequential running sum (pathtrace.comp:633-654).
// Units are numbered tile-major (8x8 pixel tiles, then sample, then pixel in
// tile), so the 64 units a wave takes at once are one tile at one sample. A
// launch that holds frame 0 ends with one more unit per pixel, its G-buffer
// camera ray, so that single rays fill the time in which the last paths finish.
//
// Shadow rays use the same ordered walk as closest hits, with the walk's
// maximum t set to maxDist and an exit on the first accepted triangle. For a
// BVH shallower than 30 levels this equals shadowRay (pathtrace.comp:211-242):
// neither stack cap (62, 30) can be reached, every box test is the same
// computation on the same operands, and an any-hit answer does not depend on
// the visiting order. The host uses this form only for such scenes.
//
// Per-lane arithmetic (raygen, walks, shading, NEE) is the shared code of
// rc_device.h / rc_shade.h, so images are bit-identical to the megakernel's.
#include <map>
#include <mutex>
#include <utility>

#include "rc_device.h"
#include "rc_internal.h"
#include "rc_shade.h"

namespace rc {

namespace {

enum : int { PH_WALK = 0, PH_SHADE = 1, PH_FETCH = 2, PH_DONE = 3, PH_START = 4 };
enum : int { RAY_GBUF = 0, RAY_PATH = 1, RAY_SHADOW = 2 };

// An ordered BVH walk in progress (traverseBVH, pathtrace.comp:142-203, as in
// closest_hit), resumable one step at a time.
struct Walk {
    v3 ro, rd, inv;
    float ht, hbu, hbv;
    int htri;
    int ref;
    float tEntry;
    int top;
    bool have;
    int li, le;  // one-triangle steps: next and end triangle of the current leaf
};

__device__ __forceinline__ void walk_begin(const SceneView &S, Walk &w, float tmax) {
    w.inv = mk(1.0f / w.rd.x, 1.0f / w.rd.y, 1.0f / w.rd.z);
    w.ht = tmax;
    w.htri = -1;
    w.hbu = 0.0f;
    w.hbv = 0.0f;
    w.top = 0;
    w.li = 0;
    w.le = 0;
    float tr;
    w.have = aabb(w.ro, w.inv, S.rootMin[0], S.rootMin[1], S.rootMin[2], S.rootMax[0], S.rootMax[1], S.rootMax[2],
                  tr);
    w.ref = S.rootRef;
    w.tEntry = tr;
}

// Stack entries below NS live in LDS, deeper ones in the private spill array.
// The LDS read is issued unconditionally (clamped index) and the spill read only
// for deep entries. Deep entries are stored as one 64-bit word: loads of two
// different types cannot be folded into one flat load through a selected
// LDS-or-scratch address, so the common pop stays a ds_read.
template <int NS>
__device__ __forceinline__ int2 stack_pop(const StackRef &st, const unsigned long long *spill, int top) {
    int2 e = st.lds[top < NS ? top : NS - 1][st.lane];
    if (top >= NS) {
        const unsigned long long v = spill[top - NS];
        e = make_int2((int)(uint32_t)v, (int)(uint32_t)(v >> 32));
    }
    return e;
}

template <int NS>
__device__ __forceinline__ void stack_push(const StackRef &st, unsigned long long *spill, int top, int2 e) {
    if (top < NS) st.lds[top][st.lane] = e;
    else spill[top - NS] = (unsigned long long)(uint32_t)e.x | ((unsigned long long)(uint32_t)e.y << 32);
}

// One node (or one leaf's triangles) of the walk; returns true when the walk has
// ended. anyHit: stop at the first triangle with t < ht (shadow rays).
template <bool STATS, int NS>
__device__ __forceinline__ bool walk_step(const SceneView &S, Walk &w, bool anyHit, const StackRef &st, unsigned long long *spill,
                                          Counters &C) {
    if (!w.have) {
        if (w.top == 0) return true;
        --w.top;
        const int2 e = stack_pop<NS>(st, spill, w.top);
        w.ref = e.x;
        w.tEntry = __int_as_float(e.y);
        if (STATS) C.nodePops++;
    }
    w.have = false;
    if (w.tEntry > w.ht) return w.top == 0;
    if (w.ref < 0) {
        int start, count;
        leaf_range(S, w.ref, start, count);
        for (int i = start; i < start + count; ++i) {
            float t, bu, bv;
            if (STATS) C.triTests++;
            if (tri_test(S, w.ro, w.rd, i, t, bu, bv) && t < w.ht) {
                w.ht = t;
                w.hbu = bu;
                w.hbv = bv;
                w.htri = i;
                if (STATS) C.hitUpdates++;
                if (anyHit) return true;
            }
        }
        return w.top == 0;
    }
    if (w.top < 62) {
        const NodeRec nd = node_rec(S, w.ref);
        const float4 a = nd.a, b = nd.b, c = nd.c;
        const int2 r = nd.r;
        if (STATS) C.childTests++;
        float tl, trr;
        const bool hl = aabb(w.ro, w.inv, a.x, a.y, a.z, a.w, b.x, b.y, tl);
        const bool hr = aabb(w.ro, w.inv, b.z, b.w, c.x, c.y, c.z, c.w, trr);
        if (hl || hr) {
            if (STATS) C.nodePops++;
            w.have = true;  // the child the reference pushes last is popped next
            if (hl && hr) {
                const bool nearLeft = tl < trr;
                const int2 far = nearLeft ? make_int2(r.y, __float_as_int(trr)) : make_int2(r.x, __float_as_int(tl));
                stack_push<NS>(st, spill, w.top, far);
                ++w.top;
                w.ref = nearLeft ? r.x : r.y;
                w.tEntry = nearLeft ? tl : trr;
            } else {
                w.ref = hl ? r.x : r.y;
                w.tEntry = hl ? tl : trr;
            }
        }
    }
    return !w.have && w.top == 0;
}

// The same walk with one triangle test per step: a leaf's triangles are tested
// over consecutive steps (same order, same acceptance), so lanes in leaves and
// lanes at inner nodes each run one short body per step.
template <bool STATS, int NS>
__device__ __forceinline__ bool walk_step_tri(const SceneView &S, Walk &w, bool anyHit, const StackRef &st,
                                              unsigned long long *spill, Counters &C) {
    if (w.li >= w.le) {
        if (!w.have) {
            if (w.top == 0) return true;
            --w.top;
            const int2 e = stack_pop<NS>(st, spill, w.top);
            w.ref = e.x;
            w.tEntry = __int_as_float(e.y);
            if (STATS) C.nodePops++;
        }
        w.have = false;
        if (w.tEntry > w.ht) return w.top == 0;
        if (w.ref >= 0) {
            if (w.top < 62) {
                const NodeRec nd = node_rec(S, w.ref);
                const float4 a = nd.a, b = nd.b, c = nd.c;
                const int2 r = nd.r;
                if (STATS) C.childTests++;
                float tl, trr;
                const bool hl = aabb(w.ro, w.inv, a.x, a.y, a.z, a.w, b.x, b.y, tl);
                const bool hr = aabb(w.ro, w.inv, b.z, b.w, c.x, c.y, c.z, c.w, trr);
                if (hl || hr) {
                    if (STATS) C.nodePops++;
                    w.have = true;
                    if (hl && hr) {
                        const bool nearLeft = tl < trr;
                        const int2 far =
                            nearLeft ? make_int2(r.y, __float_as_int(trr)) : make_int2(r.x, __float_as_int(tl));
                        stack_push<NS>(st, spill, w.top, far);
                        ++w.top;
                        w.ref = nearLeft ? r.x : r.y;
                        w.tEntry = nearLeft ? tl : trr;
                    } else {
                        w.ref = hl ? r.x : r.y;
                        w.tEntry = hl ? tl : trr;
                    }
                }
            }
            return !w.have && w.top == 0;
        }
        int start, count;
        leaf_range(S, w.ref, start, count);
        w.li = start;
        w.le = start + count;
    }
    float t, bu, bv;
    if (STATS) C.triTests++;
    const int i = w.li++;
    if (tri_test(S, w.ro, w.rd, i, t, bu, bv) && t < w.ht) {
        w.ht = t;
        w.hbu = bu;
        w.hbv = bv;
        w.htri = i;
        if (STATS) C.hitUpdates++;
        if (anyHit) return true;
    }
    return w.li >= w.le && w.top == 0;
}

// The walk with one triangle test AND one node operation per step: a lane tests
// the next triangle of its pending leaf and, once that leaf has no triangle
// left, pops or expands the next node in the same step. Leaves are tested in the
// reference's order and each leaf's triangles in order; only the cull threshold
// can be stale (not yet lowered by the pending leaf), which visits extra nodes
// whose triangles all lie beyond the final closest hit (t >= entry > threshold
// >= final t), so they can neither win nor tie. Any-hit walks do not depend on
// the order at all. Lanes in leaves and lanes at inner nodes thus share steps.
template <bool STATS, int NS>
__device__ __forceinline__ bool walk_step_dual(const SceneView &S, Walk &w, bool anyHit, const StackRef &st,
                                               unsigned long long *spill, Counters &C) {
    if (w.li < w.le) {
        float t, bu, bv;
        if (STATS) C.triTests++;
        const int i = w.li++;
        if (tri_test(S, w.ro, w.rd, i, t, bu, bv) && t < w.ht) {
            w.ht = t;
            w.hbu = bu;
            w.hbv = bv;
            w.htri = i;
            if (STATS) C.hitUpdates++;
            if (anyHit) return true;
        }
    }
    if (w.li >= w.le) {
        if (!w.have) {
            if (w.top == 0) return true;
            --w.top;
            const int2 e = stack_pop<NS>(st, spill, w.top);
            w.ref = e.x;
            w.tEntry = __int_as_float(e.y);
            if (STATS) C.nodePops++;
        }
        w.have = false;
        if (w.tEntry <= w.ht) {
            if (w.ref < 0) {
                int start, count;
                leaf_range(S, w.ref, start, count);
                w.li = start;
                w.le = start + count;
            } else if (w.top < 62) {
                const NodeRec nd = node_rec(S, w.ref);
                const float4 a = nd.a, b = nd.b, c = nd.c;
                const int2 r = nd.r;
                if (STATS) C.childTests++;
                float tl, trr;
                const bool hl = aabb(w.ro, w.inv, a.x, a.y, a.z, a.w, b.x, b.y, tl);
                const bool hr = aabb(w.ro, w.inv, b.z, b.w, c.x, c.y, c.z, c.w, trr);
                if (hl || hr) {
                    if (STATS) C.nodePops++;
                    w.have = true;
                    if (hl && hr) {
                        const bool nearLeft = tl < trr;
                        const int2 far =
                            nearLeft ? make_int2(r.y, __float_as_int(trr)) : make_int2(r.x, __float_as_int(tl));
                        stack_push<NS>(st, spill, w.top, far);
                        ++w.top;
                        w.ref = nearLeft ? r.x : r.y;
                        w.tEntry = nearLeft ? tl : trr;
                    } else {
                        w.ref = hl ? r.x : r.y;
                        w.tEntry = hl ? tl : trr;
                    }
                }
            }
        }
    }
    return w.li >= w.le && !w.have && w.top == 0;
}

// Packed per-lane path flags: ray kind (2 bits), LCA channel (2), bounce (5),
// lastDiffuse, and whether the path ends after the pending shadow ray.
constexpr uint32_t F_LASTDIFF = 1u << 9, F_CONTEND = 1u << 10;
__device__ __forceinline__ uint32_t f_pack(int kind, int ch, int bounce, bool lastDiffuse, bool contEnd) {
    return (uint32_t)kind | ((uint32_t)ch << 2) | ((uint32_t)bounce << 4) | (lastDiffuse ? F_LASTDIFF : 0u) |
           (contEnd ? F_CONTEND : 0u);
}

// All units of one view (K samples from frame frame0, every pixel; A gives the
// scene, camera, switches and images): runs until the view's counter is
// exhausted and every lane of the wave is done.
// samples: [K][tilesX * tilesY * 64] float4 (r, g, b, 0) per unit.
// cont: [3][gridDim.x * kBlock] float4, one slot per lane: the path's next ray
// and NEE weight while its shadow ray is walked (kept out of registers).
template <bool STATS, int NS, int WALK>
__device__ void lanes_view(const PassArgs &A, int frame0, int K, float4 *__restrict__ samples, int *counter,
                           int thresh, float4 *__restrict__ cont, const StackRef &st, unsigned long long *spill, Counters &C) {
    const SceneView &S = A.S;
    const CamArgs &cam = launch_camera(A);
    const int lane = threadIdx.x & 63;
    const bool lca = (fabsf(cam.lca[0]) + fabsf(cam.lca[2])) > 1e-4f;
    const LaneTile tl(cam.width, cam.height, A.tileShift);
    const int nTilePx = tl.pixels();
    // Units: K path samples of every pixel, then, when the launch holds frame 0,
    // one G-buffer camera ray per pixel (s == K). The G-buffer rays come last
    // because they are single rays: while the last paths finish, the other lanes
    // take them, so the launch's end is not the long paths' alone.
    const int nPathUnits = nTilePx * K;
    const int nUnits = nPathUnits + (frame0 == 0 ? nTilePx : 0);  // <= 2^30, checked by the host
    const unsigned long long below = (1ull << lane) - 1ull;
    const size_t nLanes = (size_t)gridDim.x * kBlock;
    float4 *const myCont = cont + (size_t)blockIdx.x * kBlock + threadIdx.x;

    // An exhausted view costs one load per wave, not an atomic.
    if (__atomic_load_n(counter, __ATOMIC_RELAXED) >= nUnits) return;

    int phase = PH_FETCH;
    uint32_t flags = 0, pxy = 0;  // pxy: px | py << 16
    int s = 0, firstTri = -1;
    uint32_t seed = 0;
    v3 thr = mk(1.0f, 1.0f, 1.0f), rad = mk(0.0f, 0.0f, 0.0f);
    Walk w;
    w.ro = mk(0.0f, 0.0f, 0.0f);
    w.rd = mk(0.0f, 0.0f, 1.0f);
    w.inv = w.rd;
    w.ht = 0.0f; w.hbu = 0.0f; w.hbv = 0.0f; w.htri = -1;
    w.ref = 0; w.tEntry = 0.0f; w.top = 0; w.have = false; w.li = 0; w.le = 0;

    while (true) {
        // (1) lanes without work take units: one atomic per wave per round.
        while (true) {
            const unsigned long long need = __ballot(phase == PH_FETCH);
            if (need == 0ull) break;
            const int leader = __ffsll((unsigned long long)need) - 1;
            int base = 0;
            if (lane == leader) base = atomicAdd(counter, __popcll(need));
            base = __shfl(base, leader, 64);
            if (phase == PH_FETCH) {
                const int u = base + __popcll(need & below);
                if (u >= nUnits) {
                    phase = PH_DONE;
                } else {
                    const int within = u & 63;
                    const int ts = u >> 6;  // tile * K + sample, or a G-buffer tile
                    int tile;
                    if (u < nPathUnits) {
                        tile = ts / K;
                        s = ts - tile * K;
                    } else {
                        tile = ts - (nPathUnits >> 6);
                        s = K;
                    }
                    int px, py;
                    tl.pixel(tile, within, px, py);
                    pxy = (uint32_t)px | ((uint32_t)py << 16);
                    if (px < cam.width && py < cam.height) phase = PH_START;
                }
            }
        }

        // (2) set-up of new units and processing of finished walks.
        if (phase == PH_START || phase == PH_SHADE) {
            const int px = (int)(pxy & 0xffffu), py = (int)(pxy >> 16);
            const size_t pi = (size_t)py * (size_t)cam.width + (size_t)px;
            const int frame = frame0 + s;
            const int kind = (int)(flags & 3u);
            int ch = (int)((flags >> 2) & 3u), bounce = (int)((flags >> 4) & 31u);
            bool lastDiffuse = (flags & F_LASTDIFF) != 0u, contEnd = (flags & F_CONTEND) != 0u;
            bool endPath = false, newPath = false, trace = false;
            int nextKind = RAY_PATH;
            float tmax = 1e30f;
            if (phase == PH_START) {
                ch = 0;
                firstTri = -1;
                if (s == K) {
                    // G-buffer from the seed-99991 camera ray (pathtrace.comp:590-595, 643-646)
                    uint32_t gs = init_seed(px, py, 99991);
                    gen_ray(cam, px, py, gs, 0.0f, w.ro, w.rd);
                    nextKind = RAY_GBUF;
                    trace = true;
                } else {
                    newPath = true;
                }
            } else if (kind == RAY_GBUF) {
                float g0 = 0.0f, g1 = 1.0f, g2 = 0.0f, g3 = 1e4f;
                if (w.htri >= 0) {
                    Hit h;
                    h.t = w.ht; h.bu = w.hbu; h.bv = w.hbv; h.tri = w.htri;
                    int m;
                    const v3 nn = hit_normal(S, h, m);
                    const v3 n = dot(w.rd, nn) < 0.0f ? nn : neg(nn);
                    g0 = n.x; g1 = n.y; g2 = n.z; g3 = h.t;
                }
                A.gbuf[pi] = make_ushort4(rc_f32_to_f16(g0), rc_f32_to_f16(g1), rc_f32_to_f16(g2), rc_f32_to_f16(g3));
                A.gbufId[pi] = w.htri;
                phase = PH_FETCH;  // a G-buffer unit is this one ray
            } else if (kind == RAY_SHADOW) {
                const float4 c0 = myCont[0], c1 = myCont[nLanes];
                if (w.htri < 0) {
                    const float wz = myCont[2 * nLanes].x;
                    rad = rad + mk(c0.w, c1.w, wz);  // unoccluded: add the NEE sample (pathtrace.comp:573)
                }
                if (contEnd) {
                    endPath = true;
                } else {
                    w.ro = mk(c0.x, c0.y, c0.z);
                    w.rd = mk(c1.x, c1.y, c1.z);
                    trace = true;
                }
            } else {
                const bool firstRec = bounce == 0 && (!lca || ch == 1);
                if (w.htri < 0) {
                    endPath = true;  // black sky (pathtrace.comp:470-473)
                    if (firstRec && frame == 0) {
                        const uint16_t one = rc_f32_to_f16(1.0f);  // miss: albedo stays (1,1,1) (:455)
                        A.galbedo[pi] = make_ushort4(one, one, one, one);
                    }
                } else {
                    if (STATS) C.shadeHits++;
                    Hit h;
                    h.t = w.ht; h.bu = w.hbu; h.bv = w.hbv; h.tri = w.htri;
                    int mi;
                    const v3 hn = hit_normal(S, h, mi);
                    const v3 hp = w.ro + w.rd * h.t;
                    const float4 *M = S.mats + 4 * (size_t)mi;
                    const float4 m0 = M[0], m1 = M[1], m2 = M[2];
                    if (firstRec) {
                        // first-hit albedo (pathtrace.comp:481-483); the G channel's under LCA
                        firstTri = h.tri;
                        if (frame == 0) {
                            const v3 fa = (m2.z > 0.5f) ? mk(1.0f, 1.0f, 1.0f) : xyz(m0);
                            A.galbedo[pi] = make_ushort4(rc_f32_to_f16(fa.x), rc_f32_to_f16(fa.y),
                                                         rc_f32_to_f16(fa.z), rc_f32_to_f16(1.0f));
                        }
                    }
                    const bool last = bounce + 1 >= A.maxBounces;
                    NeeSample ns;
                    const bool rr = shade_vertex<STATS>(S, A.nee, A.neeClamp, last, bounce, m0, m1, m2, hn, hp, w.ro,
                                                        w.rd, thr, rad, seed, lastDiffuse, ns, C);
                    ++bounce;
                    if (rr) {
                        endPath = true;
                    } else if (ns.shoot) {
                        myCont[0] = make_float4(w.ro.x, w.ro.y, w.ro.z, ns.weight.x);
                        myCont[nLanes] = make_float4(w.rd.x, w.rd.y, w.rd.z, ns.weight.y);
                        myCont[2 * nLanes] = make_float4(ns.weight.z, 0.0f, 0.0f, 0.0f);
                        contEnd = last;
                        w.ro = ns.o;
                        w.rd = ns.d;
                        tmax = ns.maxDist;
                        nextKind = RAY_SHADOW;
                        trace = true;
                    } else if (last) {
                        endPath = true;
                    } else {
                        trace = true;
                    }
                }
            }
            if (endPath) {
                // channel radiance: the sample itself, or one LCA channel whose partial
                // results wait in the sample's slot (pathtrace.comp:611-637)
                const size_t slot = (size_t)s * (size_t)nTilePx + (size_t)tl.slot(px, py);
                if (lca && ch < 2) {
                    float4 v = make_float4(rad.x, 0.0f, 0.0f, 0.0f);
                    if (ch == 1) {
                        v = samples[slot];
                        v.y = rad.y;
                    }
                    samples[slot] = v;
                    ++ch;
                    newPath = true;
                } else {
                    float sR = rad.x, sG = rad.y, sB = rad.z;
                    if (lca) {
                        const float4 v = samples[slot];
                        sR = v.x;
                        sG = v.y;
                    }
                    // sample complete (pathtrace.comp:633-637); lanes_reduce accumulates it
                    if (isnan(sR) || isnan(sG) || isnan(sB) || isinf(sR) || isinf(sG) || isinf(sB))
                        sR = sG = sB = 0.0f;
                    samples[slot] = make_float4(sR, sG, sB, 0.0f);
                    if (s == K - 1) A.primId[pi] = firstTri;
                    phase = PH_FETCH;
                }
            }
            if (newPath) {
                float fOff = 0.0f;
                if (lca) {
                    seed = init_seed(px, py, frame * 3 + ch);
                    fOff = cam.lca[ch];
                } else {
                    seed = init_seed(px, py, frame);
                }
                gen_ray(cam, px, py, seed, fOff, w.ro, w.rd);
                thr = mk(1.0f, 1.0f, 1.0f);
                rad = mk(0.0f, 0.0f, 0.0f);
                bounce = 0;
                lastDiffuse = false;
                nextKind = RAY_PATH;
                trace = true;
                if (STATS) C.paths++;
            }
            if (trace) {
                walk_begin(S, w, tmax);
                phase = PH_WALK;
                if (STATS) {
                    if (nextKind == RAY_SHADOW) {
                        C.shadowRays++;
                    } else {
                        C.rays++;
                        C.nodePops++;
                    }
                }
            }
            flags = f_pack(nextKind, ch, bounce, lastDiffuse, contEnd);
        }
        if (__ballot(phase != PH_DONE) == 0ull) break;

        // (3) walk until no lane walks or `thresh` lanes wait for shading / a unit.
        const bool anyHit = (flags & 3u) == (uint32_t)RAY_SHADOW;
        while (true) {
            if (phase == PH_WALK) {
                const bool done = WALK == 2   ? walk_step_dual<STATS, NS>(S, w, anyHit, st, spill, C)
                                  : WALK == 1 ? walk_step_tri<STATS, NS>(S, w, anyHit, st, spill, C)
                                              : walk_step<STATS, NS>(S, w, anyHit, st, spill, C);
                if (done) phase = PH_SHADE;
            }
            if (__ballot(phase == PH_WALK) == 0ull) break;
            if (__popcll(__ballot(phase == PH_SHADE || phase == PH_FETCH)) >= thresh) break;
        }
    }
}

// Kernel variants (rc_tuning.laneVariant): waves per SIMD W, LDS stack entries
// NS (W workgroups x NS x 2 KB of stacks within a CU's 160 KB) and whether a
// walk step tests a whole leaf (WALK 0), one triangle (1), or one triangle and
// one node (2, walk_step_dual):
//   0: W 7, NS 11            1: W 6, NS 13  2: W 8, NS 10  3: W 5, NS 16
//   4: W 7, NS 11, TRI       5: W 6, NS 13, TRI       6: W 5, NS 16, TRI
//   7: W 4, NS 20, TRI                                8: W 3, NS 26, TRI
//   9: W 4, NS 20, dual (default)                    10: W 5, NS 16, dual
// On the 250k-triangle frame at 80 samples: variant 0 177 ms, 4 141 ms, 6 138 ms,
// 7 121 ms (no register spills; 117 VGPRs since the camera is read through the
// constant address space; the 5-wave and higher variants spill around the shading
// block and move about 5x the HBM bytes in scratch traffic), 8 144 ms, 9 120 ms;
// at 5 samples variant 9 (118 VGPRs) takes 8.39-8.45 ms against 8.48-8.56 for 7.
template <bool STATS, int W, int NS, int WALK>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(W, 8)))
pathtrace_lanes_kernel(PassArgs A, float4 *samples, int *counter, int thresh, float4 *cont) {
    __shared__ int2 stk[NS][kBlock];
    const StackRef st{stk, (int)threadIdx.x};
    unsigned long long spill[64 - NS];
    Counters C;
    if (STATS) C = Counters{};
    lanes_view<STATS, NS, WALK>(A, A.frame0, A.nSamples, samples, counter, thresh, cont, st, spill, C);
    if (STATS) wave_add_counters(A.counters, C);
}

// Many views in one launch: every wave works through the views in table order
// (heaviest first), taking units of a view until its counter is exhausted. The
// table is read in place (uniform addresses: scalar loads), not copied.
template <bool STATS, int W, int NS, int WALK>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(W, 8)))
pathtrace_lanes_views_kernel(const PassArgs *__restrict__ views, const LaneChunk *__restrict__ chunks, int nViews,
                             int thresh, float4 *cont, unsigned long long *counters) {
    __shared__ int2 stk[NS][kBlock];
    const StackRef st{stk, (int)threadIdx.x};
    unsigned long long spill[64 - NS];
    Counters C;
    if (STATS) C = Counters{};
    for (int v = 0; v < nViews; ++v) {
        const LaneChunk &ck = chunks[v];
        if (ck.nSamples <= 0) continue;
        lanes_view<STATS, NS, WALK>(views[v], views[v].frame0 + ck.s0, ck.nSamples, ck.samples, ck.counter, thresh,
                                   cont, st, spill, C);
    }
    if (STATS) wave_add_counters(counters, C);
}

using LaneKernel = void (*)(PassArgs, float4 *, int *, int, float4 *);
using LaneViewsKernel = void (*)(const PassArgs *, const LaneChunk *, int, int, float4 *, unsigned long long *);

#define RC_LANE_VARIANT_TABLE(K, S)                                              \
    switch (variant) {                                                           \
        case 1: return S ? K<true, 6, 13, 0> : K<false, 6, 13, 0>;               \
        case 2: return S ? K<true, 8, 10, 0> : K<false, 8, 10, 0>;               \
        case 3: return S ? K<true, 5, 16, 0> : K<false, 5, 16, 0>;               \
        case 4: return S ? K<true, 7, 11, 1> : K<false, 7, 11, 1>;               \
        case 5: return S ? K<true, 6, 13, 1> : K<false, 6, 13, 1>;               \
        case 6: return S ? K<true, 5, 16, 1> : K<false, 5, 16, 1>;               \
        case 7: return S ? K<true, 4, 20, 1> : K<false, 4, 20, 1>;               \
        case 8: return S ? K<true, 3, 26, 1> : K<false, 3, 26, 1>;               \
        case 9: return S ? K<true, 4, 20, 2> : K<false, 4, 20, 2>;               \
        case 10: return S ? K<true, 5, 16, 2> : K<false, 5, 16, 2>;              \
        default: return S ? K<true, 7, 11, 0> : K<false, 7, 11, 0>;              \
    }

LaneKernel lane_kernel(int variant, bool stats) { RC_LANE_VARIANT_TABLE(pathtrace_lanes_kernel, stats) }

LaneViewsKernel lane_views_kernel(int variant, bool stats) {
    RC_LANE_VARIANT_TABLE(pathtrace_lanes_views_kernel, stats)
}

// Adds the K samples of every pixel to its running sums in frame order
// (pathtrace.comp:639-654): frame 0 stores, later frames add.
template <bool FP64>
__global__ void __launch_bounds__(256) lanes_reduce_kernel(ReduceArgs R) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= R.width * R.height) return;
    const int px = p % R.width, py = p / R.width;
    const size_t q = (size_t)LaneTile(R.width, R.height, R.tileShift).slot(px, py);
    float4 acc = make_float4(0, 0, 0, 0);
    double4 acc64 = make_double4(0, 0, 0, 0);
    float var = 0.0f;
    if (R.frame0 > 0) {
        if (FP64) acc64 = R.accum64[p];
        else acc = R.accum[p];
        var = R.variance[p];
    }
    const float4 *sp = R.samples + q;
    for (int s = 0; s < R.nSamples; ++s) {
        const float4 v = sp[(size_t)s * (size_t)R.nTilePx];
        const float luma = (v.x * 0.2126f + v.y * 0.7152f) + v.z * 0.0722f;
        const float l2 = luma * luma;
        if (R.frame0 + s == 0) {
            acc = make_float4(v.x, v.y, v.z, 1.0f);
            acc64 = make_double4(v.x, v.y, v.z, 1.0);
            var = l2;
        } else {
            if (FP64) {
                acc64.x += (double)v.x; acc64.y += (double)v.y; acc64.z += (double)v.z; acc64.w += 1.0;
            } else {
                acc.x += v.x; acc.y += v.y; acc.z += v.z; acc.w += 1.0f;
            }
            var += l2;
        }
    }
    if (FP64) {
        R.accum64[p] = acc64;
        acc = make_float4((float)acc64.x, (float)acc64.y, (float)acc64.z, (float)acc64.w);
    }
    R.accum[p] = acc;
    R.variance[p] = var;
}

// Views form: one thread per pixel of any view; view found by binary search
// over the pixel prefix sums.
template <bool FP64>
__global__ void __launch_bounds__(256) lanes_reduce_views_kernel(const PassArgs *views, const LaneChunk *chunks,
                                                                 const long long *pixOff, int nViews) {
    const long long g = (long long)blockIdx.x * 256 + threadIdx.x;
    if (g >= pixOff[nViews]) return;
    int lo = 0, hi = nViews;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pixOff[mid] <= g) lo = mid;
        else hi = mid;
    }
    const LaneChunk &ck = chunks[lo];
    if (ck.nSamples <= 0) return;
    const PassArgs &A = views[lo];
    const int p = (int)(g - pixOff[lo]);
    const int W = A.cam.width;
    const int px = p % W, py = p / W;
    const size_t q = (size_t)LaneTile(W, A.cam.height, A.tileShift).slot(px, py);
    const int frame0 = A.frame0 + ck.s0;
    float4 acc = make_float4(0, 0, 0, 0);
    double4 acc64 = make_double4(0, 0, 0, 0);
    float var = 0.0f;
    if (frame0 > 0) {
        if (FP64) acc64 = A.accum64[p];
        else acc = A.accum[p];
        var = A.variance[p];
    }
    const float4 *sp = ck.samples + q;
    for (int s = 0; s < ck.nSamples; ++s) {
        const float4 v = sp[(size_t)s * (size_t)ck.nTilePx];
        const float luma = (v.x * 0.2126f + v.y * 0.7152f) + v.z * 0.0722f;
        const float l2 = luma * luma;
        if (frame0 + s == 0) {
            acc = make_float4(v.x, v.y, v.z, 1.0f);
            acc64 = make_double4(v.x, v.y, v.z, 1.0);
            var = l2;
        } else {
            if (FP64) {
                acc64.x += (double)v.x; acc64.y += (double)v.y; acc64.z += (double)v.z; acc64.w += 1.0;
            } else {
                acc.x += v.x; acc.y += v.y; acc.z += v.z; acc.w += 1.0f;
            }
            var += l2;
        }
    }
    if (FP64) {
        A.accum64[p] = acc64;
        acc = make_float4((float)acc64.x, (float)acc64.y, (float)acc64.z, (float)acc64.w);
    }
    A.accum[p] = acc;
    A.variance[p] = var;
}

__global__ void fill_int(int *p, int n, int v) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

// Resident workgroups per CU x CUs of a kernel on the current device, queried
// once per (kernel, device): the query costs host time that a short launch
// would otherwise wait for.
int resident_blocks(const void *kernel) {
    static std::mutex mu;
    static std::map<std::pair<const void *, int>, int> cache;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find({kernel, dev});
    if (it != cache.end()) return it->second;
    int occ = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, kBlock, 0) != hipSuccess || occ < 1) occ = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
    cache[{kernel, dev}] = occ * cus;
    return occ * cus;
}

// Workgroups of a persistent launch: as many as are resident at once, but no
// more than the units need.
int persistent_blocks(const void *kernel, int units) {
    long long grid = resident_blocks(kernel);
    const long long need = ((long long)units + kBlock - 1) / kBlock;
    if (grid > need) grid = need;
    return grid < 1 ? 1 : (int)grid;
}

bool lanes_supported(const DeviceScene &D) { return D.valid && D.maxDepth < 30; }

int lanes_chunk_samples(int width, int height, int tileShift, int nSamples, long long capBytes) {
    const long long tilePx = LaneTile(width, height, tileShift).pixels();
    long long k = capBytes / (tilePx * 16);
    const long long kUnits = (1ll << 30) / tilePx - 1;  // K samples + the G-buffer rays
    if (k > kUnits) k = kUnits;
    if (k < 1) k = 1;
    return (int)(k < nSamples ? k : nSamples);
}

hipError_t ensure_lane_buffers(rc_ctx *ctx, size_t sampleBytes, int counters) {
    LaneBuffers &L = ctx->lanes;
    if (ctx->capturing && (sampleBytes > L.sampleBytes || counters > L.nCounters))
        return hipErrorStreamCaptureUnsupported;  // would allocate inside a captured frame
    if (sampleBytes > L.sampleBytes) {
        ++ctx->freeGen;
        if (L.samples) (void)hipFree(L.samples);
        L.samples = nullptr;
        L.sampleBytes = 0;
        hipError_t e = hipMalloc((void **)&L.samples, sampleBytes);
        if (e != hipSuccess) return e;
        L.sampleBytes = sampleBytes;
    }
    if (counters > L.nCounters) {
        ++ctx->freeGen;
        if (L.counters) (void)hipFree(L.counters);
        L.counters = nullptr;
        L.nCounters = 0;
        hipError_t e = hipMalloc((void **)&L.counters, sizeof(int) * (size_t)counters);
        if (e != hipSuccess) return e;
        L.nCounters = counters;
    }
    return hipSuccess;
}

void free_lane_buffers(LaneBuffers &L) {
    if (L.samples) (void)hipFree(L.samples);
    if (L.counters) (void)hipFree(L.counters);
    if (L.cont) (void)hipFree(L.cont);
    L = LaneBuffers{};
}

namespace {

// Continuation slots for every lane a launch of `kernel` can have resident.
hipError_t ensure_cont(rc_ctx *ctx, const void *kernel) {
    LaneBuffers &L = ctx->lanes;
    const size_t lanes = (size_t)resident_blocks(kernel) * kBlock;
    if (lanes <= L.contLanes) return hipSuccess;
    if (ctx->capturing) return hipErrorStreamCaptureUnsupported;  // would allocate inside a captured frame
    ++ctx->freeGen;
    if (L.cont) (void)hipFree(L.cont);
    L.cont = nullptr;
    L.contLanes = 0;
    hipError_t e = hipMalloc((void **)&L.cont, lanes * 3 * sizeof(float4));
    if (e == hipSuccess) L.contLanes = lanes;
    return e;
}

}  // namespace

hipError_t launch_lanes(rc_ctx *ctx, const PassArgs &A0, int thresh, int variant) {
    const int W = A0.cam.width, H = A0.cam.height;
    const int nTilePx = LaneTile(W, H, A0.tileShift).pixels();
    const int kMax = lanes_chunk_samples(W, H, A0.tileShift, A0.nSamples, (long long)ctx->laneBufferMiB << 20);
    hipError_t e = ensure_lane_buffers(ctx, (size_t)kMax * (size_t)nTilePx * 16, 1);
    if (e != hipSuccess) return e;
    const bool stats = ctx->countersEnabled && ctx->counters;
    const bool fp64 = ctx->accumMode == RC_ACCUM_FP64;
    const LaneKernel kernel = lane_kernel(variant, stats);
    e = ensure_cont(ctx, (const void *)kernel);
    if (e != hipSuccess) return e;
    for (int s0 = 0; s0 < A0.nSamples; s0 += kMax) {
        PassArgs A = A0;
        A.frame0 = A0.frame0 + s0;
        A.nSamples = A0.nSamples - s0 < kMax ? A0.nSamples - s0 : kMax;
        hipLaunchKernelGGL(fill_int, dim3(1), dim3(256), 0, ctx->stream, ctx->lanes.counters, 1, 0);
        const int grid = persistent_blocks((const void *)kernel, nTilePx * (A.nSamples + (A.frame0 == 0 ? 1 : 0)));
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(kBlock), 0, ctx->stream, A, ctx->lanes.samples,
                           ctx->lanes.counters, thresh, ctx->lanes.cont);
        ReduceArgs R;
        R.samples = ctx->lanes.samples;
        R.nTilePx = nTilePx;
        R.nSamples = A.nSamples;
        R.frame0 = A.frame0;
        R.width = W;
        R.height = H;
        R.tileShift = A.tileShift;
        R.accum = A.accum;
        R.accum64 = A.accum64;
        R.variance = A.variance;
        const dim3 rg((W * H + 255) / 256);
        if (fp64) hipLaunchKernelGGL(lanes_reduce_kernel<true>, rg, dim3(256), 0, ctx->stream, R);
        else hipLaunchKernelGGL(lanes_reduce_kernel<false>, rg, dim3(256), 0, ctx->stream, R);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_lanes_views(rc_ctx *ctx, const PassArgs *dViews, LaneChunk *dChunks, const long long *dPixOff,
                              int nViews, long long totalPx, long long totalUnits, int thresh, int variant,
                              bool fp64, int zeroCounters) {
    const bool stats = ctx->countersEnabled && ctx->counters;
    const LaneViewsKernel kernel = lane_views_kernel(variant, stats);
    hipError_t e = ensure_cont(ctx, (const void *)kernel);
    if (e != hipSuccess) return e;
    if (zeroCounters > 0)
        hipLaunchKernelGGL(fill_int, dim3((zeroCounters + 255) / 256), dim3(256), 0, ctx->stream, ctx->lanes.counters,
                           zeroCounters, 0);
    const int grid = persistent_blocks((const void *)kernel,
                                       totalUnits > (1ll << 30) ? (1 << 30) : (int)totalUnits);
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(kBlock), 0, ctx->stream, dViews, dChunks, nViews, thresh,
                       ctx->lanes.cont, ctx->counters);
    const dim3 rg((unsigned)((totalPx + 255) / 256));
    if (fp64) hipLaunchKernelGGL(lanes_reduce_views_kernel<true>, rg, dim3(256), 0, ctx->stream, dViews, dChunks,
                                 dPixOff, nViews);
    else hipLaunchKernelGGL(lanes_reduce_views_kernel<false>, rg, dim3(256), 0, ctx->stream, dViews, dChunks,
                            dPixOff, nViews);
    return hipGetLastError();
}

}  // namespace rc
