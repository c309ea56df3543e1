// rc_device.h — gfx950 device building blocks of the render hot path:
// vector helpers, PCG, ray/box and ray/triangle tests and the two BVH walks.
//
// Device data layout (built by rc_upload_scene from the Java float streams):
//   inner nodes  64 B  {L.min.xyz, L.max.x} {L.max.yz, R.min.xy} {R.min.z, R.max.xyz}
//                      {leftRef, rightRef, -, -}
//                a BVH2 node that holds both children's boxes, so the ordered
//                child test of pathtrace.comp:181-197 is one 64 B read.
//   child ref    >= 0: index of an inner node;  < 0: leaf, ~ref = start<<5 | count
//                (count 0 means "look up leafBigCount[start]", for leaves > 31 tris)
//   triangles    48 B  {v0, -} {v1 - v0, -} {v2 - v0, -}  (MT test operands)
//   shading      48 B  {n0, matIndex bits} {n1, -} {n2, -}
//   materials    64 B  Material.write layout;  lights 80 B + a packed cdf[] array
//
// Every arithmetic step follows the reference GLSL operand order with no FMA
// contraction (-ffp-contract=off) so that results equal the CPU oracle bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rc_fmath.h"

namespace rc {

struct v3 { float x, y, z; };

__device__ __forceinline__ v3 mk(float x, float y, float z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ v3 operator+(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ v3 operator-(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ v3 operator*(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ v3 operator*(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ v3 operator/(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ v3 cross(v3 a, v3 b) {
    return mk(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
__device__ __forceinline__ v3 normalize(v3 a) { float inv = 1.0f / sqrtf(dot(a, a)); return a * inv; }
__device__ __forceinline__ v3 reflect(v3 i, v3 n) { float d = 2.0f * dot(n, i); return i - n * d; }
__device__ __forceinline__ v3 refract(v3 i, v3 n, float eta) {
    float d = dot(n, i);
    float k = 1.0f - (eta * eta) * (1.0f - d * d);
    if (k < 0.0f) return mk(0.0f, 0.0f, 0.0f);
    float c = eta * d + sqrtf(k);
    return i * eta - n * c;
}
__device__ __forceinline__ v3 xyz(float4 f) { return mk(f.x, f.y, f.z); }

// PCG-RXS-M-XS (pathtrace.comp:80-98).
__device__ __forceinline__ uint32_t pcg(uint32_t &s) {
    s = s * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}
__device__ __forceinline__ float rnd(uint32_t &s) { return (float)pcg(s) / 4294967296.0f; }
__device__ __forceinline__ uint32_t init_seed(int x, int y, int frame) {
    uint32_t s = (uint32_t)x + (uint32_t)y * 7919u + (uint32_t)frame * 6271u;
    pcg(s);
    pcg(s);
    return s;
}

struct SceneView {
    const float *__restrict__ inner;     // 14 per inner node: left box, right box, left ref, right ref
    const float *__restrict__ triPos;    // 9 per triangle: v0, e1 = v1 - v0, e2 = v2 - v0 (36 B, packed)
    const float4 *__restrict__ triShade; // 3 per triangle
    const float4 *__restrict__ mats;     // 4 per material
    const float4 *__restrict__ lights;   // 5 per light
    const float *__restrict__ cdf;       // 1 per light
    const int *__restrict__ leafBigCount;
    float rootMin[3], rootMax[3];
    int rootRef;
    int nLights;
    float totalLightPower;
    int cdfSorted;  // cdf is non-decreasing and NaN-free: light selection may bisect
};

// Work counters (bench instrumentation); summed per wave into global memory.
struct Counters {
    uint32_t rays, shadowRays, nodePops, childTests, triTests, hitUpdates, shadeHits,
        shadowPops, shadowTriTests, lightScans, lightReads, paths;
};

template <bool STATS>
struct Ctr {
    Counters c;
    __device__ void zero() { c = Counters{}; }
};

// intersectAABB (pathtrace.comp:102-110).
__device__ __forceinline__ bool aabb(v3 ro, v3 inv, float mnx, float mny, float mnz,
                                     float mxx, float mxy, float mxz, float &tmin) {
    float t0x = (mnx - ro.x) * inv.x, t0y = (mny - ro.y) * inv.y, t0z = (mnz - ro.z) * inv.z;
    float t1x = (mxx - ro.x) * inv.x, t1y = (mxy - ro.y) * inv.y, t1z = (mxz - ro.z) * inv.z;
    float lx = fminf(t0x, t1x), ly = fminf(t0y, t1y), lz = fminf(t0z, t1z);
    float hx = fmaxf(t0x, t1x), hy = fmaxf(t0y, t1y), hz = fmaxf(t0z, t1z);
    float tn = fmaxf(fmaxf(lx, ly), lz);
    float tx = fminf(fminf(hx, hy), hz);
    tmin = tn;
    return tx >= fmaxf(tn, 0.0f);
}

// 1/x rounded to nearest, as the IEEE division computes it, for the x the
// triangle test divides by (|x| >= 1e-8): v_rcp_f32 and one FMA Newton step
// give the correctly rounded reciprocal for every finite |x| <= 2^126 (checked
// on the hardware for all such binary32 values, rc_debug_recip_check); above
// that the result is subnormal and the IEEE division runs.
__device__ __forceinline__ float recip_rn(float x) {
    if (__builtin_expect(fabsf(x) > 0x1p126f, 0)) return 1.0f / x;
    const float r = __builtin_amdgcn_rcpf(x);
    return __builtin_fmaf(__builtin_fmaf(-x, r, 1.0f), r, r);
}

// intersectTriangle (pathtrace.comp:112-130) on pre-subtracted edges.
__device__ __forceinline__ bool tri_test(const SceneView &S, v3 ro, v3 rd, int idx,
                                         float &t, float &bu, float &bv) {
    const float *p = S.triPos + 9 * (size_t)idx;
    v3 v0 = mk(p[0], p[1], p[2]), e1 = mk(p[3], p[4], p[5]), e2 = mk(p[6], p[7], p[8]);
    v3 s = ro - v0;
    // s is formed first and held by an empty asm, so that v0 is loaded together
    // with e1 and e2: with the test's early exits the compiler had moved v0's load
    // below the determinant test, a second memory round trip per tested triangle.
    asm volatile("" : "+v"(s.x), "+v"(s.y), "+v"(s.z));
    // Every value is computed for every lane and the reference's early exits are
    // one predicate at the end (the same comparisons, NaN-for-NaN): within a wave
    // some lane nearly always passes each test, so branching would only add
    // exec-mask work. The values of a lane that fails are never used.
    v3 h = cross(rd, e2);
    float det = dot(e1, h);
    float f = recip_rn(det);
    float u = f * dot(s, h);
    v3 q = cross(s, e1);
    float v = f * dot(rd, q);
    float tt = f * dot(e2, q);
    const bool ok = !(fabsf(det) < 1e-8f) && !(u < 0.0f || u > 1.0f) && !(v < 0.0f || u + v > 1.0f) &&
                    !(tt < 1e-4f);
    t = tt; bu = u; bv = v;
    return ok;
}

// Inner node record, 14 words (56 B, packed): the left child's box (min xyz,
// max xyz), the right child's, then the two child refs (>= 0: inner node index,
// < 0: ~(first triangle << 5 | count)).
struct NodeRec {
    float4 a, b, c;  // a.xyz left min, a.w b.xy left max, b.zw c.x right min, c.yzw right max
    int2 r;          // left ref, right ref
};
__device__ __forceinline__ NodeRec node_rec(const SceneView &S, int ref) {
    const float *p = S.inner + 14 * (size_t)ref;
    NodeRec n;
    n.a = make_float4(p[0], p[1], p[2], p[3]);
    n.b = make_float4(p[4], p[5], p[6], p[7]);
    n.c = make_float4(p[8], p[9], p[10], p[11]);
    n.r = make_int2(__float_as_int(p[12]), __float_as_int(p[13]));
    return n;
}

__device__ __forceinline__ void leaf_range(const SceneView &S, int ref, int &start, int &count) {
    uint32_t v = ~(uint32_t)ref;
    start = (int)(v >> 5);
    count = (int)(v & 31u);
    if (count == 0) count = S.leafBigCount[start];
}

struct Hit {
    float t, bu, bv;
    int tri;
};

// Traversal stacks: the first kShortStack entries of each lane live in LDS
// (row-major [entry][lane] so a wave's accesses hit 64 distinct banks); deeper
// entries spill to a private array. The logical stack depth and the reference's
// caps (62 closest-hit, 30 shadow) are unchanged by where entries live.
// 11 entries x 8 B x 256 lanes = 22.5 KB per workgroup: seven workgroups (7
// waves per SIMD) fit the 160 KB LDS of a CU (measured best on the full frame).
constexpr int kShortStack = 11;
constexpr int kBlock = 256;

struct StackRef {
    int2 (*lds)[kBlock];  // __shared__ int2[kShortStack][kBlock]
    int lane;             // threadIdx.x
};

// traverseBVH (pathtrace.comp:142-203). The stack holds (ref, tEntry): the child
// box test at push time is the same computation as the reference's re-test of the
// node's own box at pop, so the pop-time cull `aabbT > rec.t` reads tEntry.
// The 62-entry cap (pathtrace.comp:176) is applied to the same logical depth.
template <bool STATS>
__device__ __forceinline__ Hit closest_hit(const SceneView &S, v3 ro, v3 rd, StackRef st, Counters &C) {
    Hit h;
    h.t = 1e30f; h.tri = -1; h.bu = 0.0f; h.bv = 0.0f;
    v3 inv = mk(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    if (STATS) { C.rays++; C.nodePops++; }
    float tr;
    if (!aabb(ro, inv, S.rootMin[0], S.rootMin[1], S.rootMin[2], S.rootMax[0], S.rootMax[1],
              S.rootMax[2], tr))
        return h;
    int2 spill[64 - kShortStack];
    int top = 0;
    int ref = S.rootRef;
    float tEntry = tr;
    bool have = true;
    while (true) {
        if (!have) {
            if (top == 0) break;
            --top;
            int2 e = top < kShortStack ? st.lds[top][st.lane] : spill[top - kShortStack];
            ref = e.x;
            tEntry = __int_as_float(e.y);
            if (STATS) C.nodePops++;
        }
        have = false;
        if (tEntry > h.t) continue;
        if (ref < 0) {
            int start, count;
            leaf_range(S, ref, start, count);
            for (int i = start; i < start + count; ++i) {
                float t, bu, bv;
                if (STATS) C.triTests++;
                if (tri_test(S, ro, rd, i, t, bu, bv) && t < h.t) {
                    h.t = t; h.bu = bu; h.bv = bv; h.tri = i;
                    if (STATS) C.hitUpdates++;
                }
            }
        } else if (top < 62) {
            const NodeRec nd = node_rec(S, ref);
            const float4 a = nd.a, b = nd.b, c = nd.c;
            const int2 r = nd.r;
            if (STATS) C.childTests++;
            float tl, trr;
            bool hl = aabb(ro, inv, a.x, a.y, a.z, a.w, b.x, b.y, tl);
            bool hr = aabb(ro, inv, b.z, b.w, c.x, c.y, c.z, c.w, trr);
            if (hl || hr) {
                if (STATS) C.nodePops++;
                have = true;  // the entry the reference pushes last is popped next
                if (hl && hr) {
                    bool nearLeft = tl < trr;
                    int2 far = nearLeft ? make_int2(r.y, __float_as_int(trr)) : make_int2(r.x, __float_as_int(tl));
                    if (top < kShortStack) st.lds[top][st.lane] = far;
                    else spill[top - kShortStack] = far;
                    ++top;
                    ref = nearLeft ? r.x : r.y;
                    tEntry = nearLeft ? tl : trr;
                } else {
                    ref = hl ? r.x : r.y;
                    tEntry = hl ? tl : trr;
                }
            }
        }
    }
    return h;
}

// Box of stack entry e = (parent << 1 | side), or the root for e < 0; returns the
// child ref that box belongs to.
__device__ __forceinline__ int entry_box(const SceneView &S, int e, float &mnx, float &mny, float &mnz,
                                         float &mxx, float &mxy, float &mxz) {
    if (e < 0) {
        mnx = S.rootMin[0]; mny = S.rootMin[1]; mnz = S.rootMin[2];
        mxx = S.rootMax[0]; mxy = S.rootMax[1]; mxz = S.rootMax[2];
        return S.rootRef;
    }
    const float *n = S.inner + 14 * (size_t)(e >> 1) + 6 * (e & 1);
    mnx = n[0]; mny = n[1]; mnz = n[2]; mxx = n[3]; mxy = n[4]; mxz = n[5];
    return __float_as_int(S.inner[14 * (size_t)(e >> 1) + 12 + (e & 1)]);
}

// Closest hit of the photon shaders (photon_trace.comp:101-124,
// photon_gather.comp:85-118): unordered DFS, pop tests the node's own box and
// culls at aabbT > rec.t, pushes right then left, 62-entry cap.
__device__ __forceinline__ Hit closest_hit_unordered(const SceneView &S, v3 ro, v3 rd, StackRef st) {
    Hit h;
    h.t = 1e30f; h.tri = -1; h.bu = 0.0f; h.bv = 0.0f;
    v3 inv = mk(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    int spill[64 - kShortStack];
    int top = 0;
    int e = -1;
    bool have = true;
    while (true) {
        if (!have) {
            if (top == 0) break;
            --top;
            e = top < kShortStack ? st.lds[top][st.lane].x : spill[top - kShortStack];
        }
        have = false;
        float mnx, mny, mnz, mxx, mxy, mxz;
        int ref = entry_box(S, e, mnx, mny, mnz, mxx, mxy, mxz);
        float at;
        if (!aabb(ro, inv, mnx, mny, mnz, mxx, mxy, mxz, at)) continue;
        if (at > h.t) continue;
        if (ref < 0) {
            int start, count;
            leaf_range(S, ref, start, count);
            for (int i = start; i < start + count; ++i) {
                float t, bu, bv;
                if (tri_test(S, ro, rd, i, t, bu, bv) && t < h.t) {
                    h.t = t; h.bu = bu; h.bv = bv; h.tri = i;
                }
            }
        } else if (top < 62) {
            int right = (ref << 1) | 1;
            if (top < kShortStack) st.lds[top][st.lane].x = right;
            else spill[top - kShortStack] = right;
            ++top;
            e = ref << 1;
            have = true;
        }
    }
    return h;
}

// shadowRay (pathtrace.comp:211-242): unordered DFS with the 30-entry cap. Both
// children are pushed untested; the pop tests the node's own box. Boxes of inner
// nodes live in their parent, so each stack entry names (parent << 1 | side);
// -1 is the root.
template <bool STATS>
__device__ __forceinline__ bool unoccluded(const SceneView &S, v3 ro, v3 rd, float maxDist, StackRef st,
                                           Counters &C) {
    v3 inv = mk(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    if (STATS) C.shadowRays++;
    int spill[32 - kShortStack];
    int top = 0;
    int e = -1;
    bool have = true;
    while (true) {
        if (!have) {
            if (top == 0) break;
            --top;
            e = top < kShortStack ? st.lds[top][st.lane].x : spill[top - kShortStack];
        }
        have = false;
        if (STATS) C.shadowPops++;
        float mnx, mny, mnz, mxx, mxy, mxz;
        int ref;
        if (e < 0) {
            mnx = S.rootMin[0]; mny = S.rootMin[1]; mnz = S.rootMin[2];
            mxx = S.rootMax[0]; mxy = S.rootMax[1]; mxz = S.rootMax[2];
            ref = S.rootRef;
        } else {
            const float *n = S.inner + 14 * (size_t)(e >> 1) + 6 * (e & 1);
            mnx = n[0]; mny = n[1]; mnz = n[2]; mxx = n[3]; mxy = n[4]; mxz = n[5];
            ref = __float_as_int(S.inner[14 * (size_t)(e >> 1) + 12 + (e & 1)]);
        }
        float at;
        if (!aabb(ro, inv, mnx, mny, mnz, mxx, mxy, mxz, at)) continue;
        if (at > maxDist) continue;
        if (ref < 0) {
            int start, count;
            leaf_range(S, ref, start, count);
            for (int i = start; i < start + count; ++i) {
                float t, bu, bv;
                if (STATS) C.shadowTriTests++;
                if (tri_test(S, ro, rd, i, t, bu, bv) && t < maxDist) return false;
            }
        } else if (top < 30) {
            // push right, then left; the left entry is popped next
            int right = (ref << 1) | 1;
            if (top < kShortStack) st.lds[top][st.lane].x = right;
            else spill[top - kShortStack] = right;
            ++top;
            e = ref << 1;
            have = true;
        }
    }
    return true;
}

}  // namespace rc
