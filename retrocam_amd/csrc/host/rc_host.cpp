// rc_host.cpp — host-side scene assembly (tessellation, binned-SAH BVH, light CDF).
//
// Restates the Java host steps that feed the render hot path, with Java float
// semantics: binary32 arithmetic evaluated left to right with no contraction
// (built with -ffp-contract=off), double-precision Math.* results rounded to
// float where the Java code casts, and (int) truncation toward zero.
#include "rc_host.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

namespace {

constexpr float kFloatMax = 3.4028234663852886e38f;

// Java Math.min/Math.max for float: NaN-propagating, -0.0 < +0.0.
inline float jmin(float a, float b) {
    if (a != a) return a;
    if (b != b) return b;
    if (a == 0.0f && b == 0.0f) return std::signbit(a) ? a : b;
    return a <= b ? a : b;
}
inline float jmax(float a, float b) {
    if (a != a) return a;
    if (b != b) return b;
    if (a == 0.0f && b == 0.0f) return std::signbit(a) ? b : a;
    return a >= b ? a : b;
}
inline float jsqrtf(float x) { return (float)std::sqrt((double)x); }

struct Tri {  // GpuTriangle fields (GpuTriangle.java:16-31)
    float v0[3], v1[3], v2[3];
    float n0[3], n1[3], n2[3];
    float uv0[2], uv1[2], uv2[2];
    int32_t mat;

    float mn(int a) const { return jmin(v0[a], jmin(v1[a], v2[a])); }
    float mx(int a) const { return jmax(v0[a], jmax(v1[a], v2[a])); }
    float centroid(int a) const { return (v0[a] + v1[a] + v2[a]) / 3.0f; }

    // GpuTriangle.withFlatNormal (GpuTriangle.java:38-50)
    void flat_normal() {
        float ax = v1[0] - v0[0], ay = v1[1] - v0[1], az = v1[2] - v0[2];
        float bx = v2[0] - v0[0], by = v2[1] - v0[1], bz = v2[2] - v0[2];
        float nx = ay * bz - az * by;
        float ny = az * bx - ax * bz;
        float nz = ax * by - ay * bx;
        float len = jsqrtf(nx * nx + ny * ny + nz * nz);
        if (len > 1e-8f) { nx /= len; ny /= len; nz /= len; }
        for (float *n : {n0, n1, n2}) { n[0] = nx; n[1] = ny; n[2] = nz; }
    }

    void write(float *o) const {  // GpuTriangle.write (GpuTriangle.java:68-79)
        std::memset(o, 0, 32 * sizeof(float));
        for (int a = 0; a < 3; ++a) {
            o[0 + a] = v0[a]; o[4 + a] = v1[a]; o[8 + a] = v2[a];
            o[12 + a] = n0[a]; o[16 + a] = n1[a]; o[20 + a] = n2[a];
        }
        o[24] = uv0[0]; o[25] = uv0[1]; o[26] = uv1[0]; o[27] = uv1[1];
        o[28] = uv2[0]; o[29] = uv2[1];
        std::memcpy(&o[30], &mat, 4);
    }
};

struct Node {  // BVHNode fields (BVHNode.java:19-26)
    float mn[3], mx[3];
    int32_t left = -1, right = -1, primStart = 0, primCount = 0;

    float sa() const {
        float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
    void write(float *o) const {
        o[0] = mn[0]; o[1] = mn[1]; o[2] = mn[2]; o[3] = 0.0f;
        o[4] = mx[0]; o[5] = mx[1]; o[6] = mx[2]; o[7] = 0.0f;
        std::memcpy(&o[8], &left, 4); std::memcpy(&o[9], &right, 4);
        std::memcpy(&o[10], &primStart, 4); std::memcpy(&o[11], &primCount, 4);
    }
};

struct Bucket {  // BVHBuilder.Bucket (BVHBuilder.java:182-202)
    int count = 0;
    float mn[3] = {kFloatMax, kFloatMax, kFloatMax};
    float mx[3] = {-kFloatMax, -kFloatMax, -kFloatMax};
    void expand(const Tri &t) {
        for (int a = 0; a < 3; ++a) { mn[a] = jmin(mn[a], t.mn(a)); mx[a] = jmax(mx[a], t.mx(a)); }
    }
    void merge(const Bucket &o) {
        count += o.count;
        if (o.count == 0) return;
        for (int a = 0; a < 3; ++a) { mn[a] = jmin(mn[a], o.mn[a]); mx[a] = jmax(mx[a], o.mx[a]); }
    }
    float sa() const {
        if (count == 0) return 0.0f;
        float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

constexpr int kBuckets = 8;
constexpr int kLeafMax = 4;
constexpr float kCTrav = 1.0f;
constexpr float kCIsect = 1.5f;

// Java's (int) cast of a float: NaN -> 0, saturating at the int range, else
// truncation toward zero.
inline int64_t java_int(float f) {
    if (f != f) return 0;
    if (f >= 2147483647.0f) return 2147483647;
    if (f <= -2147483648.0f) return -2147483647 - 1;
    return (int64_t)f;
}

struct Builder {  // BVHBuilder (BVHBuilder.java:18-164)
    std::vector<Tri> &tris;
    std::vector<Node> nodes;
    // A centroid that rounds below its node's minimum by at least an eighth of
    // a thin (<= 8 ulp) extent gives a bucket index <= -1: BVHBuilder.java:117-119
    // indexes buckets[-1] and throws, so the build fails here as well.
    bool failed = false;
    explicit Builder(std::vector<Tri> &t) : tris(t) {}

    Node from_tris(int start, int count) const {
        Node n;
        for (int a = 0; a < 3; ++a) { n.mn[a] = kFloatMax; n.mx[a] = -kFloatMax; }
        for (int i = start; i < start + count; ++i)
            for (int a = 0; a < 3; ++a) {
                n.mn[a] = jmin(n.mn[a], tris[i].mn(a));
                n.mx[a] = jmax(n.mx[a], tris[i].mx(a));
            }
        return n;
    }

    bool best_split(const Node &p, int start, int count, int &bAxis, float &bPos, float &bCost) {
        float parentSA = p.sa();
        if (parentSA < 1e-8f) return false;
        bool found = false;
        for (int axis = 0; axis < 3; ++axis) {
            float lo = p.mn[axis], hi = p.mx[axis];
            if (hi - lo < 1e-6f) continue;
            Bucket b[kBuckets];
            for (int i = start; i < start + count; ++i) {
                float c = tris[i].centroid(axis);
                float f = ((float)kBuckets * (c - lo)) / (hi - lo);
                int64_t k = java_int(f);
                if (k > kBuckets - 1) k = kBuckets - 1;
                if (k < 0) { failed = true; return false; }
                b[k].count++;
                b[k].expand(tris[i]);
            }
            for (int s = 0; s < kBuckets - 1; ++s) {
                Bucket L, R;
                for (int i = 0; i <= s; ++i) L.merge(b[i]);
                for (int i = s + 1; i < kBuckets; ++i) R.merge(b[i]);
                if (L.count == 0 || R.count == 0) continue;
                float cost = kCTrav + ((L.sa() * (float)L.count) * kCIsect) / parentSA
                                    + ((R.sa() * (float)R.count) * kCIsect) / parentSA;
                float pos = lo + ((hi - lo) * (float)(s + 1)) / (float)kBuckets;
                if (!found || cost < bCost) { found = true; bAxis = axis; bPos = pos; bCost = cost; }
            }
        }
        return found;
    }

    int partition(int start, int count, int axis, float pos) {
        int lo = start, hi = start + count - 1;
        while (lo <= hi) {
            if (tris[lo].centroid(axis) < pos) {
                ++lo;
            } else {
                Tri tmp = tris[lo]; tris[lo] = tris[hi]; tris[hi] = tmp;
                --hi;
            }
        }
        return lo;
    }

    // BVHBuilder.subdivide (BVHBuilder.java:61-95) as a depth-first walk over an
    // explicit stack: nodes are numbered in the same pre-order (a node, its left
    // subtree, its right subtree) and ranges are partitioned in the same order.
    void build_all(int count) {
        struct Task { int start, count, parent, right; };
        std::vector<Task> todo;
        todo.push_back({0, count, -1, 0});
        while (!todo.empty() && !failed) {
            Task t = todo.back();
            todo.pop_back();
            int idx = (int)nodes.size();
            nodes.push_back(from_tris(t.start, t.count));
            if (t.parent >= 0) (t.right ? nodes[t.parent].right : nodes[t.parent].left) = idx;
            int mid = split(idx, t.start, t.count);
            if (mid < 0) {
                nodes[idx].primStart = t.start;
                nodes[idx].primCount = t.count;
                continue;
            }
            todo.push_back({mid, t.start + t.count - mid, idx, 1});
            todo.push_back({t.start, mid - t.start, idx, 0});
        }
    }

    // The partition point of node idx over [start, start+count), or -1 for a leaf.
    int split(int idx, int start, int count) {
        if (count <= kLeafMax) return -1;
        int axis = 0; float pos = 0.0f, cost = 0.0f;
        bool ok = best_split(nodes[idx], start, count, axis, pos, cost);
        float leafCost = kCIsect * (float)count;
        if (failed || !ok || cost >= leafCost) return -1;
        int mid = partition(start, count, axis, pos);
        if (mid == start || mid == start + count) return -1;
        return mid;
    }
};

}  // namespace

struct rcs_scene {
    std::vector<float> mats;  // 16 floats each
    std::vector<Tri> tris;
    std::vector<Node> nodes;
    std::vector<float> lights;  // 20 floats each
    float totalPower = 0.0f;
    int nLights = 0;
    int maxDepth = 0, maxLeaf = 0, nLeaves = 0;
    bool built = false;
};

extern "C" {

rcs_scene *rcs_new(void) { return new (std::nothrow) rcs_scene(); }
void rcs_free(rcs_scene *s) { delete s; }

int32_t rcs_add_material(rcs_scene *s, const float *m) {
    if (!s || !m) return -1;
    s->mats.insert(s->mats.end(), m, m + 16);
    s->built = false;
    return (int32_t)(s->mats.size() / 16 - 1);
}

static void add_quad(rcs_scene *s, const float *a, const float *b, const float *c, const float *d, int mat) {
    Tri t1{}, t2{};
    for (int k = 0; k < 3; ++k) { t1.v0[k] = a[k]; t1.v1[k] = b[k]; t1.v2[k] = c[k]; }
    t1.uv0[0] = 0; t1.uv0[1] = 0; t1.uv1[0] = 0; t1.uv1[1] = 1; t1.uv2[0] = 1; t1.uv2[1] = 1;
    t1.mat = mat; t1.flat_normal();
    for (int k = 0; k < 3; ++k) { t2.v0[k] = a[k]; t2.v1[k] = c[k]; t2.v2[k] = d[k]; }
    t2.uv0[0] = 0; t2.uv0[1] = 0; t2.uv1[0] = 1; t2.uv1[1] = 1; t2.uv2[0] = 1; t2.uv2[1] = 0;
    t2.mat = mat; t2.flat_normal();
    s->tris.push_back(t1);
    s->tris.push_back(t2);
}

int32_t rcs_add_quad(rcs_scene *s, const float *p, int32_t mat) {
    if (!s || !p) return -1;
    add_quad(s, p, p + 3, p + 6, p + 9, mat);
    s->built = false;
    return 2;
}

int32_t rcs_add_box(rcs_scene *s, float cx, float cy, float cz, float hx, float hy, float hz, int32_t mat) {
    if (!s) return -1;
    float x0 = cx - hx, x1 = cx + hx, y0 = cy - hy, y1 = cy + hy, z0 = cz - hz, z1 = cz + hz;
    // Face order and winding of Box.toTriangles (Box.java:32-43).
    const float q[6][12] = {
        {x1, y0, z1, x1, y1, z1, x1, y1, z0, x1, y0, z0},
        {x0, y0, z0, x0, y1, z0, x0, y1, z1, x0, y0, z1},
        {x0, y1, z1, x0, y1, z0, x1, y1, z0, x1, y1, z1},
        {x0, y0, z0, x0, y0, z1, x1, y0, z1, x1, y0, z0},
        {x0, y0, z1, x0, y1, z1, x1, y1, z1, x1, y0, z1},
        {x1, y0, z0, x1, y1, z0, x0, y1, z0, x0, y0, z0},
    };
    for (int f = 0; f < 6; ++f) add_quad(s, q[f], q[f] + 3, q[f] + 6, q[f] + 9, mat);
    s->built = false;
    return 12;
}

int32_t rcs_add_sphere(rcs_scene *s, float cx, float cy, float cz, float radius, int32_t mat,
                       int32_t stacks, int32_t slices) {
    if (!s || stacks < 2 || slices < 3) return -1;
    auto pt = [&](float phi, float theta, float *o) {
        float sinP = (float)std::sin((double)phi);
        o[0] = cx + (radius * sinP) * (float)std::cos((double)theta);
        o[1] = cy + radius * (float)std::cos((double)phi);
        o[2] = cz + (radius * sinP) * (float)std::sin((double)theta);
    };
    auto nrm = [&](const float *p, float *n) {
        float nx = p[0] - cx, ny = p[1] - cy, nz = p[2] - cz;
        float len = jsqrtf(nx * nx + ny * ny + nz * nz);
        if (len > 1e-8f) { nx /= len; ny /= len; nz /= len; }
        n[0] = nx; n[1] = ny; n[2] = nz;
    };
    auto mk = [&](const float *p0, const float *p1, const float *p2,
                  const float *t0, const float *t1, const float *t2) {
        Tri t{};
        for (int k = 0; k < 3; ++k) { t.v0[k] = p0[k]; t.v1[k] = p1[k]; t.v2[k] = p2[k]; }
        nrm(p0, t.n0); nrm(p1, t.n1); nrm(p2, t.n2);
        t.uv0[0] = t0[0]; t.uv0[1] = t0[1]; t.uv1[0] = t1[0]; t.uv1[1] = t1[1];
        t.uv2[0] = t2[0]; t.uv2[1] = t2[1];
        t.mat = mat;
        s->tris.push_back(t);
    };
    int added = 0;
    const double PI = 3.141592653589793;
    for (int i = 0; i < stacks; ++i) {
        float phi0 = (float)(PI * i / stacks);
        float phi1 = (float)(PI * (i + 1) / stacks);
        for (int j = 0; j < slices; ++j) {
            float th0 = (float)(2 * PI * j / slices);
            float th1 = (float)(2 * PI * (j + 1) / slices);
            float A[3], B[3], C[3], D[3];
            pt(phi0, th0, A); pt(phi0, th1, B); pt(phi1, th1, C); pt(phi1, th0, D);
            float uvA[2] = {(float)j / slices, (float)i / stacks};
            float uvB[2] = {(float)(j + 1) / slices, (float)i / stacks};
            float uvC[2] = {(float)(j + 1) / slices, (float)(i + 1) / stacks};
            float uvD[2] = {(float)j / slices, (float)(i + 1) / stacks};
            if (i != 0) { mk(A, B, C, uvA, uvB, uvC); ++added; }
            if (i != stacks - 1) { mk(A, C, D, uvA, uvC, uvD); ++added; }
        }
    }
    s->built = false;
    return added;
}

int32_t rcs_rotate_since(rcs_scene *s, int32_t first, float px, float py, float pz,
                         float degX, float degY, float degZ) {
    if (!s || first < 0 || first > (int32_t)s->tris.size()) return -1;
    const double D2R = 3.141592653589793 / 180.0;
    float cx = (float)std::cos(degX * D2R), sx = (float)std::sin(degX * D2R);
    float cy = (float)std::cos(degY * D2R), sy = (float)std::sin(degY * D2R);
    float cz = (float)std::cos(degZ * D2R), sz = (float)std::sin(degZ * D2R);
    const float r[9] = {cy * cz, cy * sz, -sy,
                        sx * sy * cz - cx * sz, sx * sy * sz + cx * cz, sx * cy,
                        cx * sy * cz + sx * sz, cx * sy * sz - sx * cz, cx * cy};
    auto rot = [&](float x, float y, float z, float *o) {
        o[0] = r[0] * x + r[3] * y + r[6] * z;
        o[1] = r[1] * x + r[4] * y + r[7] * z;
        o[2] = r[2] * x + r[5] * y + r[8] * z;
    };
    for (size_t i = (size_t)first; i < s->tris.size(); ++i) {
        Tri &t = s->tris[i];
        for (float *v : {t.v0, t.v1, t.v2}) {
            float o[3];
            rot(v[0] - px, v[1] - py, v[2] - pz, o);
            v[0] = o[0] + px; v[1] = o[1] + py; v[2] = o[2] + pz;
        }
        for (float *n : {t.n0, t.n1, t.n2}) {
            float o[3];
            rot(n[0], n[1], n[2], o);
            n[0] = o[0]; n[1] = o[1]; n[2] = o[2];
        }
    }
    s->built = false;
    return 0;
}

int32_t rcs_add_triangles(rcs_scene *s, const float *t32, int32_t n) {
    if (!s || (!t32 && n > 0) || n < 0) return -1;
    for (int i = 0; i < n; ++i) {
        const float *o = t32 + 32 * (size_t)i;
        Tri t{};
        for (int a = 0; a < 3; ++a) {
            t.v0[a] = o[a]; t.v1[a] = o[4 + a]; t.v2[a] = o[8 + a];
            t.n0[a] = o[12 + a]; t.n1[a] = o[16 + a]; t.n2[a] = o[20 + a];
        }
        t.uv0[0] = o[24]; t.uv0[1] = o[25]; t.uv1[0] = o[26]; t.uv1[1] = o[27];
        t.uv2[0] = o[28]; t.uv2[1] = o[29];
        std::memcpy(&t.mat, &o[30], 4);
        s->tris.push_back(t);
    }
    s->built = false;
    return 0;
}

int32_t rcs_triangle_count(const rcs_scene *s) { return s ? (int32_t)s->tris.size() : -1; }

static void depth_walk(const std::vector<Node> &nodes, rcs_scene *s) {
    // Iterative walk: (node, depth).
    std::vector<std::pair<int, int>> st;
    st.push_back({0, 0});
    s->maxDepth = 0; s->maxLeaf = 0; s->nLeaves = 0;
    while (!st.empty()) {
        auto [n, d] = st.back();
        st.pop_back();
        if (d > s->maxDepth) s->maxDepth = d;
        const Node &nd = nodes[n];
        if (nd.primCount > 0) {
            s->nLeaves++;
            if (nd.primCount > s->maxLeaf) s->maxLeaf = nd.primCount;
        } else {
            st.push_back({nd.left, d + 1});
            st.push_back({nd.right, d + 1});
        }
    }
}

int32_t rcs_build(rcs_scene *s) {
    if (!s || s->tris.empty()) return -1;
    int nMats = (int)(s->mats.size() / 16);
    for (const Tri &t : s->tris)
        if (t.mat < 0 || t.mat >= nMats) return -1;
    // BVHBuilder.build reorders a copy of the list (BVHBuilder.java:35-37): a
    // failed build leaves the scene's triangles as they were.
    std::vector<Tri> order = s->tris;
    Builder b(order);
    b.build_all((int)order.size());
    if (b.failed) return -1;
    s->tris.swap(order);
    s->nodes.swap(b.nodes);
    depth_walk(s->nodes, s);

    // SceneUploader.uploadLights (SceneUploader.java:144-216), triangles in BVH order.
    s->lights.clear();
    float sumPower = 0.0f;
    const double PI = 3.141592653589793;
    for (const Tri &t : s->tris) {
        const float *m = &s->mats[16 * (size_t)t.mat];
        float strength = m[7];
        if (strength <= 0.0f) continue;
        float e1x = t.v1[0] - t.v0[0], e1y = t.v1[1] - t.v0[1], e1z = t.v1[2] - t.v0[2];
        float e2x = t.v2[0] - t.v0[0], e2y = t.v2[1] - t.v0[1], e2z = t.v2[2] - t.v0[2];
        float cx = e1y * e2z - e1z * e2y;
        float cy = e1z * e2x - e1x * e2z;
        float cz = e1x * e2y - e1y * e2x;
        float area = 0.5f * jsqrtf(cx * cx + cy * cy + cz * cz);
        if (area < 1e-8f) continue;
        float er = m[4] * strength, eg = m[5] * strength, eb = m[6] * strength;
        float lum = 0.299f * er + 0.587f * eg + 0.114f * eb;
        float power = (lum * area) * (float)PI;
        sumPower += power;
        float e[20] = {t.v0[0], t.v0[1], t.v0[2], area,
                       t.v1[0], t.v1[1], t.v1[2], 0.0f,
                       t.v2[0], t.v2[1], t.v2[2], 0.0f,
                       er, eg, eb, power,
                       t.n0[0], t.n0[1], t.n0[2], 0.0f};
        s->lights.insert(s->lights.end(), e, e + 20);
    }
    s->nLights = (int)(s->lights.size() / 20);
    s->totalPower = sumPower;
    if (s->nLights > 0) {
        float cum = 0.0f;
        for (int i = 0; i < s->nLights; ++i) {
            cum += s->lights[20 * (size_t)i + 15];
            s->lights[20 * (size_t)i + 19] = (sumPower > 0.0f) ? cum / sumPower : 1.0f;
        }
        s->lights[20 * (size_t)(s->nLights - 1) + 19] = 1.0f;
    }
    s->built = true;
    return 0;
}

int32_t rcs_get_info(const rcs_scene *s, rcs_info *o) {
    if (!s || !o || !s->built) return -1;
    o->nTris = (int32_t)s->tris.size();
    o->nNodes = (int32_t)s->nodes.size();
    o->nMats = (int32_t)(s->mats.size() / 16);
    o->nLights = s->nLights;
    o->totalLightPower = s->totalPower;
    o->maxDepth = s->maxDepth;
    o->maxLeafSize = s->maxLeaf;
    o->nLeaves = s->nLeaves;
    return 0;
}

int32_t rcs_copy_tris(const rcs_scene *s, float *dst) {
    if (!s || !dst || !s->built) return -1;
    for (size_t i = 0; i < s->tris.size(); ++i) s->tris[i].write(dst + 32 * i);
    return 0;
}
int32_t rcs_copy_nodes(const rcs_scene *s, float *dst) {
    if (!s || !dst || !s->built) return -1;
    for (size_t i = 0; i < s->nodes.size(); ++i) s->nodes[i].write(dst + 12 * i);
    return 0;
}
int32_t rcs_copy_mats(const rcs_scene *s, float *dst) {
    if (!s || !dst || !s->built) return -1;
    std::memcpy(dst, s->mats.data(), s->mats.size() * sizeof(float));
    return 0;
}
int32_t rcs_copy_lights(const rcs_scene *s, float *dst) {
    if (!s || !dst || !s->built) return -1;
    if (!s->lights.empty()) std::memcpy(dst, s->lights.data(), s->lights.size() * sizeof(float));
    return 0;
}

void rcs_orbit_basis(float theta, float phi, float radius, float tx, float ty, float tz,
                     float *eye, float *fwd, float *right, float *up) {
    // OrbitCamera.getEyePosition (OrbitCamera.java:155-162)
    float sinPhi = (float)std::sin((double)phi);
    eye[0] = tx + (radius * sinPhi) * (float)std::cos((double)theta);
    eye[1] = ty + radius * (float)std::cos((double)phi);
    eye[2] = tz + (radius * sinPhi) * (float)std::sin((double)theta);
    const float tgt[3] = {tx, ty, tz};
    rcs_lookat_basis(eye, tgt, fwd, right, up);
}

void rcs_lookat_basis(const float *eye, const float *tgt, float *fwd, float *right, float *up) {
    // getCameraForward / cameraRight / cameraUp (OrbitCamera.java:167-172, 274-291)
    float fx = tgt[0] - eye[0], fy = tgt[1] - eye[1], fz = tgt[2] - eye[2];
    float fl = jsqrtf(fx * fx + fy * fy + fz * fz);
    fx /= fl; fy /= fl; fz /= fl;
    fwd[0] = fx; fwd[1] = fy; fwd[2] = fz;
    float rx = fy * 0 - fz * 1, ry = fz * 0 - fx * 0, rz = fx * 1 - fy * 0;
    float rl = jsqrtf(rx * rx + ry * ry + rz * rz);
    right[0] = rx / rl; right[1] = ry / rl; right[2] = rz / rl;
    const float *r = right;
    up[0] = r[1] * fz - r[2] * fy;
    up[1] = r[2] * fx - r[0] * fz;
    up[2] = r[0] * fy - r[1] * fx;
}

float rcs_tan_half_fov(float focalLengthMm) {
    float focalM = focalLengthMm / 1000.0f;
    return (float)std::tan(std::atan((double)(0.012f / focalM)));
}

}  // extern "C"
