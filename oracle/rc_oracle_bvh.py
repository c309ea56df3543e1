"""CPU oracle of the scene's BVH and light list: the binned-SAH builder of
BVHBuilder (src/main/java/com/retrocam/scene/BVHBuilder.java:35-202), the node
bounds of BVHNode (src/main/java/com/retrocam/scene/BVHNode.java:44-61), the
triangle bounds and centroids of GpuTriangle
(src/main/java/com/retrocam/scene/GpuTriangle.java:53-65), and the light records
and CDF of SceneUploader.uploadLights
(src/main/java/com/retrocam/scene/SceneUploader.java:144-216).

TEST INFRASTRUCTURE ONLY: used by tests/ to check librc_host.so's rcs_build
record for record. Nothing in the product imports it.

Java semantics, one operation at a time: float arithmetic is binary32 (every
intermediate is a numpy.float32), an int operand of a float product is
converted first, (int) of a float is NaN -> 0, saturating, else truncation
toward zero, Math.min / Math.max return NaN when an operand is NaN and order
-0.0 below +0.0, (float) Math.sqrt of a float is the correctly rounded binary32
square root, and an array index below 0 throws (here BucketIndexError).
"""
import numpy as np

F = np.float32
BUCKETS = 8
LEAF_MAX = 4
C_TRAV = F(1.0)
C_ISECT = F(1.5)
FMAX = F(np.finfo(np.float32).max)


class BucketIndexError(IndexError):
    """BVHBuilder.java:117-119 indexes buckets[b] with b < 0."""


def jmin(a, b):
    if a != a:
        return a
    if b != b:
        return b
    if a == 0 and b == 0:
        return a if np.signbit(a) else b
    return a if a <= b else b


def jmax(a, b):
    if a != a:
        return a
    if b != b:
        return b
    if a == 0 and b == 0:
        return b if np.signbit(a) else a
    return a if a >= b else b


def java_int(f):
    if f != f:
        return 0
    if f >= F(2147483647.0):
        return 2147483647
    if f <= F(-2147483648.0):
        return -2147483648
    return int(f)


def _surface_area(mn, mx):
    """BVHNode.surfaceArea / Bucket.sa: 2f * (dx*dy + dy*dz + dz*dx)."""
    dx = mx[0] - mn[0]
    dy = mx[1] - mn[1]
    dz = mx[2] - mn[2]
    return F(2.0) * (dx * dy + dy * dz + dz * dx)


class _Tri:
    __slots__ = ("index", "mn", "mx", "c")

    def __init__(self, index, v):
        self.index = index
        v = [[F(x) for x in row] for row in v]
        self.mn = [jmin(v[0][a], jmin(v[1][a], v[2][a])) for a in range(3)]
        self.mx = [jmax(v[0][a], jmax(v[1][a], v[2][a])) for a in range(3)]
        self.c = [(v[0][a] + v[1][a] + v[2][a]) / F(3.0) for a in range(3)]


class _Box:
    """A bucket (count + bounds), BVHBuilder.Bucket."""

    def __init__(self):
        self.count = 0
        self.mn = [FMAX] * 3
        self.mx = [-FMAX] * 3

    def expand(self, t):
        self.mn = [jmin(self.mn[a], t.mn[a]) for a in range(3)]
        self.mx = [jmax(self.mx[a], t.mx[a]) for a in range(3)]

    def merge(self, o):
        self.count += o.count
        if o.count == 0:
            return
        self.mn = [jmin(self.mn[a], o.mn[a]) for a in range(3)]
        self.mx = [jmax(self.mx[a], o.mx[a]) for a in range(3)]

    def sa(self):
        return F(0.0) if self.count == 0 else _surface_area(self.mn, self.mx)


def build(verts):
    """Builds the BVH of triangles `verts` ((n, 3, 3): triangle, vertex, axis).

    Returns (nodes, order): nodes as a list of dicts with mn, mx (3 float32
    each), left, right, primStart, primCount in depth-first pre-order; order[i]
    = the input index of the triangle at position i after the in-place
    partitions. Raises BucketIndexError where the Java builder throws."""
    tris = [_Tri(i, verts[i]) for i in range(len(verts))]
    nodes = []

    def node_box(start, count):
        b = _Box()
        for i in range(start, start + count):
            b.expand(tris[i])
        return b

    def best_split(box, start, count):
        parent_sa = _surface_area(box.mn, box.mx)
        if parent_sa < F(1e-8):
            return None
        best = None
        for axis in range(3):
            lo, hi = box.mn[axis], box.mx[axis]
            if hi - lo < F(1e-6):
                continue
            buckets = [_Box() for _ in range(BUCKETS)]
            for i in range(start, start + count):
                c = tris[i].c[axis]
                b = min(BUCKETS - 1, java_int(F(BUCKETS) * (c - lo) / (hi - lo)))
                if b < 0:
                    raise BucketIndexError(f"bucket {b} on axis {axis}")
                buckets[b].count += 1
                buckets[b].expand(tris[i])
            for s in range(BUCKETS - 1):
                left, right = _Box(), _Box()
                for i in range(s + 1):
                    left.merge(buckets[i])
                for i in range(s + 1, BUCKETS):
                    right.merge(buckets[i])
                if left.count == 0 or right.count == 0:
                    continue
                cost = (C_TRAV + left.sa() * F(left.count) * C_ISECT / parent_sa
                        + right.sa() * F(right.count) * C_ISECT / parent_sa)
                split_pos = lo + (hi - lo) * F(s + 1) / F(BUCKETS)
                if best is None or cost < best[2]:
                    best = (axis, split_pos, cost)
        return best

    def partition(start, count, axis, split_pos):
        lo, hi = start, start + count - 1
        while lo <= hi:
            if tris[lo].c[axis] < split_pos:
                lo += 1
            else:
                tris[lo], tris[hi] = tris[hi], tris[lo]
                hi -= 1
        return lo

    def subdivide(start, count):
        idx = len(nodes)
        box = node_box(start, count)
        node = {"mn": box.mn, "mx": box.mx, "left": -1, "right": -1, "primStart": 0, "primCount": 0}
        nodes.append(node)

        def leaf():
            node["primStart"], node["primCount"] = start, count
            return idx

        if count <= LEAF_MAX:
            return leaf()
        best = best_split(box, start, count)
        if best is None or best[2] >= C_ISECT * F(count):
            return leaf()
        mid = partition(start, count, best[0], best[1])
        if mid == start or mid == start + count:
            return leaf()
        node["left"] = subdivide(start, mid - start)
        node["right"] = subdivide(mid, start + count - mid)
        return idx

    subdivide(0, len(tris))
    return nodes, [t.index for t in tris]


def nodes_to_records(nodes):
    """The 12-float wire records of the nodes (BVHNode.write, BVHNode.java:31-41)."""
    out = np.zeros((len(nodes), 12), np.float32)
    ints = out.view(np.int32)
    for i, n in enumerate(nodes):
        out[i, 0:3] = n["mn"]
        out[i, 4:7] = n["mx"]
        ints[i, 8] = n["left"]
        ints[i, 9] = n["right"]
        ints[i, 10] = n["primStart"]
        ints[i, 11] = n["primCount"]
    return out


def lights(tris32, mats16):
    """uploadLights over 32-float triangle records in BVH order and 16-float
    material records. Returns ((n, 20) float32 light records, total power).
    Overflow to inf and inf / inf = NaN follow IEEE arithmetic, as in Java."""
    with np.errstate(over="ignore", invalid="ignore"):
        return _lights(tris32, mats16)


def _lights(tris32, mats16):
    pi_f = F(np.pi)
    out, total = [], F(0.0)
    for t in np.asarray(tris32, np.float32):
        m = mats16[int(t[30:31].view(np.int32)[0])]
        strength = F(m[7])
        if strength <= F(0.0):
            continue
        v0, v1, v2 = t[0:3], t[4:7], t[8:11]
        e1 = [F(v1[a]) - F(v0[a]) for a in range(3)]
        e2 = [F(v2[a]) - F(v0[a]) for a in range(3)]
        cx = e1[1] * e2[2] - e1[2] * e2[1]
        cy = e1[2] * e2[0] - e1[0] * e2[2]
        cz = e1[0] * e2[1] - e1[1] * e2[0]
        area = F(0.5) * np.sqrt(cx * cx + cy * cy + cz * cz, dtype=np.float32)
        if area < F(1e-8):
            continue
        er, eg, eb = F(m[4]) * strength, F(m[5]) * strength, F(m[6]) * strength
        lum = F(0.299) * er + F(0.587) * eg + F(0.114) * eb
        power = lum * area * pi_f
        total = total + power
        out.append([v0[0], v0[1], v0[2], area, v1[0], v1[1], v1[2], 0, v2[0], v2[1], v2[2], 0,
                    er, eg, eb, power, t[12], t[13], t[14], 0])
    rec = np.array(out, np.float32).reshape(-1, 20)
    cum = F(0.0)
    for e in rec:
        cum = cum + e[15]
        e[19] = cum / total if total > F(0.0) else F(1.0)
    if len(rec):
        rec[-1, 19] = F(1.0)
    return rec, total
