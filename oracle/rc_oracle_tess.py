"""CPU oracle of the scene tessellation: Box.toTriangles / addQuad
(src/main/java/com/retrocam/scene/Box.java:24-76), GpuTriangle.withFlatNormal
(src/main/java/com/retrocam/scene/GpuTriangle.java:38-50), Sphere.toTriangles /
spherePt / setNormal (src/main/java/com/retrocam/scene/Sphere.java:34-96) and the
rotation of TransformedPrimitive
(src/main/java/com/retrocam/scene/TransformedPrimitive.java:31-97).

TEST INFRASTRUCTURE ONLY: used by tests/ to check librc_host.so's rcs_add_box,
rcs_add_sphere and rcs_rotate_since. Nothing in the product imports it.

Java semantics: float arithmetic is binary32, left to right; the Math.PI
expressions are double; (float) Math.sqrt of a float is the correctly rounded
binary32 square root; Math.toRadians(a) = a * 0.017453292519943295 (JDK 9 and
later). Math.sin and Math.cos are specified only to within 1 ulp of the exact
result, so java_trig evaluates the exact value to 70 digits and returns the
float to which every double within 1 ulp of it rounds; it raises TrigAmbiguous
where two such doubles round to different floats (the float would then depend
on the JVM's implementation).
"""
import functools
import math
from decimal import Decimal, localcontext

import numpy as np

F = np.float32
DEG2RAD = 0.017453292519943295


class TrigAmbiguous(ValueError):
    """Doubles within 1 ulp of sin(x) or cos(x) round to different floats."""


def exact_trig(fn, x):
    """sin(x) or cos(x) of the double x as a 70-digit Decimal (Taylor series;
    |x| <= 8 keeps the cancellation under 4 digits)."""
    if abs(x) > 8.0:
        raise ValueError("argument outside the series' range")
    with localcontext() as ctx:
        ctx.prec = 80
        X = Decimal(x)
        x2 = X * X
        term, n = (X, 1) if fn == "sin" else (Decimal(1), 0)
        total = term
        tiny = Decimal(10) ** -90
        while True:
            term = -term * x2 / ((n + 1) * (n + 2))
            n += 2
            if abs(term) < tiny:
                break
            total += term
        ctx.prec = 70
        return +total


@functools.lru_cache(maxsize=None)
def java_trig(fn, x):
    """(float) Math.sin(x) / (float) Math.cos(x) for any implementation within
    1 ulp of the exact result."""
    ex = exact_trig(fn, x)
    d = float(ex)  # correctly rounded
    if Decimal(d) == ex:
        cands = [d]
    else:
        lo = d if Decimal(d) < ex else math.nextafter(d, -math.inf)
        cands = [lo, math.nextafter(lo, math.inf)]  # the doubles bracketing the exact value
    fl = {F(c).view(np.uint32).item() for c in cands}
    if len(fl) != 1:
        raise TrigAmbiguous(f"{fn}({x!r})")
    return F(cands[0])


def _sqrt(x):
    return np.sqrt(F(x), dtype=np.float32)


def _record(v0, v1, v2, n0, n1, n2, uv0, uv1, uv2, mat):
    """GpuTriangle.write (GpuTriangle.java:68-79): 32 floats."""
    o = np.zeros(32, np.float32)
    o[0:3], o[4:7], o[8:11] = v0, v1, v2
    o[12:15], o[16:19], o[20:23] = n0, n1, n2
    o[24:26], o[26:28], o[28:30] = uv0, uv1, uv2
    o[30:31].view(np.int32)[0] = mat
    return o


def flat_normal(v0, v1, v2):
    """GpuTriangle.withFlatNormal."""
    ax, ay, az = v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]
    bx, by, bz = v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]
    nx = ay * bz - az * by
    ny = az * bx - ax * bz
    nz = ax * by - ay * bx
    ln = _sqrt(nx * nx + ny * ny + nz * nz)
    if ln > F(1e-8):
        nx, ny, nz = nx / ln, ny / ln, nz / ln
    return [nx, ny, nz]


def quad(a, b, c, d, mat):
    """Box.addQuad: (a, b, c) and (a, c, d) with flat normals."""
    a, b, c, d = ([F(x) for x in p] for p in (a, b, c, d))
    n1 = flat_normal(a, b, c)
    n2 = flat_normal(a, c, d)
    return [_record(a, b, c, n1, n1, n1, (0, 0), (0, 1), (1, 1), mat),
            _record(a, c, d, n2, n2, n2, (0, 0), (1, 1), (1, 0), mat)]


def box(c, h, mat):
    """Box.toTriangles: 12 records, faces +X, -X, +Y, -Y, +Z, -Z."""
    cx, cy, cz = (F(v) for v in c)
    hx, hy, hz = (F(v) for v in h)
    x0, x1, y0, y1, z0, z1 = cx - hx, cx + hx, cy - hy, cy + hy, cz - hz, cz + hz
    faces = [
        ((x1, y0, z1), (x1, y1, z1), (x1, y1, z0), (x1, y0, z0)),
        ((x0, y0, z0), (x0, y1, z0), (x0, y1, z1), (x0, y0, z1)),
        ((x0, y1, z1), (x0, y1, z0), (x1, y1, z0), (x1, y1, z1)),
        ((x0, y0, z0), (x0, y0, z1), (x1, y0, z1), (x1, y0, z0)),
        ((x0, y0, z1), (x0, y1, z1), (x1, y1, z1), (x1, y0, z1)),
        ((x1, y0, z0), (x1, y1, z0), (x0, y1, z0), (x0, y0, z0)),
    ]
    out = []
    for f in faces:
        out += quad(*f, mat)
    return np.array(out)


def sphere(c, radius, mat, stacks, slices):
    """Sphere.toTriangles: the (stacks x slices) grid, pole rows one triangle per cell."""
    cx, cy, cz = (F(v) for v in c)
    r = F(radius)

    def pt(phi, theta):
        sin_p = java_trig("sin", float(phi))
        return [cx + r * sin_p * java_trig("cos", float(theta)),
                cy + r * java_trig("cos", float(phi)),
                cz + r * sin_p * java_trig("sin", float(theta))]

    def normal(p):
        nx, ny, nz = p[0] - cx, p[1] - cy, p[2] - cz
        ln = _sqrt(nx * nx + ny * ny + nz * nz)
        if ln > F(1e-8):
            nx, ny, nz = nx / ln, ny / ln, nz / ln
        return [nx, ny, nz]

    def tri(p0, p1, p2, t0, t1, t2):
        return _record(p0, p1, p2, normal(p0), normal(p1), normal(p2), t0, t1, t2, mat)

    out = []
    for i in range(stacks):
        phi0 = F(math.pi * i / stacks)
        phi1 = F(math.pi * (i + 1) / stacks)
        for j in range(slices):
            th0 = F(2 * math.pi * j / slices)
            th1 = F(2 * math.pi * (j + 1) / slices)
            A, B, C, D = pt(phi0, th0), pt(phi0, th1), pt(phi1, th1), pt(phi1, th0)
            uvA = (F(j) / F(slices), F(i) / F(stacks))
            uvB = (F(j + 1) / F(slices), F(i) / F(stacks))
            uvC = (F(j + 1) / F(slices), F(i + 1) / F(stacks))
            uvD = (F(j) / F(slices), F(i + 1) / F(stacks))
            if i != 0:
                out.append(tri(A, B, C, uvA, uvB, uvC))
            if i != stacks - 1:
                out.append(tri(A, C, D, uvA, uvC, uvD))
    return np.array(out)


def rotation(deg_x, deg_y, deg_z):
    """TransformedPrimitive.buildRotation: R = Rz Ry Rx, column-major 3 x 3."""
    rx, ry, rz = (float(F(d)) * DEG2RAD for d in (deg_x, deg_y, deg_z))
    cx, sx = java_trig("cos", rx), java_trig("sin", rx)
    cy, sy = java_trig("cos", ry), java_trig("sin", ry)
    cz, sz = java_trig("cos", rz), java_trig("sin", rz)
    return [cy * cz, cy * sz, -sy,
            sx * sy * cz - cx * sz, sx * sy * sz + cx * cz, sx * cy,
            cx * sy * cz + sx * sz, cx * sy * sz - sx * cz, cx * cy]


def rotate(records, pivot, rot):
    """TransformedPrimitive.toTriangles: positions about the pivot, normals by R."""
    px, py, pz = (F(v) for v in pivot)

    def rv(x, y, z):
        return [rot[0] * x + rot[3] * y + rot[6] * z,
                rot[1] * x + rot[4] * y + rot[7] * z,
                rot[2] * x + rot[5] * y + rot[8] * z]

    out = np.array(records, np.float32, copy=True)
    for o in out:
        for k in (0, 4, 8):
            v = rv(o[k] - px, o[k + 1] - py, o[k + 2] - pz)
            o[k], o[k + 1], o[k + 2] = v[0] + px, v[1] + py, v[2] + pz
        for k in (12, 16, 20):
            o[k:k + 3] = rv(o[k], o[k + 1], o[k + 2])
    return out
