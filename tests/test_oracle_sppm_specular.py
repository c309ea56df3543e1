"""CPU pins of the SPPM oracle's specular walks (rc_oracle_sppm.c) by geometry
the mathematics fixes:

- photon_trace.comp:205-250, metal: a mirror above an upward-facing light sends
  every caustic photon down; each stored photon traced back (to the mirror, then
  reflected, to the light's emission plane) lands inside the light, and its power
  is the emitted power (emission / luma * total power / N) times the mirror's
  albedo;
- photon_trace.comp:215-232, glass: photons crossing a parallel glass slab keep
  their power (glass does not change it, whatever its albedo); with ior 1 they
  keep their direction (straight back-trace to the light); with ior 1.5 they
  come out parallel to their entry direction after a lateral shift fixed by
  Snell's law through the slab's thickness;
- photon_gather.comp:140-170: the gather seen through a mirror equals the gather
  of the mirrored camera with the mirror removed (the walk reflects and adds no
  throughput), and a glass slab of ior 1 in front of the camera changes nothing.

Each pin also runs against oracle builds with one planted mistake."""
import ctypes as C

import numpy as np
import pytest

from oracle import rco
from retrocam_amd import _native as N
from retrocam_amd.scene import Material, SceneBuilder
from oracle_variants import pin_fails
from test_oracle_sppm import _bind, _trace

E = (1.0, 0.8, 0.6)
STRENGTH = 5.0
LIGHT_Y, LIGHT_S = 1.0, 0.3
MIRROR_Y = 3.0
MIRROR_ALBEDO = (0.8, 0.6, 0.4)
PHOTONS = 20000


def quad_up(b, y, s, mat, cx=0.0, cz=0.0):
    """Horizontal quad of half-size s at height y, normal +y."""
    b.add_quad((cx - s, y, cz - s), (cx - s, y, cz + s), (cx + s, y, cz + s), (cx + s, y, cz - s), mat)


def mirror_scene(with_mirror=True, dy=0.0):
    b = SceneBuilder()
    floor = b.add_material(Material.diffuse(0.7, 0.7, 0.7))
    light = b.add_material(Material.emissive(E[0], E[1], E[2], STRENGTH))
    quad_up(b, dy, 20.0, floor)
    quad_up(b, LIGHT_Y + dy, LIGHT_S, light)
    if with_mirror:
        quad_up(b, MIRROR_Y + dy, 6.0, b.add_material(Material.metal(*MIRROR_ALBEDO, 0.3)))
    return b.build()


def emitted_power(sc, n):
    em = np.array(E, np.float32) * np.float32(STRENGTH)
    lum = np.float32(0.299) * em[0] + np.float32(0.587) * em[1] + np.float32(0.114) * em[2]
    return (em / lum) * (np.float32(sc.totalLightPower) / np.float32(n))


def inside_light(x, z, tol=2e-4):
    return (np.abs(x) <= LIGHT_S + tol) & (np.abs(z) <= LIGHT_S + tol)


def pin_metal_photons_trace_back_to_the_light():
    _bind()
    sc = mirror_scene()
    assert np.allclose(sc.lights[:, 16:19], [0, 1, 0])  # the light faces +y
    ph, _, _, n = _trace(sc, photons=PHOTONS, radius=0.05)
    assert n > PHOTONS // 4, n
    p, pw, d = ph[:n, 0:3].astype(np.float64), ph[:n, 4:7].astype(np.float64), ph[:n, 8:11].astype(np.float64)
    exp = emitted_power(sc, PHOTONS) * np.array(MIRROR_ALBEDO, np.float32)
    assert np.max(np.abs(pw / exp - 1)) < 2e-6, "power = emitted power x mirror albedo"
    assert np.all(d[:, 1] < 0), "stored after the mirror: travelling down"
    # back along -d to the reflection's origin (2e-3 below the mirror), reflect, back to y = light + 1e-3
    t1 = ((MIRROR_Y - 2e-3) - p[:, 1]) / (-d[:, 1])
    m = p - d * t1[:, None]
    d0 = d * np.array([1.0, -1.0, 1.0])
    t0 = (MIRROR_Y - (LIGHT_Y + 1e-3)) / d0[:, 1]
    e = m + np.array([0.0, 2e-3, 0.0]) - d0 * t0[:, None]
    ok = inside_light(e[:, 0], e[:, 2])
    assert np.all(ok), f"{np.count_nonzero(~ok)} photons do not trace back to the light"


def quad_down(b, y, s, mat):
    """Horizontal quad of half-size s at height y, normal -y."""
    b.add_quad((-s, y, -s), (s, y, -s), (s, y, s), (-s, y, s), mat)


def slab_scene(ior, albedo=(0.5, 0.5, 0.5), half=8.0, y0=1.5, y1=2.5):
    """A glass slab between two faces with outward normals (the faces a glass
    object's surface would have; Box.toTriangles' flat normals point inwards)."""
    b = SceneBuilder()
    floor = b.add_material(Material.diffuse(0.7, 0.7, 0.7))
    light = b.add_material(Material.emissive(E[0], E[1], E[2], STRENGTH))
    ceil = b.add_material(Material.diffuse(0.7, 0.7, 0.7))
    glass = b.add_material(Material(albedo=albedo, transmission=1.0, roughness=0.0, ior=ior))
    quad_up(b, 0.0, 20.0, floor)
    quad_up(b, LIGHT_Y, 0.05, light)
    quad_up(b, 4.0, 20.0, ceil)
    quad_down(b, y0, half, glass)
    quad_up(b, y1, half, glass)
    return b.build(), 0.05, y0, y1


def _ceiling_photons(sc, n_photons=PHOTONS):
    ph, _, _, n = _trace(sc, photons=n_photons, radius=0.05)
    p, pw, d = ph[:n, 0:3].astype(np.float64), ph[:n, 4:7].astype(np.float64), ph[:n, 8:11].astype(np.float64)
    up = d[:, 1] > 0
    return p[up], pw, d[up]


def pin_glass_photons_keep_power_and_refract():
    _bind()
    for ior in (1.0, 1.5):
        sc, s, y0, y1 = slab_scene(ior)
        p, pw, d = _ceiling_photons(sc)
        assert len(p) > PHOTONS // 4
        exp = emitted_power(sc, PHOTONS)
        assert np.max(np.abs(pw / exp - 1)) < 2e-6, "glass keeps the photon power"
        if ior == 1.0:
            t = (p[:, 1] - (LIGHT_Y + 1e-3)) / d[:, 1]
            e = p - d * t[:, None]
            frac = np.mean((np.abs(e[:, 0]) <= s + 2e-4) & (np.abs(e[:, 2]) <= s + 2e-4))
            assert frac >= 0.97, frac
        else:
            # back through the slab: above and below it the photon travels along d; inside
            # along the refracted direction (same azimuth, sin_t = sin_i / ior)
            sin_i = np.sqrt(d[:, 0] ** 2 + d[:, 2] ** 2)
            sin_t = sin_i / ior
            cos_t = np.sqrt(1 - sin_t ** 2)
            hdir = np.where(sin_i[:, None] > 0, np.stack([d[:, 0], d[:, 2]], 1) / np.maximum(sin_i, 1e-30)[:, None], 0)
            dt = np.stack([hdir[:, 0] * sin_t, cos_t, hdir[:, 1] * sin_t], 1)
            a = p - d * ((p[:, 1] - y1) / d[:, 1])[:, None]          # exit point on the top face
            b_ = a - dt * ((y1 - y0) / dt[:, 1])[:, None]             # entry point on the bottom face
            e = b_ - d * ((y0 - (LIGHT_Y + 1e-3)) / d[:, 1])[:, None]
            good = (np.abs(e[:, 0]) <= s + 5e-4) & (np.abs(e[:, 2]) <= s + 5e-4)
            assert np.mean(good) >= 0.97, np.mean(good)
            straight = p - d * ((p[:, 1] - (LIGHT_Y + 1e-3)) / d[:, 1])[:, None]
            assert np.mean((np.abs(straight[:, 0]) <= s) & (np.abs(straight[:, 2]) <= s)) < 0.5


def _gather(sc, cam, ph, heads, nxt, radius, hash_size):
    L = _bind()
    acc = np.zeros((cam.width * cam.height, 4), np.float32)
    os_ = rco.OracleScene(sc)
    assert L.rco_photon_gather(C.byref(os_.c), C.byref(cam), radius, hash_size, rco._fp(ph),
                               heads.ctypes.data_as(C.POINTER(C.c_int32)), nxt.ctypes.data_as(C.POINTER(C.c_int32)),
                               rco._fp(acc), None) == 0
    return acc[:, :3].astype(np.float64)


def _cam(P, F, R, U, W=48, H=36, tan_half=0.35):
    c = N.RcCamera()
    c.camPos[:], c.camForward[:], c.camRight[:], c.camUp[:] = list(P), list(F), list(R), list(U)
    c.fovTanHalf, c.width, c.height, c.focalDist = tan_half, W, H, 2.0
    return c


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def pin_gather_walks_mirrors_and_clear_glass():
    # a large hash: two of a pixel's 27 neighbour cells sharing a bucket make the
    # reference walk that chain twice, which depends on the exact hit point
    # (and the floor lifted off y = 0: the cell hash of cells at y = -1, 0, 1 maps
    # several of a pixel's 27 neighbours to one bucket)
    radius, hs, dy = 0.12, 1 << 22, 0.61
    sc = mirror_scene(dy=dy)
    ph, heads, nxt, n = _trace(sc, photons=PHOTONS, radius=radius, hash_size=hs)
    assert n > 0
    # a camera below the mirror looking up at it sees the floor's caustic reflected
    P = np.array([0.0, 2.0 + dy, 2.5])
    F = _unit([0.0, 1.0, -0.9])
    R = np.array([1.0, 0.0, 0.0])
    U = np.cross(R, F)
    got = _gather(sc, _cam(P, F, R, U), ph, heads, nxt, radius, hs)
    M = np.array([1.0, -1.0, 1.0])
    Pm = P * M + np.array([0.0, 2 * (MIRROR_Y + dy), 0.0])
    ref = _gather(mirror_scene(False, dy), _cam(Pm, F * M, R * M, U * M), ph, heads, nxt, radius, hs)
    assert np.count_nonzero(ref.sum(1) > 0) > 50, "the view holds caustic photons"
    same = np.max(np.abs(got - ref), axis=1) <= 1e-4 * np.max(ref) + 1e-6
    assert np.mean(same) >= 0.999, np.mean(same)
    # a clear (ior 1) glass slab between the camera and the floor changes nothing
    b = SceneBuilder()
    floor = b.add_material(Material.diffuse(0.7, 0.7, 0.7))
    light = b.add_material(Material.emissive(E[0], E[1], E[2], STRENGTH))
    quad_up(b, dy, 20.0, floor)
    quad_up(b, LIGHT_Y + dy, LIGHT_S, light)
    b.add_box((0.0, 5.0 + dy, 0.0), (8.0, 0.5, 8.0), b.add_material(Material(albedo=(0.5, 0.5, 0.5), transmission=1.0,
                                                                       roughness=0.0, ior=1.0)))
    slab = b.build()
    top = _cam((0.0, 8.0 + dy, 0.0), (0.0, -1.0, 0.0), (1.0, 0.0, 0.0), (0.0, 0.0, -1.0), tan_half=0.5)
    g1 = _gather(slab, top, ph, heads, nxt, radius, hs)
    g0 = _gather(mirror_scene(False, dy), top, ph, heads, nxt, radius, hs)
    assert np.count_nonzero(g0.sum(1) > 0) > 50
    same = np.max(np.abs(g1 - g0), axis=1) <= 1e-4 * np.max(g0) + 1e-6
    assert np.mean(same) >= 0.999, np.mean(same)


PINS = {f.__name__: f for f in (pin_metal_photons_trace_back_to_the_light, pin_glass_photons_keep_power_and_refract,
                                pin_gather_walks_mirrors_and_clear_glass)}


@pytest.mark.parametrize("name", list(PINS))
def test_pin(name):
    PINS[name]()


MISTAKES = [
    ("pin_metal_photons_trace_back_to_the_light", "            power = mul(power, ld3(M));\n", ""),
    ("pin_metal_photons_trace_back_to_the_light",
     "            power = mul(power, ld3(M));\n            rd = reflect3(rd, n);",
     "            power = mul(power, ld3(M));\n            rd = reflect3(rd, h.normal);\n            rd.y = -rd.y;"),
    ("pin_metal_photons_trace_back_to_the_light", "float pp = s->totalLightPower / (float)photonCount;",
     "float pp = s->totalLightPower / (float)(photonCount + 1);"),
    ("pin_glass_photons_keep_power_and_refract", "            ro = add(h.pos, muls(rd, 2e-3f));\n            spec = 1;",
     "            power = mul(power, ld3(M));\n            ro = add(h.pos, muls(rd, 2e-3f));\n            spec = 1;"),
    ("pin_glass_photons_keep_power_and_refract",
     "            float eta = entering ? (1.0f / ior) : ior;\n            float cosI",
     "            float eta = entering ? ior : (1.0f / ior);\n            float cosI"),
    ("pin_glass_photons_keep_power_and_refract", "            if (rnd(&st) < F) {\n                rd = reflect3(rd, n);",
     "            if (rnd(&st) < 1.0f - F) {\n                rd = reflect3(rd, n);"),
    ("pin_gather_walks_mirrors_and_clear_glass",
     "                } else {\n                    rd = reflect3(rd, n);\n                }",
     "                } else {\n                    rd = reflect3(rd, h.normal);\n                    rd.x = -rd.x;\n                }"),
    ("pin_gather_walks_mirrors_and_clear_glass",
     "                    if (dot3(rf, rf) < 0.001f) rf = reflect3(rd, n);\n                    rd = normalize3(rf);\n                } else",
     "                    rf = reflect3(rd, n);\n                    rd = normalize3(rf);\n                } else"),
]


@pytest.mark.parametrize("pin,old,new", MISTAKES, ids=[f"{m[0]}-{i}" for i, m in enumerate(MISTAKES)])
def test_planted_mistakes_fail(pin, old, new):
    assert pin_fails(PINS[pin], "rc_oracle_sppm.c", old, new), f"{pin} passes with the planted mistake"
