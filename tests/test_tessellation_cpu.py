"""Scene tessellation of librc_host.so (rcs_add_box, rcs_add_sphere,
rcs_rotate_since) against the oracle's restatement of Box, Sphere, GpuTriangle
and TransformedPrimitive (src/main/java/com/retrocam/scene/Box.java:24-76,
Sphere.java:34-96, GpuTriangle.java:38-50, TransformedPrimitive.java:31-97),
record for record and bit for bit.

Math.sin / Math.cos are specified to within 1 ulp; the oracle's java_trig
proves, per angle, that every double within 1 ulp of the exact value rounds to
one float, so the comparison holds for any conforming JVM. Its exact values
are pinned by known answers (sin and cos at float pi, double pi/2, zero) and
against the platform libm."""
import math

import numpy as np
import pytest

from oracle import rc_oracle_tess as ot
from retrocam_amd.scene import Material, SceneBuilder


f32 = np.float32


def _rows(recs):
    return sorted(np.asarray(recs, np.float32).view(np.uint32).reshape(len(recs), 32).tobytes()[i * 128:(i + 1) * 128]
                  for i in range(len(recs)))


def _host(add):
    """Triangles a SceneBuilder call sequence adds (the BVH build reorders them,
    so records are compared as a multiset)."""
    b = SceneBuilder()
    m = b.add_material(Material.diffuse(0.5, 0.5, 0.5))
    add(b, m)
    return b.build().tris


def test_exact_trig_known_values():
    # sin / cos at the double nearest pi/2 and at float(pi): the arguments miss
    # pi/2 and pi by 6.123233995736766e-17 and -8.742278000372485e-08.
    assert float(ot.exact_trig("cos", math.pi / 2)) == pytest.approx(6.123233995736766e-17, rel=1e-15)
    assert float(ot.exact_trig("sin", float(f32(math.pi)))) == pytest.approx(-8.742278000372485e-08, rel=1e-12)
    assert ot.java_trig("cos", float(f32(math.pi))) == -1.0
    assert ot.java_trig("sin", 0.0) == 0.0 and ot.java_trig("cos", 0.0) == 1.0
    # sin(pi/6) with pi/6 rounded to double: just below 1/2
    from decimal import Decimal
    assert ot.exact_trig("sin", math.pi / 6) < Decimal("0.5") < ot.exact_trig("sin", math.pi / 6) + Decimal("1e-16")


def test_exact_trig_against_libm():
    # The 70-digit value rounds to what the platform's sin / cos return (both
    # are correctly rounded on all of these) and sin^2 + cos^2 = 1.
    rng = np.random.default_rng(7)
    for x in rng.uniform(0.0, 2.0 * math.pi, 300):
        s, c = ot.exact_trig("sin", float(x)), ot.exact_trig("cos", float(x))
        assert abs(s * s + c * c - 1) < 1e-60
        assert float(s) == pytest.approx(math.sin(x), rel=0, abs=2e-16)
        assert float(c) == pytest.approx(math.cos(x), rel=0, abs=2e-16)


def test_rotation_known_answer():
    # 90 degrees about z: toRadians gives the double nearest pi/2, whose cosine
    # is 6.123234e-17 as a float (not 0), sine 1. R maps x to (6.1e-17, 1, -0).
    r = ot.rotation(0.0, 0.0, 90.0)
    assert [float(v) for v in r] == [float(f32(6.123233995736766e-17)), 1.0, -0.0, -1.0,
                                     float(f32(6.123233995736766e-17)), 0.0, 0.0, 0.0, 1.0]
    assert np.signbit(r[2])
    # The unit box's first record turned about the origin: (1,-1,1) -> x' =
    # 6.1e-17 + 1 + 0 = 1, y' = 1 - 6.1e-17 = 1, z' = 1; its normal (-1,0,0) ->
    # (-6.1e-17, -1, 0).
    rec = ot.rotate(ot.box((0, 0, 0), (1, 1, 1), 0), (0, 0, 0), r)[0]
    assert list(rec[0:3]) == [1.0, 1.0, 1.0]
    assert list(rec[12:15]) == [-float(f32(6.123233995736766e-17)), -1.0, 0.0]


def test_box_known_answer():
    # Unit half-extents at the origin: the first (+X) quad (1,-1,1), (1,1,1),
    # (1,1,-1): e1 = (0,2,0), e2 = (0,2,-2), e1 x e2 = (-4,0,0) -> normal (-1,0,0).
    recs = ot.box((0, 0, 0), (1, 1, 1), 0)
    assert recs.shape == (12, 32)
    assert list(recs[0, 0:3]) == [1, -1, 1] and list(recs[0, 12:15]) == [-1, 0, 0]
    assert list(recs[0, 24:30]) == [0, 0, 0, 1, 1, 1] and list(recs[1, 24:30]) == [0, 0, 1, 1, 1, 0]


def test_sphere_known_answer():
    # 2 stacks x 3 slices: row 0 keeps (A, C, D), row 1 keeps (A, B, C). The north
    # pole (phi 0) is exactly (cx, cy + r, cz). The south pole uses phi = float(pi),
    # sin = -8.742278e-08, cos = -1: y = 2 + 2 * -1 = 0; with theta = float(2 pi / 3)
    # (cos -0.50000006, sin 0.86602537) x = 1 + 8.74e-08 rounds up to 1 + 2^-23 and
    # z = 3 - 1.514e-07 rounds down to 3 - 2^-22.
    recs = ot.sphere((1.0, 2.0, 3.0), 2.0, 0, 2, 3)
    assert recs.shape == (6, 32)
    assert list(recs[0, 0:3]) == [1.0, 4.0, 3.0]
    assert list(recs[3, 8:11]) == [1.0 + 2.0 ** -23, 0.0, 3.0 - 2.0 ** -22]
    assert list(recs[3, 24:30]) == [0.0, 0.5, float(f32(1) / f32(3)), 0.5, float(f32(1) / f32(3)), 1.0]


@pytest.mark.parametrize("stacks,slices", [(16, 32), (24, 24), (24, 48), (32, 64), (2, 3), (5, 7)])
def test_sphere_matches_oracle(stacks, slices):
    c, r = (0.37, -1.2, 2.05), 0.83
    host = _host(lambda b, m: b.add_sphere(c, r, m, stacks=stacks, slices=slices))
    assert _rows(host) == _rows(ot.sphere(c, r, 0, stacks, slices))


def test_boxes_rotated_match_oracle():
    # Boxes as the clutter scenes place them: rotations uniform in [0, 90) degrees.
    rng = np.random.default_rng(11)
    boxes = []
    for _ in range(40):
        c = tuple(float(v) for v in rng.uniform(-7, 7, 3))
        rad = float(rng.uniform(0.15, 0.6))
        h = (rad, rad * float(rng.uniform(0.5, 1.5)), rad)
        rot = tuple(float(a) for a in rng.uniform(0.0, 90.0, 3))
        boxes.append((c, h, rot))

    def add(b, m):
        for c, h, rot in boxes:
            b.add_box(c, h, m, rot_deg=rot)

    expect = []
    for c, h, rot in boxes:
        expect.extend(ot.rotate(ot.box(c, h, 0), c, ot.rotation(*rot)))
    assert _rows(_host(add)) == _rows(expect)


def test_unrotated_boxes_match_oracle():
    # The room of the reference's default scene (Scene.java:79-90), half extents 0.05 walls.
    H = 5.0
    spec = [((0, 0, 0), (H, 0.05, H)), ((0, 2 * H, 0), (H, 0.05, H)), ((-H, H, 0), (0.05, H, H)),
            ((H, H, 0), (0.05, H, H)), ((0, H, -H), (H, H, 0.05)), ((-1.5, 1.5, -1.0), (1.0, 1.5, 1.0))]

    def add(b, m):
        for c, h in spec:
            b.add_box(c, h, m)

    expect = np.concatenate([ot.box(c, h, 0) for c, h in spec])
    assert _rows(_host(add)) == _rows(expect)


def _recorded(monkeypatch, make):
    """The add_box / add_sphere calls a workload makes, and its scene."""
    calls = []
    box0, sph0 = SceneBuilder.add_box, SceneBuilder.add_sphere

    def box(self, c, h, mat, rot_deg=None):
        calls.append(("box", tuple(c), tuple(h), rot_deg))
        return box0(self, c, h, mat, rot_deg=rot_deg)

    def sph(self, c, r, mat, stacks=24, slices=24):
        calls.append(("sphere", tuple(c), r, (stacks, slices)))
        return sph0(self, c, r, mat, stacks=stacks, slices=slices)

    monkeypatch.setattr(SceneBuilder, "add_box", box)
    monkeypatch.setattr(SceneBuilder, "add_sphere", sph)
    make()
    monkeypatch.undo()
    return calls


@pytest.mark.parametrize("name", ["full_frame", "sppm_caustics"])
def test_benchmark_scene_angles_are_unambiguous(monkeypatch, name):
    # Every angle of the 250k-triangle frame's (and the caustics scene's) spheres
    # and box rotations has one float sin and cos for any Math.sin / Math.cos
    # within 1 ulp; every rotated box, and one sphere per resolution, equals the
    # oracle's records.
    from retrocam_amd import workloads
    calls = _recorded(monkeypatch, getattr(workloads, name))
    boxes = [c for c in calls if c[0] == "box"]
    spheres = [c for c in calls if c[0] == "sphere"]
    assert spheres and (boxes or name != "full_frame")
    for _, _, _, (st, sl) in spheres:
        for i in range(st + 1):
            phi = float(f32(math.pi * i / st))
            ot.java_trig("sin", phi), ot.java_trig("cos", phi)
        for j in range(sl + 1):
            th = float(f32(2 * math.pi * j / sl))
            ot.java_trig("sin", th), ot.java_trig("cos", th)
    rotated = [c for c in boxes if c[3] is not None and any(a != 0 for a in c[3])]
    one_per_res = list({c[3]: c for c in spheres}.values())

    def add(b, m):
        for _, c, h, rot in rotated:
            b.add_box(c, h, m, rot_deg=rot)
        for _, c, r, (st, sl) in one_per_res:
            b.add_sphere(c, r, m, stacks=st, slices=sl)

    expect = [ot.rotate(ot.box(c, h, 0), c, ot.rotation(*rot)) for _, c, h, rot in rotated]
    expect += [ot.sphere(c, r, 0, st, sl) for _, c, r, (st, sl) in one_per_res]
    assert _rows(_host(add)) == _rows(np.concatenate(expect))
