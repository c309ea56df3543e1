"""The binned-SAH BVH and the light list of librc_host.so (rcs_build) against
the oracle's restatements of BVHBuilder
(src/main/java/com/retrocam/scene/BVHBuilder.java:35-202) and
SceneUploader.uploadLights (src/main/java/com/retrocam/scene/SceneUploader.java:144-216),
node for node, light for light and in triangle order, bit for bit.

The oracle is pinned by hand-worked builds (bucket assignment, split costs, the
partition's swap order, the axis tie rule, the leaf rules, the exception the
Java builder raises on a thin node) and by closed-form light records (areas,
powers and CDFs of equal lights, the skip rules)."""

import numpy as np
import pytest

from oracle import rc_oracle_bvh as ob
from retrocam_amd.scene import Material, SceneBuilder


f32 = np.float32


def _records(verts, mat=None):
    """32-float triangle records; uv0.x carries the input index, normals +z."""
    n = len(verts)
    rec = np.zeros((n, 32), f32)
    v = np.asarray(verts, f32)
    rec[:, 0:3], rec[:, 4:7], rec[:, 8:11] = v[:, 0], v[:, 1], v[:, 2]
    rec[:, 14] = rec[:, 18] = rec[:, 22] = 1.0
    rec[:, 24] = np.arange(n, dtype=f32)
    rec.view(np.int32)[:, 30] = 0 if mat is None else np.asarray(mat, np.int32)
    return rec


def _host_scene(verts, mats=(Material.diffuse(0.5, 0.5, 0.5),), mat=None, normals=None):
    b = SceneBuilder()
    for m in mats:
        b.add_material(m)
    rec = _records(verts, mat)
    if normals is not None:  # (n0, n1, n2) for every triangle
        rec[:, 12:15], rec[:, 16:19], rec[:, 20:23] = normals
    b.add_triangles(rec)
    return b.build()


def _host_build(verts):
    sc = _host_scene(verts)
    return sc.nodes, sc.tris[:, 24].astype(np.int64).tolist()


def _check_same(verts):
    onodes, oorder = ob.build(np.asarray(verts, f32))
    hnodes, horder = _host_build(verts)
    assert horder == oorder
    assert np.array_equal(hnodes.view(np.uint32), ob.nodes_to_records(onodes).view(np.uint32))
    return onodes, oorder


def _strip(a):
    """Triangle (a,0,0), (a+1,0,0), (a+2,1,0): x extent [a, a+2], centroid x a+1."""
    return [(a, 0, 0), (a + 1, 0, 0), (a + 2, 1, 0)]


@pytest.mark.parametrize("plane", ["xy", "xz"])
def test_hand_worked_split_and_partition(plane):
    # Input order A0 B0 A1 B1 A2 B2 (a = 0, 10, 1, 11, 2, 12). Root box x [0, 14],
    # y [0, 1], z 0: parentSA = 28. x buckets (8c/14 truncated): c = 1 -> 0,
    # 2 and 3 -> 1, 11 and 12 -> 6, 13 -> 7. Splits s = 1..5 all give
    # 1 + 2 * (8 * 3 * 1.5) / 28 = 3.571 (the first is kept; splitPos
    # 14 * 2 / 8 = 3.5), s = 0 and 6 give 1 + 201/28. y: every centroid 1/3 ->
    # bucket 2, no split; z: zero extent, skipped. 3.571 < 1.5 * 6 -> split.
    # Partition at 3.5: lo=0 keeps A0; B0 swaps with B2 (hi 5); B2 swaps with A2
    # (hi 4); A2, A1 stay; B1 swaps with itself -> A0 A2 A1 | B1 B2 B0.
    # In the x-z plane (y and z swapped) every area is the dz * dx term instead.
    verts = [_strip(a) for a in (0, 10, 1, 11, 2, 12)]
    if plane == "xz":
        verts = [[(p[0], p[2], p[1]) for p in t] for t in verts]
    nodes, order = _check_same(verts)
    assert order == [0, 4, 2, 3, 5, 1]
    assert len(nodes) == 3
    root, left, right = nodes
    assert (root["left"], root["right"], root["primCount"]) == (1, 2, 0)
    one = [0, 1, 0] if plane == "xy" else [0, 0, 1]
    assert [float(x) for x in root["mn"]] == [0, 0, 0] and [float(x) for x in root["mx"]] == [14] + one[1:]
    assert (left["primStart"], left["primCount"]) == (0, 3)
    assert [float(x) for x in left["mx"]] == [4] + one[1:]
    assert (right["primStart"], right["primCount"]) == (3, 3)
    assert [float(x) for x in right["mn"]] == [10, 0, 0]


def _sym(a, b):
    """(a,b,0), (a+2,b+1,0), (a+1,b+2,0): its mirror under x <-> y is _sym(b, a)."""
    return [(a, b, 0), (a + 2, b + 1, 0), (a + 1, b + 2, 0)]


def test_axis_tie_keeps_the_first_axis():
    # Four 2-triangle clusters at (0,0), (10,0), (0,10), (10,10), mirror-symmetric
    # under x <-> y: the best x split and the best y split cost exactly the same
    # (1 + 2 * (78 * 4 * 1.5) / 338). Only a strict "cost < best" over axes in
    # x, y, z order keeps x: the left leaf then holds the x < 5 clusters.
    verts = []
    for (a, b) in ((0, 0), (10, 0), (0, 10), (10, 10)):
        verts += [_sym(a, b), _sym(a + 1, b + 1)]
    nodes, order = _check_same(verts)
    assert len(nodes) == 3 and nodes[1]["primCount"] == 4
    left = [verts[i] for i in order[:4]]
    assert all(max(p[0] for p in t) < 5 for t in left)
    assert {min(p[1] for p in t) for t in left} == {0, 1, 10, 11}


def test_split_costlier_than_leaf_makes_a_leaf():
    # Five triangles spanning x [-50, 50] (one to -60, one to 60) with
    # centroids x = -20, -10, 0, 10, 20 in buckets 2, 3, 4, 4, 5, all centroids
    # y = 1, z = 0. Every split has both sides with SA 2 * 110 * 2 = 440 of the
    # parent's 480: cost 1 + 7.5 * 440 / 480 = 7.875 >= 1.5 * 5 -> one leaf.
    verts = [[(-50, 0, 0), (50, 1, 0), (30 * k - 60, 2, 0)] for k in range(5)]
    nodes, order = _check_same(verts)
    assert len(nodes) == 1 and nodes[0]["primCount"] == 5 and order == [0, 1, 2, 3, 4]


def _narrow(c):
    """(c,0,0), (c,1,0), (c,0,1): zero x extent at x = c, unit y and z extents."""
    return [(c, 0, 0), (c, 1, 0), (c, 0, 1)]


def test_centroid_on_the_split_goes_right():
    # Narrow triangles at x = 0, 2, 4, 8 and T3 = (0.5,0,0), (0.5,1,0), (8,0,1)
    # with centroid x exactly 3 (input order T0 T2 T3 T4 T8). Box x [0, 8],
    # y and z [0, 1]: parentSA 2 * (8 + 1 + 8) = 34; bucket = centroid x
    # (T8 -> 7). SA of an x extent d: 2 * (2d + 1). s = 2 (pos 3): {T0, T2}
    # 2 * 10 + {T3, T4, T8} 3 * 32 = 116; s = 0/1: 2 + 4 * 32 = 130; s = 3:
    # 3 * 34 + 2 * 18 = 138; s >= 4: 4 * 34 + 2 = 138. 1 + 1.5 * 116 / 34 < 7.5.
    # T3's centroid equals the split position: "c < pos" is false, it goes right.
    verts = [_narrow(0), _narrow(2), [(0.5, 0, 0), (0.5, 1, 0), (8, 0, 1)], _narrow(4), _narrow(8)]
    nodes, order = _check_same(verts)
    assert order == [0, 1, 3, 4, 2]
    assert [(n["primStart"], n["primCount"]) for n in nodes[1:]] == [(0, 2), (2, 3)]


def _flat(x0, x1, x2):
    """Vertices at x = x0, x1, x2 with y = 0, 0, 1 and z = 0."""
    return [(x0, 0, 0), (x1, 0, 0), (x2, 1, 0)]


_WIDE = [_flat(0, 3, 6), _flat(0, 6, 6), _flat(0, 4.5, 6), _flat(0, 6, 6)]  # centroids 3, 4, 3.5, 4


def test_split_cost_equal_to_leaf_cost_makes_a_leaf():
    # Box x [0, 6], y [0, 1], z 0: parentSA 12, bucket = floor(4c / 3). L1 =
    # (0, 1, 2), centroid 1 -> bucket 1; the four wide ones span [0, 6] in
    # buckets 4, 5, 4, 5. s = 1..3: 1 + (4 * 1 * 1.5) / 12 + (12 * 4 * 1.5) / 12
    # = 1 + 0.5 + 6 = 7.5 exactly = 1.5 * 5; s = 4: 1 + 1.5 * (36 + 24) / 12 =
    # 8.5. "cost >= leafCost" makes one leaf.
    nodes, order = _check_same([_flat(0, 1, 2)] + _WIDE)
    assert len(nodes) == 1 and nodes[0]["primCount"] == 5


def test_narrower_first_triangle_splits():
    # As above with L1 = (0, 0.5, 1), centroid 0.5 -> bucket 0: s = 0..3 cost
    # 1 + 1.5 * (2 + 48) / 12 = 7.25 < 7.5, split at 6 / 8 = 0.75. (With an
    # intersection cost of 1 instead of 1.5 the same scene is a leaf:
    # 1 + 50 / 12 >= 5.)
    nodes, order = _check_same([_flat(0, 0.5, 1)] + _WIDE)
    assert len(nodes) == 3 and order[0] == 0
    assert [(n["primStart"], n["primCount"]) for n in nodes[1:]] == [(0, 1), (1, 4)]


def test_collinear_triangles_are_a_leaf():
    # Needles along x (y = z = 0): parentSA 0 < 1e-8 -> a leaf, although the x
    # extent would bin (and every split would cost 1 < 1.5 * 6).
    verts = [[(a, 0, 0), (a + 1, 0, 0), (a + 2, 0, 0)] for a in range(6)]
    nodes, order = _check_same(verts)
    assert len(nodes) == 1 and nodes[0]["primCount"] == 6


def test_axis_thinner_than_1e6_is_not_binned():
    # Four small triangles at x = 0 around (y, z) = (5, 5) and two large ones at
    # x = 5e-7 spanning y, z [0, 10] with centroids 10/3 and 20/3. The x extent
    # 5e-7 < 1e-6 is skipped; y (buckets 2, 3, 5) and z alike can only split off
    # one large triangle: 1 + 1.5 * (200 * 1 + 200 * 5) / 200 = 10 >= 1.5 * 6,
    # so one leaf. (Binning x would split small | large at a cost near 4.1.)
    e = float(f32(5e-7))
    small = [[(0, 4.5, 4.5), (0, 5.5, 4.5), (0, 4.5, 5.5)] for _ in range(4)]
    large = [[(e, 0, 0), (e, 10, 0), (e, 0, 10)], [(e, 10, 10), (e, 0, 10), (e, 10, 0)]]
    nodes, order = _check_same(small + large)
    assert len(nodes) == 1 and nodes[0]["primCount"] == 6


def test_leaf_rules_small_flat_and_coincident():
    four = [_strip(a) for a in (0, 5, 10, 15)]
    nodes, _ = _check_same(four)  # count <= 4
    assert len(nodes) == 1 and nodes[0]["primCount"] == 4
    point = [[(1, 2, 3)] * 3 for _ in range(6)]  # parentSA 0 < 1e-8
    nodes, _ = _check_same(point)
    assert len(nodes) == 1 and nodes[0]["primCount"] == 6
    same = [_strip(0) for _ in range(7)]  # one bucket per axis: no split at all
    nodes, order = _check_same(same)
    assert len(nodes) == 1 and order == list(range(7))


def test_signed_zero_bounds():
    # Math.min(-0.0f, 0.0f) is -0.0f and Math.max is +0.0f.
    verts = [[(0.0, 0, 0), (-0.0, 1, 0), (0.0, 0, 1)]]
    nodes, _ = _check_same(verts)
    assert np.signbit(nodes[0]["mn"][0]) and not np.signbit(nodes[0]["mx"][0])


def test_thin_node_bucket_below_zero_fails_like_java():
    # x = 2.6666674613952637f: (x + x + x) / 3f rounds one ulp below x. In a node
    # of x extent 5 ulp (>= 1e-6), 8 * (c - lo) / (hi - lo) = -1.6 -> (int) -1:
    # BVHBuilder indexes buckets[-1] and throws.
    lo = f32(2.6666674613952637)
    ulp = np.spacing(lo)
    assert ((lo + lo) + lo) / f32(3) == lo - ulp
    verts = [[(lo + k * ulp, 0, 0), (lo + k * ulp, 1, 0), (lo + k * ulp, 0, 1)] for k in (0, 1, 2, 3, 5)]
    verts = np.asarray(verts, f32)
    with pytest.raises(ob.BucketIndexError):
        ob.build(verts)
    with pytest.raises(RuntimeError):
        _host_build(verts)


def test_java_int_cast():
    assert [ob.java_int(f32(x)) for x in (1.9, -0.9, -1.6, 7.99)] == [1, 0, -1, 7]
    assert ob.java_int(f32("nan")) == 0
    assert ob.java_int(f32("inf")) == 2147483647 and ob.java_int(f32("-inf")) == -2147483648


def _random_scene(rng, n, kind):
    if kind == "uniform":
        c = rng.uniform(-10, 10, (n, 3))
    elif kind == "clusters":  # few distinct centres, many repeats
        c = rng.choice(rng.uniform(-5, 5, (4, 3)), n) + rng.integers(0, 3, (n, 1)) * 0.25
    elif kind == "sheet":  # every vertex on z = 0 (zero z extent)
        c = np.concatenate([rng.uniform(-3, 3, (n, 2)), np.zeros((n, 1))], 1)
    else:  # "zeros": coordinates of +-0 and small integers
        c = rng.integers(-2, 3, (n, 3)).astype(np.float64) * rng.choice([1.0, -1.0], (n, 3))
    off = rng.uniform(-0.6, 0.6, (n, 3, 3))
    if kind == "sheet":
        off[:, :, 2] = 0
    if kind == "zeros":
        off = rng.integers(-1, 2, (n, 3, 3)) * rng.choice([0.0, -0.0, 0.5], (n, 3, 3))
    return (c[:, None, :] + off).astype(f32)


@pytest.mark.parametrize("kind", ["uniform", "clusters", "sheet", "zeros"])
@pytest.mark.parametrize("n", [5, 23, 150, 2000])
def test_random_scenes_match_oracle(kind, n):
    rng = np.random.default_rng(1000 * n + len(kind))
    _check_same(_random_scene(rng, n, kind))


def test_cornell_scene_matches_oracle():
    # Boxes, quads and tessellated spheres (996 triangles), rebuilt from the
    # triangle order of a first build.
    from retrocam_amd import workloads
    t = workloads.cornell_small().scene.tris
    _check_same(np.stack([t[:, 0:3], t[:, 4:7], t[:, 8:11]], 1))


def _check_lights(sc):
    """The host's light records and total power equal the oracle's, bit for bit."""
    rec, total = ob.lights(sc.tris, sc.mats)
    assert sc.nLights == len(rec)
    if len(rec):
        assert np.array_equal(sc.lights[:len(rec)].view(np.uint32), rec.view(np.uint32))
    assert f32(sc.totalLightPower).view(np.uint32) == total.view(np.uint32)
    return rec, total


def _right(x0, z=0.0):
    """Right triangle with legs 2 and 2 in the plane z: area exactly 2."""
    return [(x0, 0, z), (x0 + 2, 0, z), (x0, 2, z)]


def test_light_record_closed_form():
    # Area 0.5 * sqrt(4^2) = 2; emission (1, 0, 0) x strength 1: lum = 0.299f,
    # power = (0.299f * 2) * (float) pi with one rounding (the doubling is
    # exact). One light: CDF 1, total = its power; the normal is n0's.
    sc = _host_scene([_right(0)], mats=[Material(emission=(1.0, 0.0, 0.0), emissionStrength=1.0)],
                     normals=((0, 0, 1), (1, 0, 0), (0, 1, 0)))
    rec, total = _check_lights(sc)
    assert rec.shape == (1, 20) and rec[0, 3] == 2.0
    assert rec[0, 15] == f32(f32(0.598) * f32(np.pi)) and total == rec[0, 15] and rec[0, 19] == 1.0
    assert list(rec[0, 16:19]) == [0, 0, 1] and list(rec[0, 12:15]) == [1, 0, 0]


def test_equal_lights_cdf():
    # Two equal lights: cumulative / sum = p / 2p = 0.5 exactly (2p is exact),
    # then 1. Four: 0.25 and 0.5 exactly, 3p / 4p within an ulp of 0.75 (3p
    # rounds), and the last entry forced to 1.
    em = Material(emission=(1.0, 1.0, 1.0), emissionStrength=4.0)
    rec, total = _check_lights(_host_scene([_right(0), _right(10)], mats=[em]))
    assert list(rec[:, 19]) == [0.5, 1.0] and total == rec[0, 15] * 2
    rec, total = _check_lights(_host_scene([_right(10 * k) for k in range(4)], mats=[em]))
    assert np.all(rec[:, 15] == rec[0, 15])
    assert list(rec[[0, 1, 3], 19]) == [0.25, 0.5, 1.0] and abs(rec[2, 19] - 0.75) <= 2.0 ** -24


def test_overflowing_power_cdf():
    # A strength of 3e38 overflows the power to +inf: the total is inf, the first
    # entry's CDF p / inf = 0 and the last one's inf / inf = NaN, replaced by 1.
    em = Material(emission=(1.0, 1.0, 1.0), emissionStrength=1.0)
    hot = Material(emission=(1.0, 1.0, 1.0), emissionStrength=3e38)
    sc = _host_scene([_right(0), _right(10)], mats=[em, hot], mat=[0, 1])
    rec, total = _check_lights(sc)
    assert np.isinf(total) and sorted(rec[:, 19].tolist()) == [0.0, 1.0]


def test_light_skip_rules():
    em = Material(emission=(1.0, 1.0, 1.0), emissionStrength=1.0)
    off = Material(emission=(1.0, 1.0, 1.0), emissionStrength=0.0)
    neg = Material(emission=(1.0, 1.0, 1.0), emissionStrength=-1.0)
    black = Material(emission=(0.0, 0.0, 0.0), emissionStrength=1.0)
    # strength <= 0 -> not a light; a collinear (zero-area) emitter -> skipped
    needle = [(0, 0, 0), (1, 1, 1), (2, 2, 2)]
    sc = _host_scene([_right(0), _right(5), needle, _right(9)], mats=[em, off, neg], mat=[1, 2, 0, 0])
    rec, _ = _check_lights(sc)
    assert len(rec) == 1 and rec[0, 0] == 9.0
    # zero luminance but positive strength: a light of power 0; all powers 0 -> CDF 1
    rec, total = _check_lights(_host_scene([_right(0), _right(5)], mats=[black]))
    assert len(rec) == 2 and total == 0 and list(rec[:, 19]) == [1.0, 1.0]
    # no emitter at all
    sc = _host_scene([_right(0)], mats=[off])
    assert sc.nLights == 0 and sc.totalLightPower == 0


@pytest.mark.parametrize("seed", [3, 4, 5])
def test_random_emitters_match_oracle(seed):
    # Lights are listed in BVH triangle order with their CDF.
    rng = np.random.default_rng(seed)
    n = 120
    verts = _random_scene(rng, n, "uniform")
    mats = [Material.diffuse(0.5, 0.5, 0.5)] + [
        Material(emission=tuple(rng.uniform(0, 2, 3)), emissionStrength=float(rng.uniform(0.1, 20)))
        for _ in range(5)]
    sc = _host_scene(verts, mats=mats, mat=rng.integers(0, 6, n))
    rec, _ = _check_lights(sc)
    assert len(rec) > 10


def test_workload_lights_match_oracle():
    from retrocam_amd import workloads
    for w in (workloads.cornell_small(), workloads.sppm_caustics()):
        _check_lights(w.scene)
