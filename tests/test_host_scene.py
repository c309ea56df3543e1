"""CPU tests of the host-side scene assembly (librc_host.so): tessellation,
binned-SAH BVH invariants and the light CDF (SceneUploader.uploadLights)."""
import numpy as np
import pytest

from retrocam_amd import workloads
from retrocam_amd.scene import Material, SceneBuilder

f32 = np.float32


def _leaves(sc):
    ni = sc.nodes.view(np.int32)
    return ni[:, 8], ni[:, 9], ni[:, 10], ni[:, 11]


def check_bvh(sc):
    L, R, PS, PC = _leaves(sc)
    n = sc.nNodes
    covered = np.zeros(sc.nTris, np.int32)
    t = sc.tris
    tmin = np.minimum(np.minimum(t[:, 0:3], t[:, 4:7]), t[:, 8:11])
    tmax = np.maximum(np.maximum(t[:, 0:3], t[:, 4:7]), t[:, 8:11])
    bmin = sc.nodes[:, 0:3]
    bmax = sc.nodes[:, 4:7]
    for i in range(n):
        if PC[i] > 0:
            assert L[i] == -1 and R[i] == -1
            covered[PS[i]:PS[i] + PC[i]] += 1
            # node bounds are exactly the union of its triangles (BVHNode.fromTriangles)
            assert np.array_equal(bmin[i], tmin[PS[i]:PS[i] + PC[i]].min(0))
            assert np.array_equal(bmax[i], tmax[PS[i]:PS[i] + PC[i]].max(0))
        else:
            # DFS pre-order: left child directly follows its parent
            assert L[i] == i + 1 and R[i] > L[i] and R[i] < n
            for c in (L[i], R[i]):
                assert np.all(bmin[i] <= bmin[c]) and np.all(bmax[i] >= bmax[c])
            assert np.array_equal(bmin[i], np.minimum(bmin[L[i]], bmin[R[i]]))
            assert np.array_equal(bmax[i], np.maximum(bmax[L[i]], bmax[R[i]]))
    assert np.all(covered == 1), "every triangle in exactly one leaf"


def test_bvh_invariants_cornell():
    w = workloads.cornell_small()
    assert w.scene.nTris == 996
    check_bvh(w.scene)


@pytest.mark.parametrize("seed", [1, 2])
def test_bvh_invariants_random(seed):
    check_bvh(workloads.tiny_random(seed, n_tris=300))


def test_bvh_two_clusters_known_split():
    # Two clusters of 4 triangles far apart: the binned SAH puts them in two
    # 4-triangle leaves (LEAF_MAX = 4), left cluster (smaller x centroids) first.
    b = SceneBuilder()
    m = b.add_material(Material.diffuse(0.5, 0.5, 0.5))
    for cx in (-10.0, -9.5, -9.0, -8.5, 8.5, 9.0, 9.5, 10.0):
        b.add_quad((cx, 0, 0), (cx + 0.1, 0, 0), (cx + 0.1, 0.1, 0), (cx, 0.1, 0), m)
    sc = b.build()
    assert sc.nTris == 16
    check_bvh(sc)
    L, R, PS, PC = _leaves(sc)
    # root splits x at the first bucket boundary that separates the clusters
    left_tris = sc.tris[PS[1]:PS[1] + PC[1]] if PC[1] > 0 else None
    assert L[0] == 1
    # every triangle under the root's left child lies in the x<0 cluster
    stack = [1]
    xs = []
    while stack:
        i = stack.pop()
        if PC[i] > 0:
            xs.extend(sc.tris[PS[i]:PS[i] + PC[i], 0].tolist())
        else:
            stack += [L[i], R[i]]
    assert len(xs) == 8 and max(xs) < 0
    del left_tris


def test_box_tessellation():
    b = SceneBuilder()
    m = b.add_material(Material.diffuse(0.5, 0.5, 0.5))
    assert b.add_box((1.0, 2.0, 3.0), (0.5, 1.0, 1.5), m) == 12
    sc = b.build()
    v = sc.tris[:, [0, 1, 2, 4, 5, 6, 8, 9, 10]].reshape(-1, 3)
    c = np.float32([1.0, 2.0, 3.0])
    h = np.float32([0.5, 1.0, 1.5])
    on_face = np.isclose(np.abs(v - c), h, rtol=0, atol=1e-6).any(axis=1)
    assert on_face.all()
    n = sc.tris[:, 12:15]
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-6)
    assert np.allclose(np.abs(n).max(axis=1), 1.0, atol=1e-6)  # axis aligned


@pytest.mark.parametrize("stacks,slices", [(16, 32), (24, 24), (2, 3)])
def test_sphere_tessellation(stacks, slices):
    b = SceneBuilder()
    m = b.add_material(Material.diffuse(0.5, 0.5, 0.5))
    n = b.add_sphere((0.5, -1.0, 2.0), 1.5, m, stacks=stacks, slices=slices)
    assert n == slices * (2 * stacks - 2)  # pole rows skip one triangle each (Sphere.java:57-58)
    sc = b.build()
    c = np.float32([0.5, -1.0, 2.0])
    for k in range(3):
        p = sc.tris[:, 4 * k:4 * k + 3]
        nn = sc.tris[:, 12 + 4 * k:15 + 4 * k]
        assert np.allclose(np.linalg.norm(p - c, axis=1), 1.5, atol=2e-6)
        assert np.allclose(nn, (p - c) / np.linalg.norm(p - c, axis=1, keepdims=True), atol=2e-6)


def test_rotation_preserves_shape():
    b = SceneBuilder()
    m = b.add_material(Material.diffuse(0.5, 0.5, 0.5))
    b.add_box((0.0, 0.0, 0.0), (1.0, 2.0, 3.0), m, rot_deg=(30.0, 45.0, 60.0))
    sc = b.build()
    v = sc.tris[:, [0, 1, 2, 4, 5, 6, 8, 9, 10]].reshape(-1, 3)
    r = np.linalg.norm(v, axis=1)
    assert np.allclose(r, np.sqrt(1 + 4 + 9), atol=1e-5)


def test_light_cdf_and_power():
    w = workloads.cornell_small()
    sc = w.scene
    assert sc.nLights == 2
    L = sc.lights
    # power = lum(E*s) * area * pi; the reference's float CDF ends at exactly 1
    e = L[:, 12:15].astype(np.float64)
    lum = 0.299 * e[:, 0] + 0.587 * e[:, 1] + 0.114 * e[:, 2]
    v0, v1, v2 = L[:, 0:3].astype(np.float64), L[:, 4:7].astype(np.float64), L[:, 8:11].astype(np.float64)
    area = 0.5 * np.linalg.norm(np.cross(v1 - v0, v2 - v0), axis=1)
    assert np.allclose(L[:, 3], area, rtol=1e-6)
    assert np.allclose(L[:, 15], lum * area * np.pi, rtol=1e-5)
    assert np.isclose(sc.totalLightPower, L[:, 15].astype(np.float64).sum(), rtol=1e-6)
    assert L[-1, 19] == np.float32(1.0)
    assert np.all(np.diff(L[:, 19]) >= 0)
    # 3x3 ceiling light at radiance 15: total power = 15 * 9 * pi
    assert np.isclose(sc.totalLightPower, 15.0 * 9.0 * np.pi, rtol=1e-5)


def test_full_frame_scene_shape():
    w = workloads.full_frame()
    sc = w.scene
    assert sc.nTris == 250_000
    assert sc.nLights == 16
    check_bvh(sc)
    mats = sc.mats
    kinds = np.where(mats[:, 10] > 0.5, 2, np.where(mats[:, 3] > 0.5, 1, 0))[7:]  # skip room + light
    frac = np.bincount(kinds, minlength=3) / len(kinds)
    assert abs(frac[0] - 0.5) < 0.1 and abs(frac[1] - 0.3) < 0.1 and abs(frac[2] - 0.2) < 0.1
