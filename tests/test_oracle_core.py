"""CPU tests that pin the oracle: PCG known answers, rc_fmath accuracy, binary16
conversion, BVH traversal semantics (brute force, tie order, stack caps)."""
import json
import os

import numpy as np
import pytest

from oracle import rco
from retrocam_amd import workloads
from retrocam_amd.scene import Material, SceneArrays, SceneBuilder

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---- PCG (pathtrace.comp:80-98) -------------------------------------------
def test_pcg_golden_vectors():
    import ctypes as C
    data = json.load(open(os.path.join(GOLDEN, "pcg_vectors.json")))
    L = rco.lib()
    for case in data["cases"]:
        s = L.rco_init_seed(case["x"], case["y"], case["frame"])
        assert s == case["seed"]
        st = C.c_uint32(s)
        got = [L.rco_pcg(C.byref(st)) for _ in case["pcg"]]
        assert got == case["pcg"]
        st = C.c_uint32(s)
        rnd = [L.rco_rand(C.byref(st)) for _ in case["rand"]]
        assert np.array_equal(np.float32(rnd), np.float32(case["rand"]))


def test_rand_can_return_one():
    # float(0xFFFFFFFF) rounds to 2^32, so rand() == 1.0 exactly for the top outputs.
    assert np.float32(np.float32(0xFFFFFFFF) / np.float32(4294967296.0)) == np.float32(1.0)
    assert np.float32(np.float32(0xFFFFFF7F) / np.float32(4294967296.0)) < np.float32(1.0)


# ---- rc_fmath accuracy against float64 ------------------------------------
def _ulp(r):
    return np.spacing(np.abs(r).astype(np.float32)).astype(np.float64)


def test_sin_cos_accuracy():
    x = np.concatenate([np.linspace(-2000, 2000, 400001), np.linspace(-1e5, 1e5, 400001)]).astype(np.float32)
    x = np.concatenate([x, np.float32([0.0, -0.0, 1e-30, np.pi / 2, np.pi, 2 * np.pi])])
    for fn, ref in ((0, np.sin), (1, np.cos)):
        y = rco.fmath(fn, x).astype(np.float64)
        r = ref(x.astype(np.float64))
        err = np.abs(y - r)
        assert np.all(err <= 2.0 * _ulp(r) + 2.0 ** -24), (fn, err.max())


def test_exp_log_pow_accuracy():
    x = np.linspace(-87, 88, 200001).astype(np.float32)
    y = rco.fmath(2, x).astype(np.float64)
    r = np.exp(x.astype(np.float64))
    assert np.all(np.abs(y - r) <= 2.0 * _ulp(r))
    x = np.exp(np.linspace(-87, 88, 200001)).astype(np.float32)
    y = rco.fmath(3, x).astype(np.float64)
    r = np.log(x.astype(np.float64))
    assert np.all(np.abs(y - r) <= 2.0 * _ulp(r) + 2.0 ** -24)
    assert rco.fmath(3, np.float32([0.0]))[0] == -np.inf
    assert np.isnan(rco.fmath(3, np.float32([-1.0]))[0])
    assert rco.fmath(2, np.float32([-200.0]))[0] == 0.0
    assert rco.fmath(2, np.float32([100.0]))[0] == np.inf
    xs = np.random.default_rng(0).uniform(0, 1, 100000).astype(np.float32)
    y = rco.fmath(4, xs, np.full_like(xs, 5.0)).astype(np.float64)
    r = xs.astype(np.float64) ** 5
    assert np.all(np.abs(y - r) <= 1e-5 * r + 1e-37)
    # GLSL pow special cases used by the shaders: pow(0, y>0) = 0
    assert rco.fmath(4, np.float32([0.0]), np.float32([5.0]))[0] == 0.0


def test_f16_round_trip_matches_numpy():
    rng = np.random.default_rng(3)
    v = np.concatenate([
        rng.standard_normal(100000).astype(np.float32) * np.float32(10.0) ** rng.integers(-9, 6, 100000).astype(np.float32),
        np.float32([0.0, -0.0, 65504.0, 65519.99, 65520.0, 1e5, -1e5, 2.0 ** -24, 2.0 ** -25, 2.0 ** -25 * 1.0000001,
                    6.1e-5, 5.96e-8, np.inf, -np.inf, 1e4, 1.0, 0.1])])
    got = rco.fmath(5, v)
    with np.errstate(over="ignore"):
        ref = v.astype(np.float16).astype(np.float32)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    nan = rco.fmath(5, np.float32([np.nan]))
    assert np.isnan(nan[0])


# ---- traversal: brute force, tie order, stack caps (pathtrace.comp:142-242)
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_closest_hit_equals_brute_force(seed):
    sc = workloads.tiny_random(seed, n_tris=64)
    os_ = rco.OracleScene(sc)
    rng = np.random.default_rng(100 + seed)
    hits = 0
    cen = (sc.tris[:, 0:3] + sc.tris[:, 4:7] + sc.tris[:, 8:11]) / 3.0
    for i in range(3000):
        ro = rng.uniform(-4, 4, 3).astype(np.float32)
        d = rng.standard_normal(3) if i % 2 else cen[rng.integers(0, sc.nTris)] - ro + 0.05 * rng.standard_normal(3)
        rd = (d / np.linalg.norm(d)).astype(np.float32)
        tri, t, _ = rco.closest_hit(os_, ro, rd)
        btri, bt = rco.brute_closest(os_, ro, rd)
        assert tri == btri
        if tri >= 0:
            assert np.float32(t) == np.float32(bt)
            hits += 1
    assert hits > 300


@pytest.mark.parametrize("seed", [4, 5])
def test_shadow_ray_equals_brute_force(seed):
    sc = workloads.tiny_random(seed, n_tris=64)
    os_ = rco.OracleScene(sc)
    rng = np.random.default_rng(200 + seed)
    blocked = 0
    cen = (sc.tris[:, 0:3] + sc.tris[:, 4:7] + sc.tris[:, 8:11]) / 3.0
    for i in range(3000):
        ro = rng.uniform(-4, 4, 3).astype(np.float32)
        d = rng.standard_normal(3) if i % 2 else cen[rng.integers(0, sc.nTris)] - ro + 0.05 * rng.standard_normal(3)
        rd = (d / np.linalg.norm(d)).astype(np.float32)
        md = float(rng.uniform(0.1, 6.0))
        tri, bt = rco.brute_closest(os_, ro, rd)
        expect_free = not (tri >= 0 and bt < np.float32(md))
        assert rco.shadow_ray(os_, ro, rd, md) == expect_free
        blocked += not expect_free
    assert blocked > 200


def _tri(z, x0=-50.0, x1=50.0, y0=-50.0, y1=50.0):
    t = np.zeros(32, np.float32)
    t[0:3] = (x0, y0, z)
    t[4:7] = (x1, y0, z)
    t[8:11] = (0.0, y1, z)
    t[12:15] = t[16:19] = t[20:23] = (0.0, 0.0, 1.0)
    return t


def _node(box, left=-1, right=-1, start=0, count=0):
    n = np.zeros(12, np.float32)
    n[0:3] = box[0]
    n[4:7] = box[1]
    n[8:12] = np.array([left, right, start, count], np.int32).view(np.float32)
    return n


def chain_scene(M: int, leaf_on_left: bool, occluder_only_deepest: bool = False) -> SceneArrays:
    """A degenerate BVH: interiors I_0..I_{M-1} in a chain, every node with the
    same box. leaf_on_left=False: I_j = (left L_j, right I_{j+1}); True:
    I_j = (left I_{j+1}, right L_j). Leaf L_k holds triangle k; triangle k lies in
    the plane z = -(10 + (M - k)), so deeper leaves are nearer to a ray from
    the origin towards -z."""
    box = ((-100.0, -100.0, -200.0), (100.0, 100.0, 5.0))
    K = M + 1
    tris = []
    for k in range(K):
        z = -(10.0 + (M - k))
        if occluder_only_deepest and k != M:
            tris.append(_tri(z, 60.0, 80.0, 60.0, 80.0))  # off the ray
        else:
            tris.append(_tri(z))
    nodes = {}
    order = []

    def emit(kind, j):
        idx = len(order)
        order.append((kind, j))
        return idx

    # DFS pre-order numbering (left subtree first) as BVHBuilder emits it.
    def build(j):
        me = emit("I", j)
        if j == M - 1:
            kids = [("L", j), ("L", M)] if not leaf_on_left else [("L", M), ("L", j)]
        else:
            kids = [("L", j), ("I", j + 1)] if not leaf_on_left else [("I", j + 1), ("L", j)]
        idx = []
        for kind, kj in kids:
            if kind == "L":
                idx.append(emit("L", kj))
            else:
                idx.append(build(kj))
        nodes[me] = idx
        return me

    build(0)
    arr = np.zeros((len(order), 12), np.float32)
    for i, (kind, j) in enumerate(order):
        if kind == "L":
            arr[i] = _node(box, start=j, count=1)
        else:
            arr[i] = _node(box, left=nodes[i][0], right=nodes[i][1])
    mats = np.stack([Material.diffuse(0.5, 0.5, 0.5).write()])
    t = np.stack(tris)
    t[:, 30] = np.zeros(K, np.int32).view(np.float32)
    return SceneArrays(t, arr, mats, np.zeros((0, 20), np.float32), 0.0, M, 1, K)


def test_closest_hit_stack_cap_drops_deep_subtrees():
    # Pop of I_j happens with top == j, so I_0..I_61 expand and I_62 is dropped
    # (pathtrace.comp:176): the nearest visited triangle is L_61's.
    ro = np.float32([0.0, 0.0, 1.0])
    d = np.array([0.01, 0.02, -1.0])
    rd = (d / np.linalg.norm(d)).astype(np.float32)
    sc = chain_scene(80, leaf_on_left=False)
    os_ = rco.OracleScene(sc)
    tri, _, _ = rco.closest_hit(os_, ro, rd)
    assert tri == 61
    assert rco.brute_closest(os_, ro, rd)[0] == 80
    # Below the cap every leaf is reached: the deepest (nearest) triangle wins.
    sc = chain_scene(40, leaf_on_left=False)
    assert rco.closest_hit(rco.OracleScene(sc), ro, rd)[0] == 40


def test_shadow_stack_cap_leaks_light():
    # Shadow walk pushes right then left (pathtrace.comp:236-237); with the
    # interior chain on the left, I_j is popped with top == j and I_30 is not
    # expanded (cap 30, :235), so an occluder only in the deepest leaf is missed.
    ro = np.float32([0.0, 0.0, 1.0])
    d = np.array([0.01, 0.02, -1.0])
    rd = (d / np.linalg.norm(d)).astype(np.float32)
    deep = rco.OracleScene(chain_scene(40, leaf_on_left=True, occluder_only_deepest=True))
    assert rco.brute_closest(deep, ro, rd)[0] == 40
    assert rco.shadow_ray(deep, ro, rd, 1000.0) is True
    shallow = rco.OracleScene(chain_scene(20, leaf_on_left=True, occluder_only_deepest=True))
    assert rco.shadow_ray(shallow, ro, rd, 1000.0) is False


def test_equal_child_distances_pop_right_first():
    # Ties push left then right (pathtrace.comp:192-194) so the right child is
    # visited first; with two coincident triangles (equal t) in different leaves
    # the right leaf's triangle is accepted first and the left one (not strictly
    # nearer) is rejected.
    box = ((-100.0, -100.0, -200.0), (100.0, 100.0, 5.0))
    tris = np.stack([_tri(-10.0), _tri(-10.0)])
    tris[:, 30] = np.zeros(2, np.int32).view(np.float32)
    nodes = np.stack([_node(box, left=1, right=2), _node(box, start=0, count=1), _node(box, start=1, count=1)])
    mats = np.stack([Material.diffuse(0.5, 0.5, 0.5).write()])
    sc = SceneArrays(tris, nodes, mats, np.zeros((0, 20), np.float32), 0.0, 1, 1, 2)
    ro = np.float32([0.0, 0.0, 1.0])
    rd = np.float32([0.0, 0.0, -1.0])
    assert rco.closest_hit(rco.OracleScene(sc), ro, rd)[0] == 1
    assert rco.brute_closest(rco.OracleScene(sc), ro, rd)[0] == 0


def test_oracle_stride_subset_matches_full():
    w = workloads.cornell_small(width=64, height=36)
    full = rco.OracleRenderer(w.scene, w.camera.to_c(), w.params.to_c(), stride=1)
    sub = rco.OracleRenderer(w.scene, w.camera.to_c(), w.params.to_c(), stride=4)
    for _ in range(2):
        full.render_pass()
        sub.render_pass()
    m = sub.mask()
    assert np.array_equal(full.accum[m], sub.accum[m])
    assert np.array_equal(full.primId[m], sub.primId[m])
