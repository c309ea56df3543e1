"""GPU parity of the path-trace path (libretrocam_hip.so through the C ABI)
against the CPU oracle on the same scenes, seeds and sample counts.

Bars: PCG streams, primary-hit and G-buffer triangle IDs and the binary16
G-buffers bit-exact; per-pixel radiance relMSE <= 1e-3 (the path's stated tolerance),
and in practice bit-identical (both sides run the same IEEE binary32 operation
sequence), which is asserted on >= 99.9% of pixels."""
import json
import os

import numpy as np
import pytest

from oracle import rco
from retrocam_amd import _native as N
from retrocam_amd import workloads
from retrocam_amd.renderer import Renderer, RenderError, SceneUploader, debug_pcg, debug_rays
from retrocam_amd.scene import CameraSpec, Material, PtParams, SceneArrays, SceneBuilder

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rel_mse(x, r):
    return float(np.mean((x - r) ** 2 / (r * r + 0.01)))


MODES = [pytest.param(N.RC_PT_MEGAKERNEL, id="megakernel"), pytest.param(N.RC_PT_WAVEFRONT, id="wavefront"),
         pytest.param(N.RC_PT_LANES, id="lanes")]


def gpu_render(ctx, scene, cam, params, spp, batch=1, fp64=False, mode=N.RC_PT_MEGAKERNEL):
    SceneUploader(ctx).upload(scene)
    r = Renderer(ctx)
    r.set_pt_mode(mode)
    r.set_accum_mode(N.RC_ACCUM_FP64 if fp64 else N.RC_ACCUM_FP32)
    r.reset()
    done = 0
    while done < spp:
        k = min(batch, spp - done)
        r.render(cam, params, n_samples=k)
        done += k
    assert r.get_total_samples() == spp
    out = {b: r.read(getattr(N, "RC_BUF_" + b.upper())) for b in
           ("accum", "gbuffer", "galbedo", "variance", "primary_id", "gbuffer_id")}
    if fp64:
        out["accum64"] = r.read(N.RC_BUF_ACCUM64)
    return out


def oracle_render(scene, cam, params, spp, stride=1, fp64=False):
    o = rco.OracleRenderer(scene, cam.to_c(), params.to_c(), stride=stride, fp64=fp64)
    for _ in range(spp):
        o.render_pass()
    return o


def assert_parity(g, o, mask=None, min_exact=0.999):
    m = np.ones(o.accum.shape[0], bool) if mask is None else mask
    assert np.array_equal(g["primary_id"][m], o.primId[m]), "primary-hit triangle IDs differ"
    assert np.array_equal(g["gbuffer_id"][m], o.gbufId[m]), "G-buffer triangle IDs differ"
    assert np.array_equal(g["gbuffer"][m], o.gbuf[m]), "G-buffer (RGBA16F) differs"
    assert np.array_equal(g["galbedo"][m], o.galbedo[m]), "first-hit albedo differs"
    ga, oa = g["accum"][m], o.accum[m]
    assert np.array_equal(ga[:, 3], oa[:, 3])
    gm = ga[:, :3] / ga[:, 3:4]
    om = oa[:, :3] / oa[:, 3:4]
    assert rel_mse(gm, om) <= 1e-3
    exact = np.mean(np.all(ga.view(np.uint32) == oa.view(np.uint32), axis=1))
    assert exact >= min_exact, f"bit-identical pixels: {exact:.6f}"
    vexact = np.mean(g["variance"][m].view(np.uint32) == o.variance[m].view(np.uint32))
    assert vexact >= min_exact


def test_pcg_streams_bit_exact(gpu_ctx):
    data = json.load(open(os.path.join(GOLDEN, "pcg_vectors.json")))
    xyf = np.array([[c["x"], c["y"], c["frame"]] for c in data["cases"]], np.int32)
    out = debug_pcg(gpu_ctx, xyf, 12)
    for row, c in zip(out, data["cases"]):
        assert int(row[0]) == c["seed"]
        assert [int(v) for v in row[1:]] == c["pcg"]
    rng = np.random.default_rng(7)
    xyf = np.stack([rng.integers(0, 2000, 4096), rng.integers(0, 2000, 4096), rng.integers(0, 1 << 20, 4096)],
                   axis=1).astype(np.int32)
    out = debug_pcg(gpu_ctx, xyf, 4)
    L = rco.lib()
    import ctypes as C
    for i in range(0, 4096, 97):
        s = L.rco_init_seed(*[int(v) for v in xyf[i]])
        assert int(out[i, 0]) == s
        st = C.c_uint32(s)
        assert [int(v) for v in out[i, 1:]] == [L.rco_pcg(C.byref(st)) for _ in range(4)]


@pytest.mark.parametrize("seed", [1, 2])
def test_ray_queries_match_oracle(gpu_ctx, seed):
    sc = workloads.tiny_random(seed, n_tris=200)
    SceneUploader(gpu_ctx).upload(sc)
    os_ = rco.OracleScene(sc)
    rng = np.random.default_rng(seed)
    n = 4000
    ro = rng.uniform(-4, 4, (n, 3)).astype(np.float32)
    d = rng.standard_normal((n, 3))
    rd = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    md = rng.uniform(0.1, 6, n).astype(np.float32)
    tri, t, un = debug_rays(gpu_ctx, ro, rd, md)
    for i in range(n):
        otri, ot, _ = rco.closest_hit(os_, ro[i], rd[i])
        assert tri[i] == otri
        assert np.float32(t[i]) == np.float32(ot)
        assert bool(un[i]) == rco.shadow_ray(os_, ro[i], rd[i], md[i])


def test_stack_caps_match_oracle(gpu_ctx):
    from test_oracle_core import chain_scene
    d = np.array([0.01, 0.02, -1.0])
    rd = (d / np.linalg.norm(d)).astype(np.float32)[None, :]
    ro = np.float32([[0.0, 0.0, 1.0]])
    sc = chain_scene(80, leaf_on_left=False)
    SceneUploader(gpu_ctx).upload(sc)
    tri, _, _ = debug_rays(gpu_ctx, ro, rd, np.float32([1000.0]))
    assert tri[0] == 61 == rco.closest_hit(rco.OracleScene(sc), ro[0], rd[0])[0]
    sc = chain_scene(40, leaf_on_left=True, occluder_only_deepest=True)
    SceneUploader(gpu_ctx).upload(sc)
    _, _, un = debug_rays(gpu_ctx, ro, rd, np.float32([1000.0]))
    assert un[0] == 1 and rco.shadow_ray(rco.OracleScene(sc), ro[0], rd[0], 1000.0)


@pytest.mark.parametrize("mode", MODES)
def test_cornell_small_parity(mode, gpu_ctx, cornell):
    w = cornell
    g = gpu_render(gpu_ctx, w.scene, w.camera, w.params, w.spp, batch=w.spp, mode=mode)
    o = oracle_render(w.scene, w.camera, w.params, w.spp)
    assert_parity(g, o)


@pytest.mark.parametrize("mode", MODES)
def test_thin_lens_sa_lca_parity(mode, gpu_ctx, cornell):
    w = cornell
    cam = CameraSpec(**{**w.camera.__dict__, "pinhole": False, "fStop": 1.4, "blades": 6, "saStrength": 0.04,
                        "focalDist": 9.0, "lcaDelta": (-0.2, 0.0, 0.3), "width": 96, "height": 54})
    g = gpu_render(gpu_ctx, w.scene, cam, w.params, 3, batch=2, mode=mode)
    o = oracle_render(w.scene, cam, w.params, 3)
    assert_parity(g, o)


def test_batched_passes_equal_single_passes(gpu_ctx, cornell):
    w = cornell
    cam = CameraSpec(**{**w.camera.__dict__, "width": 128, "height": 72})
    a = gpu_render(gpu_ctx, w.scene, cam, w.params, 6, batch=1, mode=N.RC_PT_MEGAKERNEL)
    b = gpu_render(gpu_ctx, w.scene, cam, w.params, 6, batch=4, mode=N.RC_PT_WAVEFRONT)
    c = gpu_render(gpu_ctx, w.scene, cam, w.params, 6, batch=3, mode=N.RC_PT_MEGAKERNEL)
    d = gpu_render(gpu_ctx, w.scene, cam, w.params, 6, batch=5, mode=N.RC_PT_LANES)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
        assert np.array_equal(a[k], c[k]), k
        assert np.array_equal(a[k], d[k]), k


def test_lane_tuning_changes_no_result(gpu_ctx, cornell):
    # Threshold, kernel variant and sample-buffer bound only reschedule work:
    # every setting gives the same image.
    w = cornell
    cam = CameraSpec(**{**w.camera.__dict__, "width": 100, "height": 60})
    SceneUploader(gpu_ctx).upload(w.scene)
    r = Renderer(gpu_ctx)
    ref = gpu_render(gpu_ctx, w.scene, cam, w.params, 4, batch=4, mode=N.RC_PT_MEGAKERNEL)
    for th in (1, 17, 64):
        r.set_tuning(laneThreshold=th)
        got = gpu_render(gpu_ctx, w.scene, cam, w.params, 4, batch=4, mode=N.RC_PT_LANES)
        for k in ref:
            assert np.array_equal(ref[k], got[k]), (th, k)
    for v in range(N.RC_LANE_VARIANTS):
        r.set_tuning(laneThreshold=40, laneVariant=v, laneBufferMiB=64)  # 64 MiB: several chunks
        got = gpu_render(gpu_ctx, w.scene, cam, w.params, 4, batch=4, mode=N.RC_PT_LANES)
        for k in ref:
            assert np.array_equal(ref[k], got[k]), (v, k)
    for tw in (4, 16, 32):
        r.set_tuning(laneVariant=7, laneTileWidth=tw)
        got = gpu_render(gpu_ctx, w.scene, cam, w.params, 4, batch=4, mode=N.RC_PT_LANES)
        for k in ref:
            assert np.array_equal(ref[k], got[k]), (tw, k)
    r.set_tuning(laneThreshold=32, laneVariant=9, laneBufferMiB=24576, laneTileWidth=8, ptMode=N.RC_PT_LANES)
    for bad in ({"laneThreshold": 0}, {"laneThreshold": 65}, {"laneVariant": N.RC_LANE_VARIANTS},
                {"laneBufferMiB": 63}, {"laneTileWidth": 2}, {"laneTileWidth": 12}):
        with pytest.raises(RenderError) as e:
            r.set_tuning(**bad)
        assert e.value.code == N.RC_ERR_INVALID_ARG


def test_lanes_chunked_launch(gpu_ctx, cornell):
    # 320x180 needs 0.74 MB of sample buffer per sample: with a 64 MiB bound a
    # 100-sample launch runs as two chunks, each reduced in frame order.
    w = cornell
    ref = gpu_render(gpu_ctx, w.scene, w.camera, w.params, 100, batch=100, mode=N.RC_PT_MEGAKERNEL)
    r = Renderer(gpu_ctx)
    r.set_tuning(laneBufferMiB=64)
    got = gpu_render(gpu_ctx, w.scene, w.camera, w.params, 100, batch=100, mode=N.RC_PT_LANES)
    r.set_tuning(laneBufferMiB=24576)
    for k in ref:
        assert np.array_equal(ref[k], got[k]), k


def test_autotune_keeps_a_valid_tuning(gpu_ctx, cornell):
    w = cornell
    cam = CameraSpec(**{**w.camera.__dict__, "width": 160, "height": 90})
    ref = gpu_render(gpu_ctx, w.scene, cam, w.params, 3, batch=3, mode=N.RC_PT_MEGAKERNEL)
    r = Renderer(gpu_ctx)
    r.set_pt_mode(N.RC_PT_LANES)
    got = r.autotune(cam, w.params, samples=2)
    assert 0 <= got["laneVariant"] < N.RC_LANE_VARIANTS and 1 <= got["laneThreshold"] <= 64
    assert got["laneTileWidth"] in (4, 8, 16, 32) and got["ms_per_sample"] > 0.0
    assert got["ptMode"] in (N.RC_PT_LANES, N.RC_PT_MEGAKERNEL)
    assert r.get_tuning() == {k: got[k] for k in Renderer.TUNING_FIELDS}
    assert r.get_total_samples() == 0
    out = gpu_render(gpu_ctx, w.scene, cam, w.params, 3, batch=3, mode=N.RC_PT_LANES)
    for k in ref:
        assert np.array_equal(ref[k], out[k]), k
    r.set_tuning(laneThreshold=32, laneVariant=9, laneTileWidth=8, ptMode=N.RC_PT_LANES)


def test_lanes_deep_bvh_falls_back(gpu_ctx):
    # A BVH 80 levels deep exceeds the shadow walk's 30-entry cap: RC_PT_LANES
    # must give the megakernel's image (it runs the megakernel for such scenes).
    from test_oracle_core import chain_scene
    sc = chain_scene(80, leaf_on_left=False)
    cam = CameraSpec(eye=(0.0, 0.0, 1.0), forward=(0.0, 0.0, -1.0), right=(1.0, 0.0, 0.0), up=(0.0, 1.0, 0.0),
                     width=24, height=24, pinhole=True)
    p = PtParams(1, 1.0, 4)
    a = gpu_render(gpu_ctx, sc, cam, p, 2, batch=2, mode=N.RC_PT_MEGAKERNEL)
    b = gpu_render(gpu_ctx, sc, cam, p, 2, batch=2, mode=N.RC_PT_LANES)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("mode", MODES)
def test_fp64_accumulation_parity(mode, gpu_ctx, cornell):
    w = cornell
    cam = CameraSpec(**{**w.camera.__dict__, "width": 64, "height": 36})
    g = gpu_render(gpu_ctx, w.scene, cam, w.params, 5, batch=2, fp64=True, mode=mode)
    o = oracle_render(w.scene, cam, w.params, 5, fp64=True)
    assert np.array_equal(g["accum64"], o.accum64)
    assert np.array_equal(g["accum"], o.accum64.astype(np.float32))


@pytest.mark.parametrize("mode", MODES)
def test_full_frame_parity_subset(mode, gpu_ctx, full_frame_scene):
    # 854x480 thin-lens frame over 250k triangles; oracle on every 4th pixel.
    w = full_frame_scene
    g = gpu_render(gpu_ctx, w.scene, w.camera, w.params, 2, batch=2, mode=mode)
    o = oracle_render(w.scene, w.camera, w.params, 2, stride=4)
    assert_parity(g, o, mask=o.mask())


@pytest.mark.parametrize("mode", MODES)
def test_zero_lights_and_small_leaves(mode, gpu_ctx):
    b = SceneBuilder()
    m = b.add_material(Material.diffuse(0.6, 0.5, 0.4))
    g = b.add_material(Material.glass(1.5))
    b.add_quad((-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1), m)
    b.add_quad((-1, 0.5, -1), (-1, 0.5, 1), (1, 0.5, 1), (1, 0.5, -1), g)
    sc = b.build()  # 4 triangles: the root is a leaf; no lights
    assert sc.nNodes == 1 and sc.nLights == 0
    cam = CameraSpec(eye=(0.1, 2.0, 0.2), forward=(0.0, -1.0, 0.0), right=(1.0, 0.0, 0.0), up=(0.0, 0.0, -1.0),
                     width=32, height=32, pinhole=True)
    p = PtParams(1, 1.0, 6)
    gg = gpu_render(gpu_ctx, sc, cam, p, 3, batch=3, mode=mode)
    o = oracle_render(sc, cam, p, 3)
    assert_parity(gg, o)


def test_big_leaf(gpu_ctx):
    # One leaf of 40 triangles (> 31: the device ref stores its count aside).
    sc0 = workloads.tiny_random(9, n_tris=40)
    box_min = np.minimum.reduce([sc0.tris[:, 0:3], sc0.tris[:, 4:7], sc0.tris[:, 8:11]]).min(0)
    box_max = np.maximum.reduce([sc0.tris[:, 0:3], sc0.tris[:, 4:7], sc0.tris[:, 8:11]]).max(0)
    node = np.zeros((1, 12), np.float32)
    node[0, 0:3] = box_min
    node[0, 4:7] = box_max
    node[0, 8:12] = np.array([-1, -1, 0, 40], np.int32).view(np.float32)
    sc = SceneArrays(sc0.tris, node, sc0.mats, sc0.lights, sc0.totalLightPower, 0, 40, 1)
    SceneUploader(gpu_ctx).upload(sc)
    rng = np.random.default_rng(3)
    n = 2000
    ro = rng.uniform(-4, 4, (n, 3)).astype(np.float32)
    d = rng.standard_normal((n, 3))
    rd = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    tri, t, un = debug_rays(gpu_ctx, ro, rd, np.full(n, 5.0, np.float32))
    os_ = rco.OracleScene(sc)
    for i in range(n):
        assert tri[i] == rco.brute_closest(os_, ro[i], rd[i])[0]


def test_invalid_inputs_are_rejected(gpu_ctx, cornell):
    lib = N.hip_lib()
    sc = cornell.scene
    up = SceneUploader(gpu_ctx)
    bad = sc.nodes.copy()
    bad.view(np.int32)[0, 8] = sc.nNodes + 5  # child index out of range
    with pytest.raises(RenderError) as e:
        up.upload(SceneArrays(sc.tris, bad, sc.mats, sc.lights, sc.totalLightPower, 0, 0, 0))
    assert e.value.code == N.RC_ERR_INVALID_ARG
    bad = sc.tris.copy()
    bad[5, 30] = np.array([99], np.int32).view(np.float32)[0]  # material index out of range
    with pytest.raises(RenderError):
        up.upload(SceneArrays(bad, sc.nodes, sc.mats, sc.lights, sc.totalLightPower, 0, 0, 0))
    up.upload(sc)
    r = Renderer(gpu_ctx)
    with pytest.raises(RenderError) as e:
        r.render(CameraSpec(**{**cornell.camera.__dict__, "width": 0}), cornell.params)
    assert e.value.code == N.RC_ERR_INVALID_ARG
    with pytest.raises(RenderError):
        r.render(cornell.camera, PtParams(1, 1.0, 0))
    r.render(cornell.camera, cornell.params)
    buf = np.zeros(10, np.float32)
    import ctypes as C
    assert lib.rc_read(gpu_ctx.h, N.RC_BUF_ACCUM, buf.ctypes.data_as(C.c_void_p), 40) == N.RC_ERR_INVALID_ARG


@pytest.mark.parametrize("mode", [pytest.param(N.RC_PT_LANES, id="lanes"),
                                  pytest.param(N.RC_PT_MEGAKERNEL, id="megakernel")])
def test_unsorted_light_cdf_takes_the_scan(mode, gpu_ctx, cornell):
    # The light pick bisects the CDF only when it is non-decreasing and NaN-free;
    # a host that writes another CDF gets the reference's linear scan. Both the
    # sorted (every other parity test) and the unsorted case must match the
    # oracle, which always scans.
    w = cornell
    lights = w.scene.lights.copy()
    assert lights.shape[0] >= 2
    cdf = lights[:, 19].copy()
    lights[:, 19] = cdf[::-1]  # decreasing: not sortable, picks change
    sc = SceneArrays(w.scene.tris, w.scene.nodes, w.scene.mats, lights, w.scene.totalLightPower,
                     w.scene.maxDepth, w.scene.maxLeafSize, w.scene.nLeaves)
    cam = CameraSpec(**{**w.camera.__dict__, "width": 64, "height": 36})
    g = gpu_render(gpu_ctx, sc, cam, w.params, 3, batch=3, mode=mode)
    o = oracle_render(sc, cam, w.params, 3)
    assert_parity(g, o)
    g2 = gpu_render(gpu_ctx, w.scene, cam, w.params, 3, batch=3, mode=mode)
    assert not np.array_equal(g["accum"], g2["accum"])  # the picks did change


def test_triangle_reciprocal_is_exact(gpu_ctx):
    # The triangle test divides by det with the hardware reciprocal and one FMA
    # Newton step below |det| = 2^126 (IEEE division above): every binary32 det
    # it can see must give the correctly rounded 1/det, as the oracle's division.
    import ctypes as C
    m = C.c_uint64(123)
    gpu_ctx.check(gpu_ctx._lib.rc_debug_recip_check(gpu_ctx.h, C.byref(m)))
    assert m.value == 0
