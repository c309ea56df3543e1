"""GPU parity at the full sizes of the benchmark configurations, through the C
ABI against the CPU oracle:

- SPPM caustics at its real size (854x480, 2,000,000 photons per iteration,
  2^22 hash cells, initial radius 0.02; PhotonMap.java:40-44,
  SPPMManager.java:76-101): photon multiset and hash-list structure bit-exact
  for two iterations, the gather bit-exact on the GPU's own map, and the
  per-spp frame loop within relMSE 1e-6 (only the photon-list summation order,
  which follows the atomic insertion order, differs);
- the full plain frame's post chain at 854x480 (a-trous x5, the 14 VHS passes,
  display) within 1e-5 absolute of the oracle, where the step-16 a-trous taps and
  p08's smear reach are not dominated by the image borders;
- the 32-request serving batch over 4 scenes of 50k-250k triangles: every view
  bit-identical to a single-view render of the same request, and one view per
  scene against the oracle on a 1/64 pixel subset (primary-hit triangle IDs
  bit-exact, radiance relMSE <= 1e-3);
- the U-Net at 8 x 128 x 128 (4096 32-pixel tiles at the finest level, so the
  single-group convolution kernels run) against the FP64 oracle.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import rc_oracle_unet as U
from oracle import rco
from retrocam_amd import _native as N
from retrocam_amd import workloads
from retrocam_amd.batch import ViewBatch, ViewRequest
from retrocam_amd.denoiser import NetDenoiser, rel_mse
from retrocam_amd.post import VHS_PASSES, PostProcessStack, TemporalInputs, make_params
from retrocam_amd.renderer import RenderContext, Renderer, SceneUploader
from retrocam_amd.sppm import SPPMManager
from test_gpu_sppm import sorted_records
from test_gpu_unet import oracle_features, random_raw
from test_oracle_sppm import _bind, _trace, check_hash_structure

pytestmark = pytest.mark.gpu

PHOTONS = 2_000_000
HASH = 1 << 22
R0 = 0.02


def _i32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


@pytest.fixture(scope="module")
def c3():
    w = workloads.sppm_caustics()
    assert (w.camera.width, w.camera.height) == (854, 480)
    return w


def test_c3_photon_map_matches_oracle(gpu_ctx, c3):
    _bind()
    SceneUploader(gpu_ctx).upload(c3.scene)
    sp = SPPMManager(gpu_ctx, init_radius=R0)
    for it in range(2):
        sp.trace_photons(PHOTONS, HASH)
        ph, heads, nxt, cnt = sp.read_map()
        r = sp.search_radius
        oph, _, _, ostored = _trace(c3.scene, photons=PHOTONS, radius=r, seed=it * 7919 + 6271, hash_size=HASH)
        print(f"C3 iteration {it}: {cnt} photons stored (oracle {ostored})", flush=True)
        assert cnt == ostored > 0
        assert np.array_equal(sorted_records(ph, cnt), sorted_records(oph, ostored))
        check_hash_structure(ph, heads, nxt, cnt, r, HASH)
        sp.update_radius(0.7)


def test_c3_gather_bit_exact_on_the_same_map(gpu_ctx, c3):
    L = _bind()
    SceneUploader(gpu_ctx).upload(c3.scene)
    r = Renderer(gpu_ctx)
    r.reset()
    r.render(c3.camera, c3.params)
    before = r.read(N.RC_BUF_ACCUM).copy()
    sp = SPPMManager(gpu_ctx, init_radius=R0)
    sp.trace_photons(PHOTONS, HASH)
    sp.gather_radiance(c3.camera)
    after = r.read(N.RC_BUF_ACCUM)
    ph, heads, nxt, cnt = sp.read_map()
    acc = before.copy()
    os_ = rco.OracleScene(c3.scene)
    camc = c3.camera.to_c()
    assert L.rco_photon_gather(C.byref(os_.c), C.byref(camc), sp.search_radius, HASH, rco._fp(ph), _i32p(heads),
                               _i32p(nxt), rco._fp(acc), None) == 0
    assert np.array_equal(after, acc)
    assert np.count_nonzero(np.any(after[:, :3] != before[:, :3], axis=1)) > 100


def test_c3_frame_loop_matches_oracle(gpu_ctx, c3):
    # RenderPipeline.accumulateFrame order (RenderPipeline.java:209-226): per spp
    # trace photons, render one pass, gather, update the radius.
    L = _bind()
    SceneUploader(gpu_ctx).upload(c3.scene)
    r = Renderer(gpu_ctx)
    r.reset()
    sp = SPPMManager(gpu_ctx, init_radius=R0)
    o = rco.OracleRenderer(c3.scene, c3.camera.to_c(), c3.params.to_c())
    camc = c3.camera.to_c()
    orad = np.float32(R0)
    L.rco_sppm_next_radius.restype = C.c_float
    L.rco_sppm_next_radius.argtypes = [C.c_float, C.c_int32, C.c_float]
    for it in range(3):
        sp.trace_photons(PHOTONS, HASH)
        r.render(c3.camera, c3.params)
        sp.gather_radiance(c3.camera)
        sp.update_radius(0.7)
        oph, oheads, onxt, _ = _trace(c3.scene, photons=PHOTONS, radius=float(orad), seed=it * 7919 + 6271,
                                      hash_size=HASH)
        o.render_pass()
        assert L.rco_photon_gather(C.byref(o.scene.c), C.byref(camc), float(orad), HASH, rco._fp(oph), _i32p(oheads),
                                   _i32p(onxt), rco._fp(o.accum), None) == 0
        orad = np.float32(L.rco_sppm_next_radius(float(orad), it, 0.7))
        print(f"C3 frame loop iteration {it} done", flush=True)
    assert sp.search_radius == float(orad)
    g = r.read(N.RC_BUF_ACCUM)
    gm = g[:, :3] / g[:, 3:4]
    om = o.accum[:, :3] / o.accum[:, 3:4]
    rel = float(np.mean((gm - om) ** 2 / (om * om + 0.01)))
    assert rel <= 1e-6, rel


def test_c2_post_chain_at_full_size(gpu_ctx, full_frame_scene):
    w = full_frame_scene
    assert (w.camera.width, w.camera.height) == (854, 480)
    SceneUploader(gpu_ctx).upload(w.scene)
    r = Renderer(gpu_ctx)
    r.reset()
    r.render(w.camera, w.params, n_samples=4)
    acc = r.read(N.RC_BUF_ACCUM)
    gb = r.read(N.RC_BUF_GBUFFER)
    ga = r.read(N.RC_BUF_GALBEDO)
    var = r.read(N.RC_BUF_VARIANCE)
    ps = PostProcessStack(gpu_ctx)
    W, H = w.camera.width, w.camera.height
    # the full plain frame's chain (a-trous x5 + every VHS pass) at two time
    # inputs, and a long smear with strong tracking and dropout events
    cases = [
        (make_params(enable=VHS_PASSES, atrous=True, iterations=5), TemporalInputs(0.0, 1.0, 0)),
        (make_params(enable=VHS_PASSES, atrous=True, iterations=5), TemporalInputs(7.3, 1.6, 219)),
        (make_params(enable=VHS_PASSES, atrous=True, iterations=5, ccdSmearLength=150, ccdSmearThreshold=0.3,
                     trackingSeverity=0.9, dropoutBurstRate=1.0, tapeAge=0.9, headSwitchLines=40),
         TemporalInputs(3.1, 2.0, 902)),
    ]
    for p, t in cases:
        print(f"C2 post chain frameIndex {t.frameIndex}", flush=True)
        g = ps.run_on_accum(p, t)
        hdr, rgba8 = rco.post_chain(acc, W, H, 4, p, t.to_c(), gb, ga, var)
        d = np.abs(g.hdr.astype(np.float64) - hdr.astype(np.float64))
        assert np.max(d) <= 1e-5, np.max(d)
        d8 = np.abs(g.rgba8.astype(np.int32) - rgba8.astype(np.int32))
        assert d8.max() <= 1 and np.mean(d8 == 0) >= 0.999


BUFS = [N.RC_BUF_ACCUM, N.RC_BUF_GBUFFER, N.RC_BUF_GALBEDO, N.RC_BUF_VARIANCE, N.RC_BUF_PRIMARY_ID]


def test_c5_serving_batch_views_equal_single_renders(gpu_ctx):
    scenes, raw = workloads.serving_batch()
    reqs = [ViewRequest(slot, cam, params, spp) for slot, cam, params, spp in raw]
    assert len(reqs) == 32 and len(scenes) == 4
    Renderer(gpu_ctx).set_pt_mode(N.RC_PT_LANES)
    vb = ViewBatch(gpu_ctx)
    for k, sc in enumerate(scenes):
        vb.upload_scene(k, sc)
    vb.render(reqs, reset=True)
    got = {i: {b: vb.read(i, b) for b in BUFS} for i in range(len(reqs))}
    print("C5 batch rendered", flush=True)
    with RenderContext(0) as ctx:
        r = Renderer(ctx)
        r.set_pt_mode(N.RC_PT_LANES)
        for k, sc in enumerate(scenes):
            SceneUploader(ctx).upload(sc)
            for i, q in enumerate(reqs):
                if q.scene != k:
                    continue
                r.reset()
                r.render(q.camera, q.params, n_samples=q.spp)
                for b in BUFS:
                    assert np.array_equal(got[i][b], r.read(b)), (i, b)
            print(f"C5 scene {k}: single-view renders equal", flush=True)
    # one view per scene (its smallest sample count) against the oracle on every 8th pixel
    for k, sc in enumerate(scenes):
        i = min((i for i, q in enumerate(reqs) if q.scene == k), key=lambda i: reqs[i].spp)
        q = reqs[i]
        o = rco.OracleRenderer(sc, q.camera.to_c(), q.params.to_c(), stride=8)
        for _ in range(q.spp):
            o.render_pass()
        m = o.mask()
        assert np.array_equal(got[i][N.RC_BUF_PRIMARY_ID][m], o.primId[m]), (k, i)
        ga, oa = got[i][N.RC_BUF_ACCUM][m], o.accum[m]
        gm, om = ga[:, :3] / ga[:, 3:4], oa[:, :3] / oa[:, 3:4]
        rel = float(np.mean((gm - om) ** 2 / (om * om + 0.01)))
        exact = float(np.mean(np.all(ga.view(np.uint32) == oa.view(np.uint32), axis=1)))
        print(f"C5 scene {k} view {i} ({q.camera.width}x{q.camera.height}, {q.spp} spp): relMSE {rel:.2e}, "
              f"bit-identical {exact:.4f}", flush=True)
        assert rel <= 1e-3 and exact >= 0.999, (k, rel, exact)


def test_unet_at_4096_finest_tiles_matches_oracle(gpu_ctx):
    net = NetDenoiser(gpu_ctx).init(13)
    theta = net.get_params().astype(np.float64)
    raw, tgt = random_raw(5, 8, 128, 128)
    assert 8 * 128 * 128 // 32 == 4096
    feats = oracle_features(raw)
    out = net.forward(raw)
    ref, _ = U.forward(theta, feats, bf=True)
    assert rel_mse(out, np.maximum(ref, 0.0)) <= 1e-4
    loss, g = net.grad(raw, tgt)
    rloss, rg = U.loss_and_grad(theta, feats, tgt.astype(np.float64), bf=True)
    assert abs(loss - rloss) <= 1e-3 * abs(rloss)
    off = 0
    for name, ws, bs in U.layer_shapes():
        n = int(np.prod(ws)) + bs[0]
        a, b = g[off:off + n].astype(np.float64), rg[off:off + n]
        off += n
        cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30))
        rel = float(np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30))
        assert cos >= 0.999 and rel <= 3e-2, (name, cos, rel)
