"""GPU parity of SPPM (photon trace/insert, gather, radius schedule) through
the C ABI against the CPU oracle.

The photon set is deterministic (each photon's path depends only on its index
and the iteration seed) but its slot and list order follow the atomic order, so
photon records are compared as multisets (bit-exact), the hash lists by their
structure, and the gather is checked bit-exactly by running the oracle's gather
on the GPU's own photon map."""
import ctypes as C

import numpy as np
import pytest

from oracle import rco
from retrocam_amd import _native as N
from retrocam_amd import workloads
from retrocam_amd.renderer import Renderer, SceneUploader
from retrocam_amd.sppm import SPPMManager
from test_oracle_sppm import _bind, _trace, check_hash_structure

pytestmark = pytest.mark.gpu


def sorted_records(ph, n):
    recs = ph[:n].view(np.uint32)
    order = np.lexsort(recs.T[::-1])
    return recs[order]


@pytest.fixture(scope="module")
def caustics():
    return workloads.sppm_caustics(width=160, height=90, spp=3, photons=200_000, hash_size=1 << 18)


def test_photon_map_matches_oracle(gpu_ctx, caustics):
    w = caustics
    _bind()
    SceneUploader(gpu_ctx).upload(w.scene)
    sp = SPPMManager(gpu_ctx, init_radius=0.05)
    for it in range(2):
        sp.trace_photons(200_000, 1 << 18)
        ph, heads, nxt, cnt = sp.read_map()
        r = sp.search_radius
        oph, oheads, onxt, ostored = _trace(w.scene, photons=200_000, radius=r, seed=it * 7919 + 6271,
                                            hash_size=1 << 18)
        assert cnt == ostored > 0
        assert np.array_equal(sorted_records(ph, cnt), sorted_records(oph, ostored))
        check_hash_structure(ph, heads, nxt, cnt, r, 1 << 18)
        sp.update_radius(0.7)


def test_gather_bit_exact_on_the_same_map(gpu_ctx, caustics):
    w = caustics
    L = _bind()
    SceneUploader(gpu_ctx).upload(w.scene)
    r = Renderer(gpu_ctx)
    r.reset()
    r.render(w.camera, w.params)
    before = r.read(N.RC_BUF_ACCUM).copy()
    sp = SPPMManager(gpu_ctx, init_radius=0.05)
    sp.trace_photons(200_000, 1 << 18)
    sp.gather_radiance(w.camera)
    after = r.read(N.RC_BUF_ACCUM)
    ph, heads, nxt, cnt = sp.read_map()
    acc = before.copy()
    os_ = rco.OracleScene(w.scene)
    camc = w.camera.to_c()
    assert L.rco_photon_gather(C.byref(os_.c), C.byref(camc), sp.search_radius, 1 << 18, rco._fp(ph),
                               heads.ctypes.data_as(C.POINTER(C.c_int32)), nxt.ctypes.data_as(C.POINTER(C.c_int32)),
                               rco._fp(acc), None) == 0
    assert np.array_equal(after, acc)
    assert np.any(after[:, :3] != before[:, :3])  # caustics were added somewhere


def test_sppm_frame_loop_matches_oracle(gpu_ctx, caustics):
    # RenderPipeline.accumulateFrame order (RenderPipeline.java:209-226):
    # per spp trace photons, render one pass, gather, update the radius.
    w = caustics
    L = _bind()
    SceneUploader(gpu_ctx).upload(w.scene)
    r = Renderer(gpu_ctx)
    r.reset()
    sp = SPPMManager(gpu_ctx, init_radius=0.05)
    o = rco.OracleRenderer(w.scene, w.camera.to_c(), w.params.to_c())
    camc = w.camera.to_c()
    orad = np.float32(0.05)
    L.rco_sppm_next_radius.restype = C.c_float
    L.rco_sppm_next_radius.argtypes = [C.c_float, C.c_int32, C.c_float]
    for it in range(w.spp):
        sp.trace_photons(200_000, 1 << 18)
        r.render(w.camera, w.params)
        sp.gather_radiance(w.camera)
        sp.update_radius(0.7)
        oph, oheads, onxt, _ = _trace(w.scene, photons=200_000, radius=float(orad), seed=it * 7919 + 6271,
                                      hash_size=1 << 18)
        o.render_pass()
        assert L.rco_photon_gather(C.byref(o.scene.c), C.byref(camc), float(orad), 1 << 18, rco._fp(oph),
                                   oheads.ctypes.data_as(C.POINTER(C.c_int32)),
                                   onxt.ctypes.data_as(C.POINTER(C.c_int32)), rco._fp(o.accum), None) == 0
        orad = np.float32(L.rco_sppm_next_radius(float(orad), it, 0.7))
    assert sp.search_radius == float(orad)
    g = r.read(N.RC_BUF_ACCUM)
    gm = g[:, :3] / g[:, 3:4]
    om = o.accum[:, :3] / o.accum[:, 3:4]
    rel = float(np.mean((gm - om) ** 2 / (om * om + 0.01)))
    assert rel <= 1e-6, rel  # only the photon-list summation order differs
