"""CPU checks that pin the SPPM oracle (rc_oracle_sppm.c): photon-seed and
cell-hash known answers, the radius schedule's closed form, caustics-only
storage, the energy bound, hash-list structure, and the gather's density
estimate against an analytic Epanechnikov sum."""
import ctypes as C

import numpy as np
import pytest

from oracle import rco
from retrocam_amd import workloads
from retrocam_amd.scene import CameraSpec, Material, SceneBuilder

M = 0xFFFFFFFF


def py_pcg(s):
    s = (s * 747796405 + 2891336453) & M
    return s


def py_photon_seed(idx, seed):
    # photon_trace.comp:68-72, Python integers masked to 32 bits
    s = (idx * 1664525 + (seed & M) * 1013904223 + 2531011) & M
    return py_pcg(py_pcg(s))


def py_cell_hash(cx, cy, cz, size):
    h = ((cx * 73856093) & M) ^ ((cy * 19349663) & M) ^ ((cz * 83492791) & M)
    return h & (size - 1)


def test_photon_seed_known_answers():
    L = rco.lib()
    L.rco_photon_seed.restype = C.c_uint32
    L.rco_photon_seed.argtypes = [C.c_uint32, C.c_int32]
    for it in (0, 1, 63):
        seed = it * 7919 + 6271  # SPPMManager.java:88
        for idx in (0, 1, 999_999, 1_999_999, 123457):
            assert L.rco_photon_seed(idx, seed) == py_photon_seed(idx, seed)


def test_cell_hash_known_answers():
    L = rco.lib()
    L.rco_cell_hash.restype = C.c_uint32
    L.rco_cell_hash.argtypes = [C.c_int32] * 4
    rng = np.random.default_rng(1)
    for _ in range(500):
        c = [int(v) for v in rng.integers(-100000, 100000, 3)]
        assert L.rco_cell_hash(*c, 1 << 22) == py_cell_hash(*c, 1 << 22)
    assert L.rco_cell_hash(-1, -1, -1, 1 << 22) == py_cell_hash(-1, -1, -1, 1 << 22)


def test_radius_schedule_closed_form():
    L = rco.lib()
    L.rco_sppm_next_radius.restype = C.c_float
    L.rco_sppm_next_radius.argtypes = [C.c_float, C.c_int32, C.c_float]
    r, r0, a = np.float32(0.02), 0.02, np.float32(0.7)
    prod = 1.0
    for i in range(64):
        r = np.float32(L.rco_sppm_next_radius(float(r), i, float(a)))
        prod *= np.sqrt((i + float(a)) / (i + 1.0))
        assert abs(float(r) - r0 * prod) <= 1e-6 * r0 * prod * (i + 1)


def _trace(sc, photons=20000, radius=0.05, seed=6271, hash_size=1 << 16):
    os_ = rco.OracleScene(sc)
    ph = np.zeros((photons, 12), np.float32)
    heads = np.zeros(hash_size, np.int32)
    nxt = np.zeros(photons, np.int32)
    stored = C.c_int32()
    rc = rco.lib().rco_photon_trace(C.byref(os_.c), photons, radius, seed, hash_size, rco._fp(ph),
                                    heads.ctypes.data_as(C.POINTER(C.c_int32)),
                                    nxt.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(stored))
    assert rc == 0
    return ph, heads, nxt, stored.value


def _bind():
    return rco.lib()  # the SPPM entry points are declared by rco.load


def check_hash_structure(ph, heads, nxt, stored, radius, hash_size):
    n = min(stored, ph.shape[0])
    seen = np.zeros(n, np.int32)
    for h in np.nonzero(heads >= 0)[0]:
        idx = heads[h]
        steps = 0
        while idx >= 0:
            assert idx < n
            seen[idx] += 1
            c = np.floor(ph[idx, 0:3] / np.float32(radius)).astype(np.int64)
            assert py_cell_hash(int(c[0]), int(c[1]), int(c[2]), hash_size) == h
            idx = nxt[idx]
            steps += 1
            assert steps <= n
    assert np.all(seen == 1), "every stored photon is in exactly one list"


def test_caustic_photons_structure_and_energy():
    _bind()
    w = workloads.sppm_caustics(width=64, height=36)
    ph, heads, nxt, stored = _trace(w.scene, photons=40000, radius=0.05)
    assert 0 < stored < 40000
    check_hash_structure(ph, heads, nxt, stored, 0.05, 1 << 16)
    lum = 0.299 * ph[:stored, 4] + 0.587 * ph[:stored, 5] + 0.114 * ph[:stored, 6]
    assert lum.sum() <= w.scene.totalLightPower * (1 + 1e-5) * stored / 40000 + 1e-3
    assert np.all(lum <= w.scene.totalLightPower / 40000 * (1 + 1e-5))


def test_no_specular_surfaces_store_no_photons():
    _bind()
    w = workloads.cornell_small()
    _, heads, _, stored = _trace(w.scene, photons=20000)
    assert stored == 0 and np.all(heads == -1)


def test_gather_matches_analytic_epanechnikov_sum():
    # One diffuse floor quad (albedo a) seen by a pinhole looking straight down;
    # hand-made photons around the hit point of pixel (0, 0). The gather must
    # return a * sum_i (1 - d_i^2/r^2) Phi_i / (pi r^2) over photons within r.
    L = _bind()
    a = 0.6
    b = SceneBuilder()
    fm = b.add_material(Material.diffuse(a, a, a))
    b.add_quad((-50, 0, -50), (-50, 0, 50), (50, 0, 50), (50, 0, -50), fm)
    sc = b.build()
    cam = CameraSpec(eye=(0.3, 2.0, -0.4), forward=(0.0, -1.0, 0.0), right=(1.0, 0.0, 0.0), up=(0.0, 0.0, -1.0),
                     width=1, height=1, focalLengthMm=35.0, pinhole=True)
    hit = np.array([0.3, 0.0, -0.4])
    r = 0.05
    rng = np.random.default_rng(3)
    n = 300
    pos = hit + rng.uniform(-0.07, 0.07, (n, 3)) * np.array([1, 0, 1])
    power = rng.uniform(0.01, 0.1, (n, 3))
    hs = 1 << 12
    ph = np.zeros((n, 12), np.float32)
    ph[:, 0:3] = pos
    ph[:, 4:7] = power
    heads = np.full(hs, -1, np.int32)
    nxt = np.full(n, -1, np.int32)
    for i in range(n):
        c = np.floor(ph[i, 0:3] / np.float32(r)).astype(np.int64)
        h = py_cell_hash(int(c[0]), int(c[1]), int(c[2]), hs)
        nxt[i] = heads[h]
        heads[h] = i
    acc = np.zeros((1, 4), np.float32)
    os_ = rco.OracleScene(sc)
    camc = cam.to_c()
    assert L.rco_photon_gather(C.byref(os_.c), C.byref(camc), r, hs, rco._fp(ph),
                               heads.ctypes.data_as(C.POINTER(C.c_int32)), nxt.ctypes.data_as(C.POINTER(C.c_int32)),
                               rco._fp(acc), None) == 0
    d2 = np.sum((ph[:, 0:3].astype(np.float64) - hit) ** 2, axis=1)
    w_ = np.where(d2 <= r * r, 1 - d2 / (r * r), 0.0)
    # the 27 cells around the hit hash to distinct buckets here, so no photon counts twice
    cc = np.floor(hit / r).astype(np.int64)
    bk = {py_cell_hash(int(cc[0] + dx), int(cc[1] + dy), int(cc[2] + dz), hs)
          for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)}
    assert len(bk) == 27
    exp = a * (w_[:, None] * power).sum(0) / (np.pi * r * r)
    assert np.allclose(acc[0, :3], exp, rtol=2e-5)
