"""ctypes loader of the CPU oracle (oracle/_build/librc_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker. The product never imports it.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "librc_oracle.so")

_FP = C.POINTER(C.c_float)
_VP = C.c_void_p


class RcoScene(C.Structure):
    _fields_ = [("tris", _FP), ("nodes", _FP), ("mats", _FP), ("lights", _FP),
                ("nTris", C.c_int32), ("nNodes", C.c_int32), ("nMats", C.c_int32), ("nLights", C.c_int32),
                ("totalLightPower", C.c_float)]


class RcoFrame(C.Structure):
    _fields_ = [("accum", _VP), ("variance", _VP), ("gbuf", _VP), ("galbedo", _VP),
                ("primId", _VP), ("gbufId", _VP), ("accum64", _VP)]


STAT_FIELDS = ["paths", "rays", "shadowRays", "nodePops", "childTests", "triTests", "hitUpdates", "shadeHits",
               "shadowPops", "shadowTriTests", "lightScans", "lightReads"]


class RcoStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in STAT_FIELDS]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n in STAT_FIELDS}


def build() -> str:
    r = subprocess.run(["make", "-s", "-C", HERE], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout)
    return LIB_PATH


_lib = None


def load(path: str) -> C.CDLL:
    """Loads an oracle library built from these sources (the default build, or a
    test's variant build) and declares its entry points."""
    from retrocam_amd._native import RcCamera, RcPostParams, RcPtParams, RcTemporal
    L = C.CDLL(path)
    U32P = C.POINTER(C.c_uint32)
    L.rco_pcg.restype = C.c_uint32
    L.rco_pcg.argtypes = [U32P]
    L.rco_init_seed.restype = C.c_uint32
    L.rco_init_seed.argtypes = [C.c_int32, C.c_int32, C.c_int32]
    L.rco_rand.restype = C.c_float
    L.rco_rand.argtypes = [U32P]
    L.rco_closest_hit.restype = C.c_int32
    L.rco_closest_hit.argtypes = [C.POINTER(RcoScene), _FP, _FP, _FP, _FP]
    L.rco_brute_closest.restype = C.c_int32
    L.rco_brute_closest.argtypes = [C.POINTER(RcoScene), _FP, _FP, _FP]
    L.rco_shadow_ray.restype = C.c_int32
    L.rco_shadow_ray.argtypes = [C.POINTER(RcoScene), _FP, _FP, C.c_float]
    L.rco_render_pass.restype = C.c_int32
    L.rco_render_pass.argtypes = [C.POINTER(RcoScene), C.POINTER(RcCamera), C.POINTER(RcPtParams),
                                  C.c_int32, C.c_int32, C.POINTER(RcoFrame), C.POINTER(RcoStats)]
    L.rco_post_chain.restype = C.c_int32
    L.rco_post_chain.argtypes = [_FP, _VP, _VP, _FP, C.c_int32, C.c_int32, C.c_int32, C.POINTER(RcPostParams),
                                 C.POINTER(RcTemporal), _FP, _VP]
    L.rco_post_pass.restype = C.c_int32
    L.rco_post_pass.argtypes = [C.c_int32, _FP, _FP, C.c_int32, C.c_int32, C.POINTER(RcPostParams),
                                C.POINTER(RcTemporal), _VP, _VP, _FP, C.c_int32]
    L.rco_hash.restype = C.c_uint32
    L.rco_hash.argtypes = [C.c_int32, C.c_uint32, C.c_uint32]
    L.rco_fmath.restype = C.c_int32
    L.rco_fmath.argtypes = [C.c_int32, _FP, _FP, _FP, C.c_int32]
    L.rco_num_threads.restype = C.c_int32
    L.rco_num_threads.argtypes = []
    L.rco_gen_ray.restype = C.c_int32
    L.rco_gen_ray.argtypes = [C.POINTER(RcCamera), C.c_int32, C.c_int32, U32P, C.c_float, _FP, _FP]
    L.rco_sample_ngon.restype = None
    L.rco_sample_ngon.argtypes = [U32P, C.c_float, C.c_int32, _FP]
    L.rco_sample_vndf.restype = None
    L.rco_sample_vndf.argtypes = [_FP, C.c_float, C.c_float, C.c_float, _FP]
    L.rco_smith_g1.restype = C.c_float
    L.rco_smith_g1.argtypes = [C.c_float, C.c_float]
    L.rco_smith_g2.restype = C.c_float
    L.rco_smith_g2.argtypes = [C.c_float, C.c_float, C.c_float]
    L.rco_sample_cosine.restype = None
    L.rco_sample_cosine.argtypes = [C.c_float, C.c_float, _FP, _FP]
    L.rco_scatter.restype = C.c_int32
    L.rco_scatter.argtypes = [_FP, _FP, _FP, _FP, U32P, _FP, _FP, _FP, _FP]
    I32P = C.POINTER(C.c_int32)
    L.rco_photon_seed.restype = C.c_uint32
    L.rco_photon_seed.argtypes = [C.c_uint32, C.c_int32]
    L.rco_cell_hash.restype = C.c_uint32
    L.rco_cell_hash.argtypes = [C.c_int32] * 4
    L.rco_photon_trace.restype = C.c_int32
    L.rco_photon_trace.argtypes = [C.POINTER(RcoScene), C.c_int32, C.c_float, C.c_int32, C.c_int32, _FP, I32P, I32P,
                                   I32P]
    L.rco_photon_gather.restype = C.c_int32
    L.rco_photon_gather.argtypes = [C.POINTER(RcoScene), C.POINTER(RcCamera), C.c_float, C.c_int32, _FP, I32P, I32P,
                                    _FP, C.c_void_p]
    L.rco_sppm_next_radius.restype = C.c_float
    L.rco_sppm_next_radius.argtypes = [C.c_float, C.c_int32, C.c_float]
    return L


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = load(LIB_PATH)
    return _lib


@contextlib.contextmanager
def using(library: C.CDLL):
    """Routes every call of this module through `library` (a variant build of the
    oracle) for the duration of the block."""
    global _lib
    saved = lib()
    _lib = library
    try:
        yield library
    finally:
        _lib = saved


def _fp(a):
    return a.ctypes.data_as(_FP)


def _vp(a):
    return None if a is None else a.ctypes.data_as(_VP)


class OracleScene:
    """Holds the float streams alive and the rco_scene view of them."""

    def __init__(self, sc):
        self.tris = np.ascontiguousarray(sc.tris, dtype=np.float32)
        self.nodes = np.ascontiguousarray(sc.nodes, dtype=np.float32)
        self.mats = np.ascontiguousarray(sc.mats, dtype=np.float32)
        lights = sc.lights if sc.lights.shape[0] > 0 else np.zeros((1, 20), np.float32)
        self.lights = np.ascontiguousarray(lights, dtype=np.float32)
        self.c = RcoScene(_fp(self.tris), _fp(self.nodes), _fp(self.mats), _fp(self.lights),
                          self.tris.shape[0], self.nodes.shape[0], self.mats.shape[0], sc.lights.shape[0],
                          float(sc.totalLightPower))


class OracleRenderer:
    """Progressive accumulation on the CPU with the semantics of Renderer.render."""

    def __init__(self, scene, cam_c, params_c, stride: int = 1, fp64: bool = False):
        self.scene = OracleScene(scene)
        self.cam = cam_c
        self.params = params_c
        self.stride = stride
        W, H = cam_c.width, cam_c.height
        n = W * H
        self.accum = np.zeros((n, 4), np.float32)
        self.variance = np.zeros(n, np.float32)
        self.gbuf = np.zeros((n, 4), np.uint16)
        self.galbedo = np.zeros((n, 4), np.uint16)
        self.primId = np.full(n, -7, np.int32)
        self.gbufId = np.full(n, -7, np.int32)
        self.accum64 = np.zeros((n, 4), np.float64) if fp64 else None
        self.frame = 0
        self.stats = RcoStats()

    def render_pass(self, collect_stats: bool = False):
        f = RcoFrame(_vp(self.accum), _vp(self.variance), _vp(self.gbuf), _vp(self.galbedo), _vp(self.primId),
                     _vp(self.gbufId), _vp(self.accum64))
        st = C.byref(self.stats) if collect_stats else None
        rc = lib().rco_render_pass(C.byref(self.scene.c), C.byref(self.cam), C.byref(self.params), self.frame,
                                   self.stride, C.byref(f), st)
        if rc != 0:
            raise RuntimeError("rco_render_pass failed")
        self.frame += 1

    def mask(self):
        W, H = self.cam.width, self.cam.height
        m = np.zeros((H, W), dtype=bool)
        m[::self.stride, ::self.stride] = True
        return m.reshape(-1)


def post_chain(accum, W, H, total, params, temporal, gbuf=None, galbedo=None, variance=None):
    """rco_post_chain: returns (hdr float32 (H*W,4), rgba8 uint8 (H*W,4))."""
    accum = np.ascontiguousarray(accum, np.float32).reshape(-1, 4)
    hdr = np.zeros((W * H, 4), np.float32)
    rgba = np.zeros((W * H, 4), np.uint8)
    gb = None if gbuf is None else np.ascontiguousarray(gbuf, np.uint16)
    ga = None if galbedo is None else np.ascontiguousarray(galbedo, np.uint16)
    va = None if variance is None else np.ascontiguousarray(variance, np.float32)
    rc = lib().rco_post_chain(_fp(accum), _vp(gb), _vp(ga), None if va is None else _fp(va), W, H, int(total),
                              C.byref(params), C.byref(temporal), _fp(hdr), rgba.ctypes.data_as(_VP))
    if rc != 0:
        raise RuntimeError("rco_post_chain failed")
    return hdr, rgba


def post_pass(pass_id, src, W, H, params, temporal, gbuf=None, galbedo=None, variance=None, total=1):
    src = np.ascontiguousarray(src, np.float32).reshape(-1, 4)
    dst = np.zeros_like(src)
    gb = None if gbuf is None else np.ascontiguousarray(gbuf, np.uint16)
    ga = None if galbedo is None else np.ascontiguousarray(galbedo, np.uint16)
    va = None if variance is None else np.ascontiguousarray(variance, np.float32)
    rc = lib().rco_post_pass(int(pass_id), _fp(src), _fp(dst), W, H, C.byref(params), C.byref(temporal), _vp(gb),
                             _vp(ga), None if va is None else _fp(va), int(total))
    if rc != 0:
        raise RuntimeError("rco_post_pass failed")
    return dst


def hash_u32(kind: int, a: int, b: int = 0) -> int:
    """0 pcgHash(a), 1 one-round uHash(a), 2 two-round uHash(a),
    3 float bits of fHash1(a, b), 4 float bits of fHash2(a, b)."""
    return int(lib().rco_hash(int(kind), int(a) & 0xFFFFFFFF, int(b) & 0xFFFFFFFF))


def fmath(fn: int, x, y=None):
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(np.zeros_like(x) if y is None else y, np.float32)
    out = np.zeros_like(x)
    if lib().rco_fmath(int(fn), _fp(x), _fp(y), _fp(out), x.size) != 0:
        raise ValueError("bad fmath function id")
    return out


def closest_hit(scene: OracleScene, ro, rd):
    ro = np.asarray(ro, np.float32)
    rd = np.asarray(rd, np.float32)
    t = np.zeros(1, np.float32)
    n = np.zeros(3, np.float32)
    tri = lib().rco_closest_hit(C.byref(scene.c), _fp(ro), _fp(rd), _fp(t), _fp(n))
    return tri, float(t[0]), n


def brute_closest(scene: OracleScene, ro, rd):
    ro = np.asarray(ro, np.float32)
    rd = np.asarray(rd, np.float32)
    t = np.zeros(1, np.float32)
    tri = lib().rco_brute_closest(C.byref(scene.c), _fp(ro), _fp(rd), _fp(t))
    return tri, float(t[0])


def shadow_ray(scene: OracleScene, ro, rd, max_dist):
    ro = np.asarray(ro, np.float32)
    rd = np.asarray(rd, np.float32)
    return bool(lib().rco_shadow_ray(C.byref(scene.c), _fp(ro), _fp(rd), float(max_dist)))


# ---- sub-step probes of pathtrace.comp (for the oracle's own pins) ----------
def _u32(v: int):
    return C.c_uint32(int(v) & 0xFFFFFFFF)


def rand_stream(seed: int, n: int):
    """The next n rand() values of a seed stream, and the state after them."""
    s = _u32(seed)
    out = [float(lib().rco_rand(C.byref(s))) for _ in range(n)]
    return out, int(s.value)


def gen_ray(cam_c, px: int, py: int, seed: int, focal_offset: float = 0.0):
    """generateRay: (origin, direction, seed state after the ray)."""
    s = _u32(seed)
    ro = np.zeros(3, np.float32)
    rd = np.zeros(3, np.float32)
    if lib().rco_gen_ray(C.byref(cam_c), int(px), int(py), C.byref(s), float(focal_offset), _fp(ro), _fp(rd)) != 0:
        raise ValueError("rco_gen_ray rejected its arguments")
    return ro, rd, int(s.value)


def sample_ngon(seed: int, radius: float, blades: int):
    s = _u32(seed)
    xy = np.zeros(2, np.float32)
    lib().rco_sample_ngon(C.byref(s), float(radius), int(blades), _fp(xy))
    return xy, int(s.value)


def sample_vndf(wo, alpha: float, u1: float, u2: float):
    wo = np.ascontiguousarray(wo, np.float32)
    h = np.zeros(3, np.float32)
    lib().rco_sample_vndf(_fp(wo), float(alpha), float(u1), float(u2), _fp(h))
    return h


def smith_g1(nv: float, alpha: float) -> float:
    return float(lib().rco_smith_g1(float(nv), float(alpha)))


def smith_g2(nl: float, nv: float, alpha: float) -> float:
    return float(lib().rco_smith_g2(float(nl), float(nv), float(alpha)))


def sample_cosine(u1: float, u2: float, n):
    n = np.ascontiguousarray(n, np.float32)
    d = np.zeros(3, np.float32)
    lib().rco_sample_cosine(float(u1), float(u2), _fp(n), _fp(d))
    return d


def scatter(mat, rd, geom_normal, pos, seed: int):
    """Specular scattering step: (origin, direction, throughput factor, Fresnel
    term or -1, seed state after)."""
    mat = np.ascontiguousarray(mat, np.float32)
    rd = np.ascontiguousarray(rd, np.float32)
    gn = np.ascontiguousarray(geom_normal, np.float32)
    pos = np.ascontiguousarray(pos, np.float32)
    s = _u32(seed)
    ro = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    thr = np.zeros(3, np.float32)
    F = C.c_float()
    if lib().rco_scatter(_fp(mat), _fp(rd), _fp(gn), _fp(pos), C.byref(s), _fp(ro), _fp(d), _fp(thr), C.byref(F)) != 0:
        raise ValueError("rco_scatter: not a specular material")
    return ro, d, thr, float(F.value), int(s.value)
