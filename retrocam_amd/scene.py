"""Scene assembly for tests, benchmarks and C++-less callers.

Materials mirror the factories of src/main/java/com/retrocam/scene/Material.java:37-60;
geometry, BVH and lights come from librc_host.so (a Java-semantics restatement of
Box/Sphere/TransformedPrimitive/BVHBuilder/SceneUploader.uploadLights).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N

f32 = np.float32


@dataclass
class Material:
    """Material.java:15-26 fields; write() is Material.write (Material.java:61-66)."""
    albedo: tuple = (0.8, 0.8, 0.8)
    metallic: float = 0.0
    emission: tuple = (0.0, 0.0, 0.0)
    emissionStrength: float = 0.0
    roughness: float = 0.5
    ior: float = 1.5
    transmission: float = 0.0
    albedoTex: int = -1

    def write(self) -> np.ndarray:
        o = np.zeros(16, dtype=f32)
        o[0:3] = self.albedo
        o[3] = self.metallic
        o[4:7] = self.emission
        o[7] = self.emissionStrength
        o[8] = self.roughness
        o[9] = self.ior
        o[10] = self.transmission
        o[11] = np.array([self.albedoTex], dtype=np.int32).view(f32)[0]
        return o

    @staticmethod
    def diffuse(r, g, b):
        return Material(albedo=(r, g, b))

    @staticmethod
    def emissive(r, g, b, strength):
        return Material(albedo=(r, g, b), emission=(r, g, b), emissionStrength=strength)

    @staticmethod
    def metal(r, g, b, roughness):
        return Material(albedo=(r, g, b), metallic=1.0, roughness=roughness)

    @staticmethod
    def glass(ior):
        return Material(albedo=(1.0, 1.0, 1.0), transmission=1.0, roughness=0.0, ior=ior)


@dataclass
class SceneArrays:
    """Built scene: the float streams handed to rc_upload_scene."""
    tris: np.ndarray      # (nTris, 32) float32
    nodes: np.ndarray     # (nNodes, 12) float32
    mats: np.ndarray      # (nMats, 16) float32
    lights: np.ndarray    # (nLights, 20) float32
    totalLightPower: float
    maxDepth: int
    maxLeafSize: int
    nLeaves: int

    @property
    def nTris(self):
        return self.tris.shape[0]

    @property
    def nNodes(self):
        return self.nodes.shape[0]

    @property
    def nLights(self):
        return self.lights.shape[0]


class SceneBuilder:
    """Accumulates materials and primitives, then builds BVH + lights in C++."""

    def __init__(self):
        self._lib = N.host_lib()
        self._s = self._lib.rcs_new()
        if not self._s:
            raise MemoryError("rcs_new failed")

    def __del__(self):
        try:
            if self._s:
                self._lib.rcs_free(self._s)
                self._s = None
        except Exception:
            pass

    def add_material(self, m: Material) -> int:
        w = m.write()
        idx = self._lib.rcs_add_material(self._s, N.fptr(w))
        if idx < 0:
            raise RuntimeError("rcs_add_material failed")
        return idx

    def add_box(self, c, h, mat: int, rot_deg=None) -> int:
        first = self.triangle_count()
        n = self._lib.rcs_add_box(self._s, *[float(v) for v in c], *[float(v) for v in h], int(mat))
        if n < 0:
            raise RuntimeError("rcs_add_box failed")
        if rot_deg is not None and any(r != 0 for r in rot_deg):
            self.rotate_since(first, c, rot_deg)
        return n

    def add_sphere(self, c, r: float, mat: int, stacks=24, slices=24) -> int:
        n = self._lib.rcs_add_sphere(self._s, *[float(v) for v in c], float(r), int(mat), int(stacks), int(slices))
        if n < 0:
            raise RuntimeError("rcs_add_sphere failed")
        return n

    def add_quad(self, a, b, c, d, mat: int) -> int:
        p = np.array([*a, *b, *c, *d], dtype=f32)
        if self._lib.rcs_add_quad(self._s, N.fptr(p), int(mat)) < 0:
            raise RuntimeError("rcs_add_quad failed")
        return 2

    def rotate_since(self, first: int, pivot, deg):
        if self._lib.rcs_rotate_since(self._s, int(first), *[float(v) for v in pivot], *[float(v) for v in deg]) < 0:
            raise RuntimeError("rcs_rotate_since failed")

    def add_triangles(self, tris32: np.ndarray):
        t = np.ascontiguousarray(tris32, dtype=f32).reshape(-1, 32)
        if self._lib.rcs_add_triangles(self._s, N.fptr(t), t.shape[0]) < 0:
            raise RuntimeError("rcs_add_triangles failed")

    def triangle_count(self) -> int:
        return self._lib.rcs_triangle_count(self._s)

    def build(self) -> SceneArrays:
        if self._lib.rcs_build(self._s) != 0:
            raise RuntimeError("rcs_build failed: empty scene, a bad material index, or a node only a few "
                               "ulps thick whose bucket index falls below 0 (where BVHBuilder throws)")
        info = N.RcsInfo()
        self._lib.rcs_get_info(self._s, C.byref(info))
        tris = np.zeros((info.nTris, 32), dtype=f32)
        nodes = np.zeros((info.nNodes, 12), dtype=f32)
        mats = np.zeros((info.nMats, 16), dtype=f32)
        lights = np.zeros((max(info.nLights, 0), 20), dtype=f32)
        self._lib.rcs_copy_tris(self._s, N.fptr(tris))
        self._lib.rcs_copy_nodes(self._s, N.fptr(nodes))
        self._lib.rcs_copy_mats(self._s, N.fptr(mats))
        if info.nLights > 0:
            self._lib.rcs_copy_lights(self._s, N.fptr(lights))
        return SceneArrays(tris, nodes, mats, lights, float(info.totalLightPower), info.maxDepth,
                           info.maxLeafSize, info.nLeaves)


@dataclass
class CameraSpec:
    """Camera uniforms of ThinLensCamera.uploadTo (ThinLensCamera.java:27-58)."""
    eye: tuple
    forward: tuple
    right: tuple
    up: tuple
    width: int
    height: int
    focalDist: float = 2.0
    focalLengthMm: float = 35.0
    fStop: float = 2.0
    blades: int = 5
    saStrength: float = 0.04
    lcaDelta: tuple = (0.0, 0.0, 0.0)
    pinhole: bool = False

    def lens_radius(self) -> float:
        if self.pinhole:
            return 0.0
        focal_m = f32(self.focalLengthMm) / f32(1000.0)
        return float(focal_m / (f32(2.0) * f32(self.fStop)))

    def to_c(self) -> N.RcCamera:
        c = N.RcCamera()
        c.camPos[:] = [float(v) for v in self.eye]
        c.camRight[:] = [float(v) for v in self.right]
        c.camUp[:] = [float(v) for v in self.up]
        c.camForward[:] = [float(v) for v in self.forward]
        c.focalDist = float(self.focalDist)
        c.lensRadius = self.lens_radius()
        c.fovTanHalf = float(N.host_lib().rcs_tan_half_fov(float(self.focalLengthMm)))
        c.apertureBlades = int(self.blades)
        c.saStrength = float(self.saStrength)
        c.lcaDelta[:] = [float(v) for v in self.lcaDelta]
        c.width = int(self.width)
        c.height = int(self.height)
        return c


def orbit_camera(width, height, theta=None, phi=1.2, radius=10.0, target=(0.0, 5.0, 0.0), **kw) -> CameraSpec:
    """Default OrbitCamera state (OrbitCamera.java:27-33): theta = (float)(PI/2)."""
    if theta is None:
        theta = float(f32(np.pi / 2.0))
    lib = N.host_lib()
    eye, fwd, right, up = (np.zeros(3, dtype=f32) for _ in range(4))
    lib.rcs_orbit_basis(float(theta), float(phi), float(radius), *[float(t) for t in target],
                        N.fptr(eye), N.fptr(fwd), N.fptr(right), N.fptr(up))
    return CameraSpec(tuple(eye), tuple(fwd), tuple(right), tuple(up), width, height, **kw)


def lookat_camera(width, height, eye, target, **kw) -> CameraSpec:
    lib = N.host_lib()
    e = np.array(eye, dtype=f32)
    t = np.array(target, dtype=f32)
    fwd, right, up = (np.zeros(3, dtype=f32) for _ in range(3))
    lib.rcs_lookat_basis(N.fptr(e), N.fptr(t), N.fptr(fwd), N.fptr(right), N.fptr(up))
    return CameraSpec(tuple(e), tuple(fwd), tuple(right), tuple(up), width, height, **kw)


@dataclass
class PtParams:
    """Renderer NEE uniforms (Renderer.java:93-96); maxBounces = pathtrace.comp:465."""
    neeEnabled: int = 1
    neeFireflyClamp: float = 1.0
    maxBounces: int = 6

    def to_c(self) -> N.RcPtParams:
        p = N.RcPtParams()
        p.neeEnabled = int(self.neeEnabled)
        p.neeFireflyClamp = float(self.neeFireflyClamp)
        p.maxBounces = int(self.maxBounces)
        return p


@dataclass
class Workload:
    name: str
    scene: SceneArrays
    camera: CameraSpec
    params: PtParams
    spp: int
    extra: dict = field(default_factory=dict)
