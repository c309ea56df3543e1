"""Seeded synthetic workloads of the render hot path.

  cornell_small   320x180, 16 spp, pinhole: 5 wall quads, 2 boxes, a ceiling
                  quad light (radiance 15), a 32x16 UV sphere; Lambertian
                  albedos U[0.1, 0.8] from the seed.
  full_frame      854x480, 128 spp: ~250k triangles of jittered spheres/boxes
                  (50% diffuse, 30% GGX metal alpha U[0.05, 0.6], 20% glass
                  ior 1.5) in a closed room with 8 quad area lights; thin lens
                  35 mm f/2.8, hexagon bokeh, focus 2.0 m, SA 0.04.
  serving_batch   4 scenes (50k-250k triangles) + 32 uneven view requests.
  training_scene  seeded size x size denoiser-training views (rooms with random
                  objects, lights and lenses).
  reference_default  the reference's default Cornell scene
                  (src/main/java/com/retrocam/scene/SceneEditor.java:111-134) at 854x480.
"""
from __future__ import annotations

import numpy as np

from .scene import Material, PtParams, SceneBuilder, Workload, lookat_camera, orbit_camera


def cornell_small(seed: int = 1, width: int = 320, height: int = 180, spp: int = 16,
                  clamp: float = 1.0) -> Workload:
    rng = np.random.default_rng(seed)
    b = SceneBuilder()
    alb = rng.uniform(0.1, 0.8, size=(8, 3)).astype(np.float32)
    m = [b.add_material(Material.diffuse(*alb[i])) for i in range(8)]
    light = b.add_material(Material.emissive(1.0, 1.0, 1.0, 15.0))
    x0, x1, y0, y1, z0, z1 = -5.0, 5.0, 0.0, 10.0, -5.0, 5.0
    b.add_quad((x0, y0, z0), (x0, y0, z1), (x1, y0, z1), (x1, y0, z0), m[0])  # floor
    b.add_quad((x0, y1, z0), (x1, y1, z0), (x1, y1, z1), (x0, y1, z1), m[1])  # ceiling
    b.add_quad((x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0), m[2])  # back
    b.add_quad((x0, y0, z0), (x0, y1, z0), (x0, y1, z1), (x0, y0, z1), m[3])  # left
    b.add_quad((x1, y0, z0), (x1, y0, z1), (x1, y1, z1), (x1, y1, z0), m[4])  # right
    b.add_box((-1.5, 1.5, -1.0), (1.0, 1.5, 1.0), m[5])
    b.add_box((1.5, 1.0, 0.5), (1.0, 1.0, 1.0), m[6])
    ly = 9.99
    b.add_quad((-1.5, ly, -1.5), (1.5, ly, -1.5), (1.5, ly, 1.5), (-1.5, ly, 1.5), light)
    b.add_sphere((0.3, 1.2, -3.4), 1.2, m[7], stacks=16, slices=32)
    scene = b.build()
    cam = orbit_camera(width, height, pinhole=True)
    return Workload("cornell_small", scene, cam, PtParams(1, clamp, 6), spp, {"seed": seed})


def full_frame(seed: int = 1, n_tris: int = 250_000, width: int = 854, height: int = 480, spp: int = 128,
               clamp: float = 1.0) -> Workload:
    rng = np.random.default_rng(seed)
    b = SceneBuilder()
    walls = [b.add_material(Material.diffuse(*rng.uniform(0.1, 0.8, 3))) for _ in range(6)]
    lightm = b.add_material(Material.emissive(1.0, 0.95, 0.85, 12.0))
    X, Y, Z = 8.0, 6.0, 8.0
    # Room faces as 8x8 quad grids: large single triangles would share BVH nodes
    # with dense object clusters and, with the builder's node-bounds binning
    # (BVHBuilder.java:107-118), end up in leaves of hundreds of triangles.
    g = 8

    def grid(o, du, dv, mat):
        for i in range(g):
            for j in range(g):
                a = [o[k] + du[k] * i / g + dv[k] * j / g for k in range(3)]
                bb = [o[k] + du[k] * (i + 1) / g + dv[k] * j / g for k in range(3)]
                c = [o[k] + du[k] * (i + 1) / g + dv[k] * (j + 1) / g for k in range(3)]
                d = [o[k] + du[k] * i / g + dv[k] * (j + 1) / g for k in range(3)]
                b.add_quad(a, bb, c, d, mat)

    grid((-X, 0, -Z), (0, 0, 2 * Z), (2 * X, 0, 0), walls[0])
    grid((-X, Y, -Z), (2 * X, 0, 0), (0, 0, 2 * Z), walls[1])
    grid((-X, 0, -Z), (2 * X, 0, 0), (0, Y, 0), walls[2])
    grid((-X, 0, Z), (0, Y, 0), (2 * X, 0, 0), walls[3])
    grid((-X, 0, -Z), (0, Y, 0), (0, 0, 2 * Z), walls[4])
    grid((X, 0, -Z), (0, 0, 2 * Z), (0, Y, 0), walls[5])
    ly = Y - 0.01
    for i in range(4):
        for j in range(2):
            cx, cz = -4.5 + 3.0 * i, -3.0 + 5.0 * j
            b.add_quad((cx - 0.5, ly, cz - 0.5), (cx + 0.5, ly, cz - 0.5), (cx + 0.5, ly, cz + 0.5),
                       (cx - 0.5, ly, cz + 0.5), lightm)
    sphere_res = [(12, 24), (16, 32), (24, 48)]  # 552, 960, 2208 triangles (multiples of 12)
    count = b.triangle_count()
    assert (n_tris - count) % 12 == 0

    def material():
        u = rng.random()
        if u < 0.5:
            return b.add_material(Material.diffuse(*rng.uniform(0.1, 0.8, 3)))
        if u < 0.8:
            return b.add_material(Material.metal(*rng.uniform(0.5, 0.95, 3), float(rng.uniform(0.05, 0.6))))
        return b.add_material(Material.glass(1.5))

    while True:
        remaining = n_tris - count
        if remaining <= 0:
            break
        x, z = rng.uniform(-7.0, 7.0), rng.uniform(-7.0, 5.0)
        r = float(rng.uniform(0.15, 0.6))
        y = r + float(rng.uniform(0.0, 3.0)) * (rng.random() < 0.3)
        if remaining >= 2208 and rng.random() < 0.6:
            st, sl = sphere_res[int(rng.integers(0, 3))]
            count += b.add_sphere((x, y, z), r, material(), stacks=st, slices=sl)
        else:
            rot = tuple(float(a) for a in rng.uniform(0.0, 90.0, 3))
            count += b.add_box((x, y, z), (r, r * float(rng.uniform(0.5, 1.5)), r), material(), rot_deg=rot)
    scene = b.build()
    cam = lookat_camera(width, height, eye=(0.0, 2.2, 7.0), target=(0.0, 1.0, 0.0), focalDist=2.0,
                        focalLengthMm=35.0, fStop=2.8, blades=6, saStrength=0.04)
    return Workload("full_frame", scene, cam, PtParams(1, clamp, 6), spp, {"seed": seed, "n_tris": n_tris})


def sppm_caustics(seed: int = 1, width: int = 854, height: int = 480, spp: int = 64, photons: int = 2_000_000,
                  hash_size: int = 1 << 22, init_radius: float = 0.02, alpha: float = 0.7,
                  sphere_res=(32, 64)) -> Workload:
    """The cornell_small scene plus three ior-1.5 glass spheres (32 stacks x 64
    slices = 3968 triangles each) under a small 1x1 quad light."""
    w = cornell_small(seed, width=width, height=height, spp=spp)
    rng = np.random.default_rng(seed)
    b = SceneBuilder()
    # rebuild the Cornell scene (same materials and geometry), then add the spheres
    alb = rng.uniform(0.1, 0.8, size=(8, 3)).astype(np.float32)
    m = [b.add_material(Material.diffuse(*alb[i])) for i in range(8)]
    light = b.add_material(Material.emissive(1.0, 1.0, 1.0, 15.0))
    x0, x1, y0, y1, z0, z1 = -5.0, 5.0, 0.0, 10.0, -5.0, 5.0
    b.add_quad((x0, y0, z0), (x0, y0, z1), (x1, y0, z1), (x1, y0, z0), m[0])
    b.add_quad((x0, y1, z0), (x1, y1, z0), (x1, y1, z1), (x0, y1, z1), m[1])
    b.add_quad((x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0), m[2])
    b.add_quad((x0, y0, z0), (x0, y1, z0), (x0, y1, z1), (x0, y0, z1), m[3])
    b.add_quad((x1, y0, z0), (x1, y0, z1), (x1, y1, z1), (x1, y1, z0), m[4])
    b.add_box((-1.5, 1.5, -1.0), (1.0, 1.5, 1.0), m[5])
    b.add_box((1.5, 1.0, 0.5), (1.0, 1.0, 1.0), m[6])
    ly = 9.99
    b.add_quad((-1.5, ly, -1.5), (1.5, ly, -1.5), (1.5, ly, 1.5), (-1.5, ly, 1.5), light)
    b.add_sphere((0.3, 1.2, -3.4), 1.2, m[7], stacks=16, slices=32)
    glass = b.add_material(Material.glass(1.5))
    small = b.add_material(Material.emissive(1.0, 0.9, 0.7, 40.0))
    st, sl = sphere_res
    for cx, cz, r in ((-2.5, 2.0, 0.7), (0.0, 2.6, 0.8), (2.6, 2.2, 0.6)):
        b.add_sphere((cx, r + 0.6, cz), r, glass, stacks=st, slices=sl)
    sy = 6.0
    b.add_quad((-0.5, sy, 1.9), (0.5, sy, 1.9), (0.5, sy, 2.9), (-0.5, sy, 2.9), small)
    scene = b.build()
    return Workload("sppm_caustics", scene, w.camera, w.params, spp,
                    {"seed": seed, "photons": photons, "hash_size": hash_size, "init_radius": init_radius,
                     "alpha": alpha})


def training_scene(seed: int, size: int = 128, spp: int = 8) -> Workload:
    """A seeded denoiser-training view (size x size): a closed room with random
    wall albedos, 3-10 large (even seeds) or 40-160 small (odd seeds) jittered
    spheres/boxes with the full-frame material mix (50% diffuse, 30% GGX metal,
    20% glass), 1-3 quad lights of random strength and a random thin-lens or
    pinhole camera looking into the room."""
    rng = np.random.default_rng(100_000 + seed)
    b = SceneBuilder()
    walls = [b.add_material(Material.diffuse(*rng.uniform(0.1, 0.8, 3))) for _ in range(6)]
    X, Y, Z = float(rng.uniform(4, 8)), float(rng.uniform(4, 7)), float(rng.uniform(4, 8))
    clutter = seed % 2 == 1  # odd seeds: many small objects, as in the full frame
    b.add_quad((-X, 0, -Z), (-X, 0, Z), (X, 0, Z), (X, 0, -Z), walls[0])
    b.add_quad((-X, Y, -Z), (X, Y, -Z), (X, Y, Z), (-X, Y, Z), walls[1])
    b.add_quad((-X, 0, -Z), (X, 0, -Z), (X, Y, -Z), (-X, Y, -Z), walls[2])
    b.add_quad((-X, 0, Z), (-X, Y, Z), (X, Y, Z), (X, 0, Z), walls[3])
    b.add_quad((-X, 0, -Z), (-X, Y, -Z), (-X, Y, Z), (-X, 0, Z), walls[4])
    b.add_quad((X, 0, -Z), (X, 0, Z), (X, Y, Z), (X, Y, -Z), walls[5])
    for _ in range(int(rng.integers(1, 4))):
        lm = b.add_material(Material.emissive(*rng.uniform(0.7, 1.0, 3), float(rng.uniform(5.0, 30.0))))
        cx, cz = rng.uniform(-0.7 * X, 0.7 * X), rng.uniform(-0.7 * Z, 0.7 * Z)
        hx, hz = rng.uniform(0.3, 1.2, 2)
        ly = Y - 0.01
        b.add_quad((cx - hx, ly, cz - hz), (cx + hx, ly, cz - hz), (cx + hx, ly, cz + hz), (cx - hx, ly, cz + hz), lm)

    def material():
        u = rng.random()
        if u < 0.5:
            return b.add_material(Material.diffuse(*rng.uniform(0.1, 0.8, 3)))
        if u < 0.8:
            return b.add_material(Material.metal(*rng.uniform(0.5, 0.95, 3), float(rng.uniform(0.05, 0.6))))
        return b.add_material(Material.glass(1.5))

    for _ in range(int(rng.integers(40, 160)) if clutter else int(rng.integers(3, 11))):
        x, z = rng.uniform(-0.8 * X, 0.8 * X), rng.uniform(-0.8 * Z, 0.8 * Z)
        r = float(rng.uniform(0.15, 0.6)) if clutter else float(rng.uniform(0.3, 1.2))
        y = r + float(rng.uniform(0.0, 1.5)) * (rng.random() < 0.3)
        if rng.random() < 0.5:
            b.add_sphere((x, y, z), r, material(), stacks=16, slices=32)
        else:
            rot = tuple(float(a) for a in rng.uniform(0.0, 90.0, 3))
            b.add_box((x, y, z), (r, r * float(rng.uniform(0.5, 1.5)), r), material(), rot_deg=rot)
    scene = b.build()
    eye = (float(rng.uniform(-0.6 * X, 0.6 * X)), float(rng.uniform(0.8, 0.8 * Y)), float(0.9 * Z))
    tgt = (float(rng.uniform(-0.3 * X, 0.3 * X)), float(rng.uniform(0.3, 0.5 * Y)), float(rng.uniform(-0.5 * Z, 0.2 * Z)))
    if rng.random() < 0.6:
        dist = float(np.linalg.norm(np.subtract(eye, tgt)))
        cam = lookat_camera(size, size, eye=eye, target=tgt, focalDist=float(dist * rng.uniform(0.5, 1.2)),
                            focalLengthMm=35.0, fStop=float(rng.choice([1.4, 2.0, 2.8, 5.6])),
                            blades=int(rng.integers(5, 8)), saStrength=0.04)
    else:
        cam = lookat_camera(size, size, eye=eye, target=tgt, pinhole=True)
    return Workload("training_scene", scene, cam, PtParams(1, 1.0, 6), spp, {"seed": seed})


def serving_batch(seed: int = 1, n_requests: int = 32):
    """Batched-serving workload: 4 seeded full-frame-style scenes of ~50k, 100k,
    150k and 250k triangles and n_requests view requests with resolutions drawn
    from {320x180, 640x360, 854x480, 1280x720}, spp from U{4..128}, a random
    scene and a jittered thin-lens camera. Returns (scenes, [(slot, camera,
    params, spp)])."""
    sizes = (50_008, 100_012, 150_004, 250_000)  # room + lights = 784 triangles, objects in 12s
    scenes = [full_frame(seed=21 + i, n_tris=n, spp=1).scene for i, n in enumerate(sizes)]
    rng = np.random.default_rng(1000 + seed)
    res = [(320, 180), (640, 360), (854, 480), (1280, 720)]
    reqs = []
    for _ in range(n_requests):
        w, h = res[int(rng.integers(0, 4))]
        spp = int(rng.integers(4, 129))
        slot = int(rng.integers(0, 4))
        eye = (float(rng.uniform(-3, 3)), float(rng.uniform(1.2, 3.0)), float(rng.uniform(5.0, 7.5)))
        tgt = (float(rng.uniform(-2, 2)), float(rng.uniform(0.5, 1.5)), float(rng.uniform(-2, 1)))
        cam = lookat_camera(w, h, eye=eye, target=tgt, focalDist=float(rng.uniform(1.5, 5.0)), focalLengthMm=35.0,
                            fStop=2.8, blades=6, saStrength=0.04)
        reqs.append((slot, cam, PtParams(1, 1.0, 6), spp))
    return scenes, reqs


def reference_default(width: int = 854, height: int = 480, spp: int = 16) -> Workload:
    """SceneEditor.createDefault (SceneEditor.java:111-134), default orbit camera and lens."""
    b = SceneBuilder()
    white = b.add_material(Material.diffuse(0.73, 0.73, 0.73))
    red = b.add_material(Material.diffuse(0.65, 0.05, 0.05))
    green = b.add_material(Material.diffuse(0.12, 0.45, 0.15))
    light = b.add_material(Material.emissive(1.0, 0.95, 0.85, 15.0))
    metal = b.add_material(Material.metal(0.95, 0.85, 0.55, 0.1))
    glass = b.add_material(Material.glass(1.5))
    H = 5.0
    b.add_box((0, 0, 0), (H, 0.05, H), white)
    b.add_box((0, 2 * H, 0), (H, 0.05, H), white)
    b.add_box((-H, H, 0), (0.05, H, H), red)
    b.add_box((H, H, 0), (0.05, H, H), green)
    b.add_box((0, H, -H), (H, H, 0.05), white)
    b.add_box((0, 2 * H - 0.1, 0), (1.5, 0.05, 1.5), light)
    b.add_box((-1.5, 1.5, -1.0), (1.0, 1.5, 1.0), white)
    b.add_box((1.5, 1.0, 0.5), (1.0, 1.0, 1.0), metal)
    b.add_sphere((0.0, 1.2, 1.5), 1.2, glass, stacks=24, slices=24)
    scene = b.build()
    cam = orbit_camera(width, height, focalDist=2.0, focalLengthMm=35.0, fStop=2.0, blades=5, saStrength=0.04)
    return Workload("reference_default", scene, cam, PtParams(1, 1.0, 6), spp, {})


def tiny_random(seed: int, n_tris: int = 48, extent: float = 2.0) -> "SceneArrays":
    """Random triangle soup with random materials (for brute-force traversal tests)."""
    rng = np.random.default_rng(seed)
    b = SceneBuilder()
    mats = [b.add_material(Material.diffuse(0.5, 0.5, 0.5)),
            b.add_material(Material.emissive(1.0, 1.0, 1.0, 5.0))]
    tris = np.zeros((n_tris, 32), dtype=np.float32)
    c = rng.uniform(-extent, extent, size=(n_tris, 3))
    for k in range(3):
        tris[:, 4 * k:4 * k + 3] = c + rng.uniform(-0.6, 0.6, size=(n_tris, 3))
    e1 = tris[:, 4:7] - tris[:, 0:3]
    e2 = tris[:, 8:11] - tris[:, 0:3]
    n = np.cross(e1, e2)
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-12)
    for k in range(3):
        tris[:, 12 + 4 * k:15 + 4 * k] = n
    mi = rng.integers(0, len(mats), size=n_tris).astype(np.int32)
    tris[:, 30] = mi.view(np.float32)
    b.add_triangles(tris)
    return b.build()
