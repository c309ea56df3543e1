"""CPU pins of the oracle's camera and specular BSDF code (rc_oracle.c) by what
the mathematics fixes, independent of the oracle's own formulas:

- generateRay (pathtrace.comp:401-439): the pinhole ray against its closed form;
  every thin-lens ray through the focal point of its pixel's pinhole ray, on a
  focal plane offset by the LCA channel offset and by spherical aberration
  saStrength * (lens radius fraction)^2 (clamped at 0.01); the draw count;
- sampleNGon (pathtrace.comp:310-321): samples inside the regular polygon with
  uniform area density (chi-squared over fan sectors, radial and angular
  barycentric coordinates);
- LCA (pathtrace.comp:611-613): the channel whose focal offset is 0 images a
  point on the focal plane sharply, the others blur it;
- sampleGGXVNDF / smithG1GGX / smithG2GGX (pathtrace.comp:252-292): sampled half
  vectors against the GGX visible-normal density (Heitz 2018), G1 and the
  height-correlated G2 against their tan-form closed forms (Heitz 2014); the
  metal step's weight F*G2/G1 in [0, 1] and against those forms, and its
  roughness-to-alpha mapping from the normal-incidence half-angle distribution;
- sampleCosineHemisphere (pathtrace.comp:296-306): cos^2 theta and the azimuth
  uniform;
- glass (pathtrace.comp:504-523): ior 1 is a pass-through, Schlick's F with
  F(0) = r0, Snell's law, total internal reflection beyond the critical angle,
  and reflection chosen with probability F.

Each pin is also run against oracle builds with one planted mistake
(test_planted_mistakes_fail) to show that it can fail."""
import numpy as np
import pytest
from scipy import stats

from oracle import rco
from retrocam_amd import _native as N
from retrocam_amd.scene import Material, SceneBuilder
from oracle_variants import pin_fails

F32 = np.float32


def make_cam(P, F, R, U, tan_half, W, H, focal=2.0, lens=0.0, blades=6, sa=0.0, lca=(0.0, 0.0, 0.0)):
    c = N.RcCamera()
    for name, v in (("camPos", P), ("camForward", F), ("camRight", R), ("camUp", U)):
        arr = getattr(c, name)
        for i in range(3):
            arr[i] = float(v[i])
    c.fovTanHalf, c.width, c.height = float(tan_half), int(W), int(H)
    c.focalDist, c.lensRadius, c.apertureBlades, c.saStrength = float(focal), float(lens), int(blades), float(sa)
    for i in range(3):
        c.lcaDelta[i] = float(lca[i])
    return c


def cam_vec(c, name):
    return np.array(list(getattr(c, name)), np.float64)  # the float32 values the kernel reads


def random_frame(rng):
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    return q[:, 0], q[:, 1], np.cross(q[:, 0], q[:, 1])  # F, R, U orthonormal


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def random_camera(rng, lens=0.0, sa=0.0):
    F, R, U = random_frame(rng)
    W, H = (int(v) for v in rng.integers(8, 400, 2))
    return make_cam(rng.uniform(-5, 5, 3), F, R, U, rng.uniform(0.05, 1.2), W, H, focal=rng.uniform(0.3, 10.0),
                    lens=lens, blades=int(rng.integers(3, 9)), sa=sa)


def pinhole_dir(c, px, py, u1, u2):
    """Direction through the jittered pixel position (px + u1, py + u2)."""
    W, H = c.width, c.height
    sx = 2.0 * (px + u1) / W - 1.0
    sy = 2.0 * (py + u2) / H - 1.0
    t = float(F32(c.fovTanHalf))
    d = cam_vec(c, "camForward") + cam_vec(c, "camRight") * (sx * (W / H) * t) + cam_vec(c, "camUp") * (sy * t)
    return unit(d)


# ---- generateRay -----------------------------------------------------------
def pin_pinhole_closed_form():
    rng = np.random.default_rng(21)
    for _ in range(300):
        c = random_camera(rng, lens=0.0, sa=float(rng.uniform(-1, 1)))
        px, py = int(rng.integers(0, c.width)), int(rng.integers(0, c.height))
        seed = int(rng.integers(0, 2**32))
        ro, rd, after = rco.gen_ray(c, px, py, seed)
        (u1, u2), s2 = rco.rand_stream(seed, 2)
        assert after == s2, "a pinhole ray draws exactly the two jitter values"
        assert np.array_equal(ro, np.asarray(list(c.camPos), np.float32)), "pinhole rays start at the camera"
        d = pinhole_dir(c, px, py, u1, u2)
        assert np.max(np.abs(rd - d)) < 3e-6, (rd, d)


def _focal_plane_check(with_sa):
    rng = np.random.default_rng(22 if with_sa else 23)
    for i in range(300):
        sa = float(rng.choice([0.0, rng.uniform(-3, 3)])) if with_sa else 0.0
        c = random_camera(rng, lens=float(rng.uniform(0.001, 0.5)), sa=sa)
        f_off = float(rng.uniform(-0.3, 0.3)) if i % 2 else 0.0
        px, py = int(rng.integers(0, c.width)), int(rng.integers(0, c.height))
        seed = int(rng.integers(0, 2**32))
        ro, rd, after = rco.gen_ray(c, px, py, seed, f_off)
        (u1, u2, *_), s5 = rco.rand_stream(seed, 5)
        assert after == s5, "a thin-lens ray draws 2 jitter + 3 aperture values"
        P, R, U, Fw = (cam_vec(c, k) for k in ("camPos", "camRight", "camUp", "camForward"))
        off = ro.astype(np.float64) - P
        lx, ly = off @ R, off @ U
        assert abs(off @ Fw) < 1e-5, "the lens point lies in the lens plane"
        _, s2 = rco.rand_stream(seed, 2)
        ng, _ = rco.sample_ngon(s2, c.lensRadius, c.apertureBlades)
        assert abs(lx - ng[0]) < 2e-6 and abs(ly - ng[1]) < 2e-6, "the lens point is the aperture sample"
        r = float(F32(c.lensRadius))
        rho2 = min(max((lx * lx + ly * ly) / (r * r), 0.0), 1.0) if r > 1e-6 else 0.0
        ef = max(float(F32(c.focalDist)) + f_off + sa * rho2, 0.01)
        pin = pinhole_dir(c, px, py, u1, u2)
        Q = P + pin * (ef / (pin @ Fw))             # the pinhole ray on the focal plane
        t = (ef - off @ Fw) / (rd.astype(np.float64) @ Fw)
        X = ro + t * rd.astype(np.float64)          # where this ray meets that plane
        assert np.linalg.norm(X - Q) <= 3e-5 * np.linalg.norm(Q - P) + 1e-5, (i, X, Q)


def pin_thin_lens_focus():
    _focal_plane_check(False)


def pin_spherical_aberration():
    _focal_plane_check(True)


# ---- sampleNGon ------------------------------------------------------------
def pin_ngon_uniform_area_density():
    rng = np.random.default_rng(24)
    for n in (3, 5, 6, 8):
        radius = 0.7
        seeds = rng.integers(0, 2**32, 40000)
        pts = np.array([rco.sample_ngon(int(s), radius, n)[0] for s in seeds], np.float64)
        ang = np.mod(np.arctan2(pts[:, 1], pts[:, 0]), 2 * np.pi)
        k = np.minimum((ang / (2 * np.pi / n)).astype(int), n - 1)
        a0, a1 = 2 * np.pi * k / n, 2 * np.pi * (k + 1) / n
        v0 = radius * np.stack([np.cos(a0), np.sin(a0)], 1)
        v1 = radius * np.stack([np.cos(a1), np.sin(a1)], 1)
        det = v0[:, 0] * v1[:, 1] - v0[:, 1] * v1[:, 0]
        a = (pts[:, 0] * v1[:, 1] - pts[:, 1] * v1[:, 0]) / det
        b = (v0[:, 0] * pts[:, 1] - v0[:, 1] * pts[:, 0]) / det
        assert np.all(a >= -1e-5) and np.all(b >= -1e-5) and np.all(a + b <= 1 + 1e-5), "outside the polygon"
        s = np.clip(a + b, 1e-12, 1.0)
        for name, v, bins in (("sector", k, np.arange(n + 1) - 0.5), ("radial", s * s, np.linspace(0, 1, 11)),
                              ("angular", np.clip(a / s, 0, 1), np.linspace(0, 1, 11))):
            cnt, _ = np.histogram(v, bins)
            p = stats.chisquare(cnt).pvalue
            assert p > 1e-4, (n, name, cnt, p)


# ---- LCA ---------------------------------------------------------------------
def _lca_scene():
    b = SceneBuilder()
    m = b.add_material(Material.emissive(0.5, 1.0, 1.5, 2.0))  # radiance (1, 2, 3): channels distinguishable
    s, D = 0.01, 5.0
    b.add_quad((-s, -s, -D), (s, -s, -D), (s, s, -D), (-s, s, -D), m)
    return b.build()


def pin_lca_sharp_channel():
    sc = _lca_scene()
    params = N.RcPtParams(0, 1e30, 1)
    for lca, sharp in (((0.5, 0.0, -0.5), 1), ((0.0, 0.5, -0.5), 0), ((0.5, -0.5, 0.0), 2)):
        c = make_cam((0, 0, 0), (0, 0, -1), (1, 0, 0), (0, 1, 0), 1e-4, 1, 1, focal=5.0, lens=0.4, blades=6,
                     lca=lca)
        o = rco.OracleRenderer(sc, c, params)
        for _ in range(48):
            o.render_pass()
        mean = o.accum[0, :3] / o.accum[0, 3]
        E = (1.0, 2.0, 3.0)
        assert mean[sharp] == E[sharp], (lca, mean)   # every sample of the focused channel hits the emitter
        for ch in range(3):
            if ch != sharp:
                assert mean[ch] < 0.5 * E[ch], (lca, mean)


# ---- GGX ---------------------------------------------------------------------
def ggx_vndf_pdf(h, wo, a):
    """Heitz 2018: D_wo(h) = G1(wo) max(0, wo.h) D(h) / (wo.n), local frame with y = n."""
    hy = h[..., 1]
    D = a * a / (np.pi * (hy * hy * (a * a - 1) + 1) ** 2)
    g1 = 2 * wo[1] / (wo[1] + np.sqrt(a * a + (1 - a * a) * wo[1] ** 2))
    return g1 * np.maximum(0.0, h @ wo) * D / wo[1]


def pin_vndf_matches_visible_normal_density():
    rng = np.random.default_rng(25)
    for alpha, theta_o in ((0.25, 0.3), (0.5, 1.0), (0.9, 1.35), (0.5, 0.0)):
        wo = np.array([np.sin(theta_o), np.cos(theta_o), 0.0])
        u = rng.random((40000, 2)).astype(np.float32)
        h = np.array([rco.sample_vndf(wo.astype(np.float32), alpha, a, b) for a, b in u], np.float64)
        assert np.all(h[:, 1] >= 0) and np.allclose(np.linalg.norm(h, axis=1), 1, atol=1e-5)
        qmap = lambda hy: (lambda t2: t2 / (alpha * alpha + t2))((1 - hy * hy) / np.maximum(hy * hy, 1e-300))
        nq, nphi = 8, 8
        qs = np.minimum((qmap(h[:, 1]) * nq).astype(int), nq - 1)
        ph = np.minimum((np.mod(np.arctan2(h[:, 2], h[:, 0]), 2 * np.pi) / (2 * np.pi) * nphi).astype(int), nphi - 1)
        obs = np.bincount(qs * nphi + ph, minlength=nq * nphi).astype(np.float64)
        # expected bin masses: midpoint rule on a fine (mu, phi) grid
        mu = (np.arange(6000) + 0.5) / 6000
        phi = (np.arange(720) + 0.5) / 720 * 2 * np.pi
        M, P = np.meshgrid(mu, phi, indexing="ij")
        s = np.sqrt(1 - M * M)
        H = np.stack([s * np.cos(P), M, s * np.sin(P)], -1)
        dens = ggx_vndf_pdf(H, wo, alpha) * (1 / 6000) * (2 * np.pi / 720)
        assert abs(dens.sum() - 1) < 2e-3, dens.sum()
        gq = np.minimum((qmap(M) * nq).astype(int), nq - 1)
        gp = np.minimum((P / (2 * np.pi) * nphi).astype(int), nphi - 1)
        exp = np.bincount((gq * nphi + gp).ravel(), weights=dens.ravel(), minlength=nq * nphi)
        exp = exp / exp.sum() * len(h)
        keep = exp >= 20
        o_, e_ = np.append(obs[keep], obs[~keep].sum()), np.append(exp[keep], exp[~keep].sum())
        if e_[-1] == 0:
            o_, e_ = o_[:-1], e_[:-1]
        p = stats.chisquare(o_, e_ * o_.sum() / e_.sum()).pvalue
        assert p > 1e-4, (alpha, theta_o, p)


def _lam(mu, a):
    t2 = (1 - mu * mu) / (mu * mu)
    return (np.sqrt(1 + a * a * t2) - 1) / 2


def pin_smith_closed_forms():
    for a in (0.001, 0.05, 0.3, 0.7, 1.0):
        for nv in (0.02, 0.2, 0.5, 0.9, 1.0):
            g1 = 1 / (1 + _lam(nv, a))
            assert abs(rco.smith_g1(nv, a) - g1) <= 1e-5 * g1, (a, nv)
            for nl in (0.03, 0.4, 0.8, 1.0):
                g2 = 1 / (1 + _lam(nv, a) + _lam(nl, a))
                assert abs(rco.smith_g2(nl, nv, a) - g2) <= 1e-5 * g2, (a, nv, nl)


def _metal(rough, albedo):
    return np.array([albedo[0], albedo[1], albedo[2], 1.0, 0, 0, 0, 0, rough, 1.5, 0.0, 0, 0, 0, 0, 0], np.float32)


def pin_metal_weight_and_closed_form():
    rng = np.random.default_rng(26)
    checked = 0
    for _ in range(3000):
        gn = unit(rng.standard_normal(3)).astype(np.float32)
        rd = unit(rng.standard_normal(3)).astype(np.float32)
        albedo = rng.uniform(0, 1, 3)
        rough = float(rng.choice([0.0, rng.uniform(0, 1)]))
        pos = rng.uniform(-3, 3, 3).astype(np.float32)
        ro, d, thr, _, _ = rco.scatter(_metal(rough, albedo), rd, gn, pos, int(rng.integers(0, 2**32)))
        n = gn.astype(np.float64) if rd @ gn < 0 else -gn.astype(np.float64)
        wo = -rd.astype(np.float64)
        assert np.all(thr >= 0) and np.all(thr <= 1 + 1e-6), thr      # F * G2 / G1 in [0, 1]
        assert d @ n > 0, "the reflection leaves above the surface"
        assert np.max(np.abs(ro - (pos + n * 2e-3))) < 1e-5
        mirror = wo - 2 * (wo @ n) * n
        mirror = -mirror  # reflect(-wo, n)
        if np.max(np.abs(d - mirror)) < 1e-5 or rough == 0.0:
            continue  # the below-surface fallback or a near-delta lobe: no half vector to recover
        h = unit(wo + d)
        a = max(rough, 1e-3)
        F = albedo + (1 - albedo) * (1 - max(h @ wo, 1e-4)) ** 5
        nl, nv = max(n @ d, 1e-4), max(n @ wo, 1e-4)
        g = (1 / (1 + _lam(nv, a) + _lam(nl, a))) * (1 + _lam(nv, a))
        exp = F * g
        assert np.max(np.abs(thr - exp)) <= 2e-3 * max(1e-3, exp.max()) + 2e-5, (thr, exp, rough)
        checked += 1
    assert checked > 1000


def pin_metal_alpha_is_roughness():
    # Normal incidence: the visible-normal density is D(h) cos(theta_h), so
    # P(tan^2 theta_h <= T) = T / (a^2 + T); reflections that would leave below
    # the surface (theta_h > 45 deg, mass a^2 / (a^2 + 1)) fall back to the mirror.
    rng = np.random.default_rng(27)
    gn = np.array([0, 0, 1], np.float32)
    rd = np.array([0, 0, -1], np.float32)
    for rough in (0.15, 0.4):
        th = []
        for s in rng.integers(0, 2**32, 20000):
            _, d, _, _, _ = rco.scatter(_metal(rough, (1, 1, 1)), rd, gn, np.zeros(3, np.float32), int(s))
            th.append(0.5 * np.arccos(np.clip(d[2], -1, 1)))
        t2 = np.tan(np.array(th)) ** 2
        a2 = rough * rough
        for T in (0.25 * a2, a2, 4 * a2, 0.5):
            exp = T / (a2 + T) + a2 / (a2 + 1)
            assert abs(np.mean(t2 <= T) - exp) < 0.015, (rough, T, np.mean(t2 <= T), exp)


def pin_cosine_hemisphere():
    rng = np.random.default_rng(28)
    n = unit(rng.standard_normal(3)).astype(np.float32)
    t = unit(np.cross(n, [0.3, 0.5, 0.7]))
    b = np.cross(n, t)
    u = rng.random((40000, 2)).astype(np.float32)
    d = np.array([rco.sample_cosine(a, c, n) for a, c in u], np.float64)
    c = d @ n
    assert np.all(c >= -1e-6)
    phi = np.mod(np.arctan2(d @ b, d @ t), 2 * np.pi)
    for v, lo, hi in ((c * c, 0, 1), (phi, 0, 2 * np.pi)):
        cnt, _ = np.histogram(v, np.linspace(lo, hi, 11))
        assert stats.chisquare(cnt).pvalue > 1e-4, cnt


# ---- glass -------------------------------------------------------------------
def _glass(ior, albedo=(1.0, 1.0, 1.0)):
    return np.array([albedo[0], albedo[1], albedo[2], 0.0, 0, 0, 0, 0, 0.0, ior, 1.0, 0, 0, 0, 0, 0], np.float32)


def _seed_with_draw(rng, lo, hi):
    while True:
        s = int(rng.integers(0, 2**32))
        (u,), _ = rco.rand_stream(s, 1)
        if lo <= u < hi:
            return s


def pin_glass_ior_one_passes_through():
    rng = np.random.default_rng(29)
    for _ in range(500):
        gn = unit(rng.standard_normal(3)).astype(np.float32)
        rd = unit(rng.standard_normal(3)).astype(np.float32)
        albedo = rng.uniform(0.2, 1.0, 3)
        cosI = abs(float(rd.astype(np.float64) @ gn))
        Fx = (1 - cosI) ** 5                      # Schlick with r0 = 0
        seed = _seed_with_draw(rng, min(Fx + 1e-6, 1.0), 1.0)
        pos = rng.uniform(-2, 2, 3).astype(np.float32)
        ro, d, thr, F, _ = rco.scatter(_glass(1.0, albedo), rd, gn, pos, seed)
        # refract(I, N, 1) forms sqrt(1 - (1 - cos^2)): float rounding of 1 - cos^2
        # leaves up to ~1e-7 / cos in the tangential part
        assert np.max(np.abs(d - rd)) < 2e-6 + 2e-7 / max(cosI, 1e-3), (d, rd)
        assert np.array_equal(thr, albedo.astype(np.float32))
        assert np.max(np.abs(ro - (pos + d * 2e-3))) < 1e-6


def pin_glass_fresnel_schlick():
    rng = np.random.default_rng(30)
    for ior in (1.2, 1.5, 2.4):
        r0 = ((1 - ior) / (1 + ior)) ** 2
        gn = np.array([0, 1, 0], np.float32)
        _, _, _, F, _ = rco.scatter(_glass(ior), np.array([0, -1, 0], np.float32), gn, np.zeros(3, np.float32), 1)
        assert abs(F - r0) < 1e-7, (ior, F, r0)   # normal incidence: F = r0 (0.04 at 1.5)
        for _ in range(100):
            th = rng.uniform(0, np.pi / 2)
            rd = np.array([np.sin(th), -np.cos(th), 0], np.float32)
            _, _, _, F, _ = rco.scatter(_glass(ior), rd, gn, np.zeros(3, np.float32), 1)
            cosI = abs(float(rd.astype(np.float64) @ gn))
            assert abs(F - (r0 + (1 - r0) * (1 - cosI) ** 5)) < 2e-6


def pin_glass_snell_and_total_internal_reflection():
    rng = np.random.default_rng(31)
    ior = 1.5
    for _ in range(600):
        gn = unit(rng.standard_normal(3)).astype(np.float32)
        rd = unit(rng.standard_normal(3)).astype(np.float32)
        g, r = gn.astype(np.float64), rd.astype(np.float64)
        entering = r @ g < 0
        n = g if entering else -g
        eta = 1 / ior if entering else ior
        sin_i = np.linalg.norm(np.cross(n, r))
        mirror = r - 2 * (r @ n) * n
        if eta * sin_i > 1.0 + 1e-6:                  # beyond the critical angle
            for _ in range(4):
                _, d, _, _, _ = rco.scatter(_glass(ior), rd, gn, np.zeros(3, np.float32), int(rng.integers(0, 2**32)))
                assert np.max(np.abs(d - mirror)) < 2e-6, "total internal reflection"
            continue
        if eta * sin_i > 1.0 - 1e-4:
            continue
        cosI = abs(r @ n)
        r0 = ((1 - ior) / (1 + ior)) ** 2
        Fx = r0 + (1 - r0) * (1 - cosI) ** 5
        seed = _seed_with_draw(rng, min(Fx + 1e-5, 1.0), 1.0)
        _, d, _, _, _ = rco.scatter(_glass(ior), rd, gn, np.zeros(3, np.float32), seed)
        d = d.astype(np.float64)
        assert d @ n < 0, "the refracted ray continues through the surface"
        assert abs(np.linalg.norm(np.cross(n, d)) - eta * sin_i) < 3e-6, "Snell's law"
        assert abs(np.linalg.det(np.stack([r, n, d]))) < 3e-6, "in the plane of incidence"
        if Fx < 0.999:
            seed = _seed_with_draw(rng, 0.0, max(Fx - 1e-5, 0.0)) if Fx > 2e-5 else None
            if seed is not None:
                _, d, _, _, _ = rco.scatter(_glass(ior), rd, gn, np.zeros(3, np.float32), seed)
                assert np.max(np.abs(d - mirror)) < 2e-6, "a draw below F reflects"


def pin_glass_reflects_with_probability_F():
    rng = np.random.default_rng(32)
    gn = np.array([0, 1, 0], np.float32)
    for th in (0.3, 1.2, 1.45):
        rd = np.array([np.sin(th), -np.cos(th), 0], np.float32)
        mirror = np.array([np.sin(th), np.cos(th), 0])
        refl = 0
        n = 6000
        for s in rng.integers(0, 2**32, n):
            _, d, _, F, _ = rco.scatter(_glass(1.5), rd, gn, np.zeros(3, np.float32), int(s))
            refl += np.max(np.abs(d - mirror)) < 1e-5
        sd = np.sqrt(F * (1 - F) / n)
        assert abs(refl / n - F) < 5 * sd + 1e-3, (th, refl / n, F)


PINS = {f.__name__: f for f in (
    pin_pinhole_closed_form, pin_thin_lens_focus, pin_spherical_aberration, pin_ngon_uniform_area_density,
    pin_lca_sharp_channel, pin_vndf_matches_visible_normal_density, pin_smith_closed_forms,
    pin_metal_weight_and_closed_form, pin_metal_alpha_is_roughness, pin_cosine_hemisphere,
    pin_glass_ior_one_passes_through, pin_glass_fresnel_schlick, pin_glass_snell_and_total_internal_reflection,
    pin_glass_reflects_with_probability_F)}


@pytest.mark.parametrize("name", list(PINS))
def test_pin(name):
    PINS[name]()


# Planted mistakes: (pin, file, exact snippet, replacement). Each must make its pin fail.
MISTAKES = [
    ("pin_pinhole_closed_form", "float ux = (((float)px + 0.5f) + jx) / (float)c->width;",
     "float ux = ((float)px + jx) / (float)c->width;"),
    ("pin_pinhole_closed_form", "float a = (sx * aspect) * c->fovTanHalf;", "float a = sx * c->fovTanHalf;"),
    ("pin_pinhole_closed_form", "float sx = ux * 2.0f - 1.0f, sy = uy * 2.0f - 1.0f;",
     "float sx = ux * 2.0f - 1.0f, sy = 1.0f - uy * 2.0f;"),
    ("pin_thin_lens_focus", "float tf = ef / dot3(pin, F);", "float tf = ef;"),
    ("pin_thin_lens_focus", "float ef = (c->focalDist + fOff) + c->saStrength * nr2;",
     "float ef = c->focalDist + c->saStrength * nr2;"),
    ("pin_thin_lens_focus", "v3 o = add(add(P, muls(R, lx)), muls(U, ly));",
     "v3 o = add(add(P, muls(R, ly)), muls(U, lx));"),
    ("pin_spherical_aberration", "nr2 = rc_clampf((lx * lx + ly * ly) / (c->lensRadius * c->lensRadius), 0.0f, 1.0f);",
     "nr2 = rc_clampf((lx * lx + ly * ly) / c->lensRadius, 0.0f, 1.0f);"),
    ("pin_spherical_aberration", "ef = fmaxf(ef, 0.01f);", "ef = ef;"),
    ("pin_ngon_uniform_area_density", "if (r1 + r2 > 1.0f) { r1 = 1.0f - r1; r2 = 1.0f - r2; }\n    int sector",
     "int sector"),
    ("pin_ngon_uniform_area_density", "int sector = (int)(rco_rand(seed) * (float)n);",
     "int sector = (int)(rco_rand(seed) * (float)(n - 1));"),
    ("pin_ngon_uniform_area_density", "*ox = (r1 * c0 + r2 * c1) * radius;",
     "*ox = (sqrtf(r1) * c0 + r2 * c1) * radius;"),
    ("pin_lca_sharp_channel", "gen_ray(c, x, y, &sr, c->lcaDelta[0], &ro, &rd);",
     "gen_ray(c, x, y, &sr, c->lcaDelta[1], &ro, &rd);"),
    ("pin_lca_sharp_channel", "float b = trace_radiance(s, p, ro, rd, &sb, &aB, &tB, st).z;",
     "float b = trace_radiance(s, p, ro, rd, &sb, &aB, &tB, st).y;"),
    ("pin_vndf_matches_visible_normal_density", "v3 v = normalize3(V(wo.x * alpha, wo.y, wo.z * alpha));",
     "v3 v = normalize3(V(wo.x, wo.y, wo.z));"),
    ("pin_vndf_matches_visible_normal_density", "float r = sqrtf(ux);\n    float phi = (uy < a)",
     "float r = ux;\n    float phi = (uy < a)"),
    ("pin_vndf_matches_visible_normal_density", "float p2 = (r * rc_sinf(phi)) * ((uy < a) ? 1.0f : v.y);",
     "float p2 = (r * rc_sinf(phi)) * ((uy < a) ? v.y : 1.0f);"),
    ("pin_smith_closed_forms", "float denom = nv + sqrtf(a2 + ((1.0f - a2) * nv) * nv);",
     "float denom = nv + sqrtf(a2 + (1.0f - a2) * nv);"),
    ("pin_smith_closed_forms", "float da = nv * sqrtf(a2 + ((1.0f - a2) * nl) * nl);",
     "float da = nl * sqrtf(a2 + ((1.0f - a2) * nl) * nl);"),
    ("pin_metal_weight_and_closed_form", "float g = g2 / fmaxf(g1, 1e-7f);", "float g = g1 / fmaxf(g2, 1e-7f);"),
    ("pin_metal_weight_and_closed_form", "float pw = rc_powf(1.0f - hv_, 5.0f);", "float pw = rc_powf(1.0f - nv, 5.0f);"),
    ("pin_metal_alpha_is_roughness", "float alpha = fmaxf(rough, 1e-3f);", "float alpha = fmaxf(rough * rough, 1e-3f);"),
    ("pin_cosine_hemisphere", "float r = sqrtf(ux);\n    float phi = 6.28318530718f * uy;",
     "float r = ux;\n    float phi = 6.28318530718f * uy;"),
    ("pin_glass_ior_one_passes_through", "float c = eta * d + sqrtf(k);", "float c = eta * d - sqrtf(k);"),
    ("pin_glass_fresnel_schlick", "    r0 = r0 * r0;\n    float F = r0", "    float F = r0"),
    ("pin_glass_snell_and_total_internal_reflection", "float eta = entering ? (1.0f / ior) : ior;",
     "float eta = entering ? ior : (1.0f / ior);"),
    ("pin_glass_snell_and_total_internal_reflection", "if (dot3(rf, rf) < 0.001f) rf = reflect3(d, n);",
     "if (dot3(rf, rf) < 0.001f) rf = d;"),
    ("pin_glass_reflects_with_probability_F", "if (rco_rand(seed) < F) {\n        d = reflect3(d, n);",
     "if (rco_rand(seed) > F) {\n        d = reflect3(d, n);"),
]


@pytest.mark.parametrize("pin,old,new", MISTAKES, ids=[f"{m[0]}-{i}" for i, m in enumerate(MISTAKES)])
def test_planted_mistakes_fail(pin, old, new):
    assert pin_fails(PINS[pin], "rc_oracle.c", old, new), f"{pin} passes with the planted mistake"
