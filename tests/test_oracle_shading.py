"""Pins of the CPU oracle's shading, sampling and light sampling (R6, R9, R12, R13).

The reference ships no outputs for these shaders, so each oracle function is checked
against what the mathematics fixes, written independently of its code:

  * GGX: D in its tan form, 1 / (pi a^2 cos^4 (1 + tan^2 / a^2)^2), integrates to one
    over the half-vector hemisphere (weighted by cos); the masking term is the
    separable Smith form 2 / (1 + sqrt(1 + a^2 tan^2)); the Fresnel term is within
    the spherical-Gaussian fit's error of Schlick's (1 - cos)^5; the BRDF is
    reciprocal; the specular lobe loses energy but never creates it;
  * sampleBRDF draws half vectors distributed as D cos (chi-square over equal-mass
    bins computed by quadrature of the tan form) with uniform azimuth, and cosine-
    weighted directions for the diffuse lobe;
  * sampleLight: the mean of the light estimate over many draws equals, for every
    light sampling mode, the closed-form unoccluded integral over the emitters -
    Lambert's formula for the cosine-weighted solid angle of a polygon, checked
    against brute-force quadrature of the same integral;
  * closest hit: the hit position is the ray-plane intersection, the interpolated
    normal goes through the inverse-transpose of the instance matrix (numpy), the
    tangent frame is orthonormal and follows +u, bilinear texture fetches return
    texels at texel centres and averages between them with repeat addressing, and
    the metallic conversion maps a non-specular material to its own albedo with
    f0 = 0.04 and a white specular material to metalness 1.
"""
import ctypes as C

import numpy as np
import pytest
from scipy import integrate, stats

from oracle import oracle
from vulkanpbrt_amd import capi


# ---------------------------------------------------------------------------
# textbook forms (independent of the oracle's code)
# ---------------------------------------------------------------------------
def ggx_d(cos_h, a):
    cos_h = np.asarray(cos_h, np.float64)
    tan2 = (1.0 - cos_h ** 2) / cos_h ** 2
    return 1.0 / (np.pi * a * a * cos_h ** 4 * (1.0 + tan2 / (a * a)) ** 2)


def smith_g1(cos_t, a):
    tan2 = (1.0 - cos_t ** 2) / cos_t ** 2
    return 2.0 / (1.0 + np.sqrt(1.0 + a * a * tan2))


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def hemisphere_dirs(rng, n, min_cos=0.05):
    d = unit(rng.normal(size=(4 * n, 3)))
    d[:, 2] = np.abs(d[:, 2])
    return d[d[:, 2] > min_cos][:n]


def specular_only(alpha, r0=1.0):
    """A surface whose BRDF is the GGX specular lobe alone: no diffuse albedo (so the
    specular weight is 1) and r90 = 1."""
    return oracle.surface(perceptual_roughness=np.sqrt(alpha), alpha_roughness=alpha, metalness=1.0,
                          r0=(r0,) * 3, r90=(1.0,) * 3, diffuse=(0.0,) * 3, specular=(1.0,) * 3)


# ---------------------------------------------------------------------------
# BRDF and its pdf
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("alpha", [0.05, 0.2, 0.5, 1.0])
def test_specular_pdf_integrates_to_one(oracle_lib, alpha):
    """pdf_brdf with specular weight 1 is the half-vector density D(h) cos(h): over the
    hemisphere of h it integrates to one."""
    s = specular_only(alpha)

    def pdf_at(theta):
        h = np.array([np.sin(theta), 0.0, np.cos(theta)])
        v = np.array([0.0, 0.0, 1.0])
        l = 2 * (v @ h) * h - v  # normalize(v + l) == h
        _, pdf, sw = oracle.brdf(s, v, l)
        assert sw[0] == 1.0
        return pdf[0]

    total, _ = integrate.quad(lambda t: pdf_at(t) * 2 * np.pi * np.sin(t), 0, np.pi / 2 - 1e-9,
                              points=[alpha / 4, alpha, 4 * alpha], limit=400, epsabs=1e-10)
    assert abs(total - 1.0) < 1e-6


def test_diffuse_pdf_is_cosine_over_pi(oracle_lib):
    s = oracle.surface(specular=(0.0,) * 3, r0=(0.0,) * 3, diffuse=(0.5, 0.4, 0.3))
    rng = np.random.default_rng(1)
    l = hemisphere_dirs(rng, 64)
    v = np.tile([0.0, 0.0, 1.0], (len(l), 1))
    _, pdf, sw = oracle.brdf(s, v, l)
    assert np.all(sw == 0.0)
    # the shaders' 1/pi literal 0.31830988618 is good to 1.2e-11
    np.testing.assert_allclose(pdf, l[:, 2] / np.pi, rtol=1e-10)


@pytest.mark.parametrize("alpha", [0.1, 0.35, 0.8])
def test_specular_lobe_is_ggx_times_smith(oracle_lib, alpha):
    """F = 1 (r0 = r90 = 1): the BRDF is D G1(l) G1(v) / (4 cos_l cos_v)."""
    s = specular_only(alpha)
    rng = np.random.default_rng(2)
    v, l = hemisphere_dirs(rng, 300), hemisphere_dirs(rng, 300)
    f, _, _ = oracle.brdf(s, v, l)
    h = unit(v + l)
    want = ggx_d(h[:, 2], alpha) * smith_g1(l[:, 2], alpha) * smith_g1(v[:, 2], alpha) / (4 * l[:, 2] * v[:, 2])
    np.testing.assert_allclose(f[:, 0], want, rtol=1e-9)
    np.testing.assert_array_equal(f[:, 0], f[:, 2])


def test_fresnel_is_schlick_within_fit_error(oracle_lib):
    """With r90 = 1 the Fresnel factor is the spherical-Gaussian fit of Schlick's
    r0 + (1 - r0)(1 - cos_vh)^5: within 1e-2 of it at every angle."""
    alpha, r0 = 0.4, 0.2
    s = specular_only(alpha, r0)
    rng = np.random.default_rng(3)
    v, l = hemisphere_dirs(rng, 500), hemisphere_dirs(rng, 500)
    f, _, _ = oracle.brdf(s, v, l)
    h = unit(v + l)
    gd = ggx_d(h[:, 2], alpha) * smith_g1(l[:, 2], alpha) * smith_g1(v[:, 2], alpha) / (4 * l[:, 2] * v[:, 2])
    fresnel = f[:, 0] / gd
    vh = np.sum(v * h, axis=1)
    schlick = r0 + (1 - r0) * (1 - vh) ** 5
    assert np.max(np.abs(fresnel - schlick)) < 1e-2
    assert np.all(fresnel >= r0 - 1e-12) and np.all(fresnel <= 1 + 1e-12)


def test_brdf_is_reciprocal(oracle_lib):
    rng = np.random.default_rng(4)
    for k in range(6):
        pr = rng.uniform(0.1, 1.0)
        s = oracle.surface(perceptual_roughness=pr, metalness=rng.uniform(0, 1), r0=rng.uniform(0.02, 1, 3),
                           r90=(1.0,) * 3, diffuse=rng.uniform(0, 1, 3), specular=rng.uniform(0, 1, 3))
        v, l = hemisphere_dirs(rng, 200), hemisphere_dirs(rng, 200)
        a, _, _ = oracle.brdf(s, v, l)
        b, _, _ = oracle.brdf(s, l, v)
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-300)


def test_diffuse_at_normal_incidence_is_lambert(oracle_lib):
    """No specular reflectance (r0 = r90 = 0): at v = l = n the Burley diffuse term
    reduces to albedo / pi, times the reference's constant 1 - 0.1."""
    alb = np.array([0.7, 0.5, 0.2])
    for pr in (0.2, 0.5, 1.0):  # alpha = 0 makes D 0/0 at the mirror direction
        s = oracle.surface(perceptual_roughness=pr, r0=(0.0,) * 3, r90=(0.0,) * 3, diffuse=alb,
                           specular=(0.0,) * 3)
        f, _, _ = oracle.brdf(s, [0, 0, 1.0], [0, 0, 1.0])
        np.testing.assert_allclose(f[0], alb * 0.9 / np.pi, rtol=1e-10)


@pytest.mark.parametrize("alpha", [0.15, 0.5, 1.0])
@pytest.mark.parametrize("cos_v", [1.0, 0.6, 0.2])
def test_specular_lobe_conserves_energy(oracle_lib, alpha, cos_v):
    """Directional albedo of the F = 1 lobe, integrated over half vectors on a fine grid
    (dw_l = 4 (v.h) dw_h): never above one, close to one for smooth surfaces at normal
    incidence."""
    s = specular_only(alpha)
    v = np.array([np.sqrt(1 - cos_v ** 2), 0.0, cos_v])
    nt, nphi = 1500, 256
    st = (np.arange(nt) + 0.5) / nt
    theta = st ** 2 * (np.pi / 2)  # dense near the normal, where D peaks
    dtheta = 2 * st * (np.pi / 2) / nt
    phi = (np.arange(nphi) + 0.5) / nphi * 2 * np.pi
    T, P = np.meshgrid(theta, phi, indexing="ij")
    h = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    vh = h @ v
    l = 2 * vh[:, None] * h - v
    ok = (vh > 0) & (l[:, 2] > 0)
    f, _, _ = oracle.brdf(s, np.tile(v, (int(ok.sum()), 1)), l[ok])
    w = (np.sin(T) * dtheta[:, None] * (2 * np.pi / nphi)).reshape(-1)[ok]
    albedo = np.sum(f[:, 0] * l[ok, 2] * 4 * vh[ok] * w)
    assert albedo <= 1.0 + 1e-3
    if alpha <= 0.2 and cos_v == 1.0:
        assert albedo > 0.95


# ---------------------------------------------------------------------------
# sampleBRDF
# ---------------------------------------------------------------------------
def chi2_pvalue(counts, probs):
    exp = probs * counts.sum()
    return stats.chisquare(counts, exp).pvalue


@pytest.mark.parametrize("alpha,cos_v", [(0.1, 1.0), (0.3, 0.7), (0.7, 0.4)])
def test_specular_samples_follow_ggx(oracle_lib, alpha, cos_v):
    s = specular_only(alpha)
    v = np.array([np.sqrt(1 - cos_v ** 2), 0.0, cos_v])
    n = 200_000
    l, pdf, _ = oracle.sample_brdf(s, v, n)
    # l is the mirror of v about the drawn h (which lies in the upper hemisphere), so
    # v + l = 2 (v.h) h: h up to the sign of v.h
    h = unit(v + l)
    h[h[:, 2] < 0] *= -1
    cos_h = h[:, 2]
    # equal-mass bins of cos(theta_h) under D cos, from quadrature of the tan form
    grid = np.linspace(0, np.pi / 2, 20001)
    dens = ggx_d(np.cos(grid[:-1] + 1e-12), alpha) * np.cos(grid[:-1]) * np.sin(grid[:-1]) * 2 * np.pi
    cdf = np.concatenate([[0], np.cumsum(dens * np.diff(grid))])
    cdf /= cdf[-1]
    edges_theta = np.interp(np.linspace(0, 1, 21), cdf, grid)
    counts, _ = np.histogram(np.arccos(np.clip(cos_h, -1, 1)), bins=edges_theta)
    assert chi2_pvalue(counts, np.full(20, 1 / 20)) > 1e-4
    # azimuth of h uniform
    az = np.arctan2(h[:, 1], h[:, 0])
    counts, _ = np.histogram(az, bins=16, range=(-np.pi, np.pi))
    assert chi2_pvalue(counts, np.full(16, 1 / 16)) > 1e-4
    # the returned pdf is the half-vector density of the drawn h
    np.testing.assert_allclose(pdf, ggx_d(cos_h, alpha) * cos_h, rtol=1e-6)


def test_diffuse_samples_are_cosine_weighted(oracle_lib):
    s = oracle.surface(specular=(0.0,) * 3, r0=(0.0,) * 3, diffuse=(0.6,) * 3)
    l, pdf, _ = oracle.sample_brdf(s, np.array([0.3, 0.1, 0.95]) / np.linalg.norm([0.3, 0.1, 0.95]), 200_000)
    # cos^2(theta) of a cosine-weighted direction is uniform on [0, 1]
    counts, _ = np.histogram(l[:, 2] ** 2, bins=20, range=(0, 1))
    assert chi2_pvalue(counts, np.full(20, 1 / 20)) > 1e-4
    counts, _ = np.histogram(np.arctan2(l[:, 1], l[:, 0]), bins=16, range=(-np.pi, np.pi))
    assert chi2_pvalue(counts, np.full(16, 1 / 16)) > 1e-4
    np.testing.assert_allclose(pdf, l[:, 2] / np.pi, rtol=1e-9)


# ---------------------------------------------------------------------------
# sampleLight against the unoccluded emitter integral
# ---------------------------------------------------------------------------
class Desc:
    """A pbrt_scene_desc built from Python arrays (kept alive with it)."""

    def __init__(self, meshes, lights=(), textures=()):
        self.keep = []
        md = (capi.MeshDesc * len(meshes))()
        for i, m in enumerate(meshes):
            pos = np.ascontiguousarray(m["positions"], np.float32)
            idx = np.ascontiguousarray(m["indices"], np.uint32)
            nor = None if m.get("normals") is None else np.ascontiguousarray(m["normals"], np.float32)
            uv = None if m.get("uvs") is None else np.ascontiguousarray(m["uvs"], np.float32)
            self.keep += [pos, idx, nor, uv]
            mat = capi.Material()
            mat.diffuse_ior[:] = list(m.get("diffuse", (0.8, 0.8, 0.8))) + [1.0]
            mat.specular_dissolve[:] = list(m.get("specular", (0.0, 0.0, 0.0))) + [1.0]
            mat.ambient_roughness[:] = [0.0, 0.0, 0.0, m.get("roughness", 0.5)]
            mat.emission_texture_id[:] = list(m.get("emission", (0.0, 0.0, 0.0))) + [0.0]
            md[i] = capi.MeshDesc(len(pos), pos.ctypes.data, None if nor is None else nor.ctypes.data,
                                  None if uv is None else uv.ctypes.data, idx.size, 4, idx.ctypes.data, mat,
                                  m.get("diffuse_texture", -1), -1, -1, -1)
        inst = (capi.InstanceDesc * len(meshes))()
        for i, m in enumerate(meshes):
            t = np.asarray(m.get("transform", np.eye(4)), np.float32)
            inst[i].transform[:] = t.T.ravel().tolist()  # column-major
            inst[i].mesh = i
        tex = (capi.TextureDesc * max(1, len(textures)))()
        for i, t in enumerate(textures):
            a = np.ascontiguousarray(t, np.uint8)
            self.keep.append(a)
            tex[i] = capi.TextureDesc(a.shape[1], a.shape[0], a.ctypes.data)
        li = (capi.Light * max(1, len(lights)))()
        for i, x in enumerate(lights):
            li[i] = x
        self.keep += [md, inst, tex, li]
        self.desc = capi.SceneDesc(len(meshes), md, len(meshes), inst, len(textures), tex, len(lights), li, 0, 0)


def projected_solid_angle_lambert(poly, p, n):
    """Lambert's formula: (1/2) sum over edges of the edge's angle at p times the cosine
    between n and the edge plane's normal (polygon wholly above p's horizon)."""
    r = unit(np.asarray(poly, np.float64) - p)
    total = 0.0
    for i in range(len(r)):
        a, b = r[i], r[(i + 1) % len(r)]
        g = np.cross(a, b)
        total += np.arccos(np.clip(a @ b, -1, 1)) * (g / np.linalg.norm(g)) @ n
    return abs(total) / 2


def projected_solid_angle_quadrature(tri, p, n, m=400):
    """The same integral, sum of cos_p cos_l / d^2 dA over a fine barycentric grid."""
    a, b, c = (np.asarray(x, np.float64) for x in tri)
    ln = unit(np.cross(b - a, c - a))
    area = 0.5 * np.linalg.norm(np.cross(b - a, c - a))
    i, j = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    keep = i + j < m
    pts = []
    for (u, v) in ((i[keep] + 1 / 3, j[keep] + 1 / 3), (i[keep & (i + j < m - 1)] + 2 / 3,
                                                       j[keep & (i + j < m - 1)] + 2 / 3)):
        pts.append(a + np.outer(u / m, b - a) + np.outer(v / m, c - a))
    q = np.concatenate(pts)
    d = q - p
    dist = np.linalg.norm(d, axis=1)
    w = d / dist[:, None]
    f = np.clip(w @ n, 0, None) * np.clip(-(w @ ln), 0, None) / dist ** 2
    return f.mean() * area


def emitter_scene():
    """Three emissive triangles of different size, emission and tilt above the origin."""
    tris = [
        ([[-0.6, 0.2, 2.0], [0.5, 0.1, 2.2], [0.1, 0.9, 1.9]], (4.0, 3.0, 2.0)),
        ([[1.0, -0.5, 1.2], [1.6, -0.4, 1.6], [1.2, 0.3, 1.4]], (1.0, 1.5, 0.5)),
        ([[-1.5, -1.0, 0.8], [-0.9, -1.4, 1.1], [-1.0, -0.6, 1.5]], (0.3, 0.2, 6.0)),
    ]
    meshes = []
    for pts, em in tris:
        p = np.asarray(pts, np.float64)
        # wind so the emitting side faces the receiver at the origin
        nrm = np.cross(p[1] - p[0], p[2] - p[0])
        if nrm @ (-p.mean(0)) < 0:
            p = p[[0, 2, 1]]
        meshes.append(dict(positions=p, indices=[0, 1, 2], emission=em))
    return meshes


def light_strength_selection(prefix):
    """Which light sampleLight's LIGHT_STRENGTH search (lighting.glsl:103-111) ends on, for
    three lights with strength prefix sums P0 < P1 < P2 = S: the search sets begin = mid
    (not mid + 1) when the drawn strength exceeds P[mid], and never narrows a two-wide
    range when it does not, leaving at the iteration cap.  From (begin, end) = (0, 3):
    pk <= P1 ends on light 0 and pk > P1 on light 1; light 2 is never picked.  Returns
    the selection probabilities."""
    p0, p1, s = prefix
    return np.array([p1 / s, 1 - p1 / s, 0.0])


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_light_sampling_estimates_the_emitter_integral(oracle_lib, mode):
    meshes = emitter_scene()
    d = Desc(meshes)
    sc = oracle.OracleScene(d.desc)
    assert sc.light_count == 3
    p = np.array([0.05, -0.1, 0.0])
    n = unit([0.15, -0.1, 1.0])
    want = np.zeros(3)
    power = np.array([3 * sum(m["emission"]) for m in meshes])  # ambient + diffuse + specular
    prefix = np.cumsum(power)
    np.testing.assert_allclose(sc.lights()[:, 31], prefix, rtol=1e-6)
    select = light_strength_selection(prefix)
    for k, m in enumerate(meshes):
        tri = np.asarray(m["positions"], np.float64)
        lam = projected_solid_angle_lambert(tri, p, n)
        quad = projected_solid_angle_quadrature(tri, p, n)
        assert abs(lam - quad) < 2e-4 * lam  # the two closed forms agree
        area = 0.5 * np.linalg.norm(np.cross(tri[1] - tri[0], tri[2] - tri[0]))
        # emission enters three times (ambient + diffuse + specular colour of the light);
        # SURFACE_STRENGTH's estimate omits the triangle area
        em = 3 * np.asarray(m["emission"])
        if mode == 1:
            # estimate of light k: its term times S / strength_k, picked with select[k]
            want += select[k] * em * lam * prefix[-1] / power[k]
        else:
            want += em * lam / (area if mode == 0 else 1.0)
    nsamp = 400_000
    r = sc.sample_light(mode, p, n, nsamp, frame=7)
    assert r["ok"].mean() > 0.99
    est = np.where(r["ok"][:, None], r["col"] / np.where(r["ok"], r["pdf"], 1)[:, None], 0.0)
    mean, se = est.mean(0), est.std(0) / np.sqrt(nsamp)
    print("mode", mode, "estimate", mean, "+-", se, "closed form", want)
    assert np.all(np.abs(mean - want) < 5 * se + 1e-3 * want)
    # the shadow-ray direction points at the sampled emitter point, tmax short of it
    assert np.allclose(np.linalg.norm(r["l"][r["ok"]], axis=1), 1.0, atol=1e-9)


def test_point_and_directional_lights(oracle_lib):
    """Explicit lights: a point light attenuated by 1/(c0 + c1 d + c2 d^2) and a
    directional light, both estimated exactly (no randomness in their terms) by the
    uniform and light-strength modes, whose mean is their sum."""
    floor = dict(positions=[[-5, -5, -1], [5, -5, -1], [5, 5, -1]], indices=[0, 1, 2])
    pl = capi.Light()
    pl.v0_type[:] = [1.0, 2.0, 3.0, 2.0]
    pl.col_diffuse[:] = [2.0, 1.0, 0.5, 0.0]
    pl.strengths[:] = [1.0, 0.5, 0.25, 0.0]
    dl = capi.Light()
    dl.v0_type[:] = [0.0, 0.0, 0.0, 1.0]
    dl.dir_angle2[:] = list(-unit([0.3, 0.2, 1.0])) + [0.0]
    dl.col_ambient[:] = [0.5, 0.5, 0.5, 0.0]
    dl.strengths[:] = [1.0, 0.0, 0.0, 0.0]
    d = Desc([floor], lights=[pl, dl])
    sc = oracle.OracleScene(d.desc)
    p, n = np.array([0.2, -0.3, 0.0]), unit([0.1, 0.0, 1.0])
    dist = np.linalg.norm(np.array([1.0, 2.0, 3.0]) - p)
    to = unit(np.array([1.0, 2.0, 3.0]) - p)
    point = np.array([2.0, 1.0, 0.5]) * max(n @ to, 0) / (1 + 0.5 * dist + 0.25 * dist ** 2)
    sun = np.array([0.5, 0.5, 0.5]) * max(n @ unit([0.3, 0.2, 1.0]), 0)
    # uniform: each light's term times the light count, picked with probability 1/2
    r = sc.sample_light(2, p, n, 100_000, frame=3)
    est = r["col"] / r["pdf"][:, None]
    np.testing.assert_allclose(est.mean(0), point + sun, rtol=2e-2)
    # light strength, two lights: the search of lighting.glsl:103-111 starts with
    # (begin, end) = (0, 2) and never narrows it, so light 0 (the point light, power
    # 3.5 of 5) is always picked and weighted by 5 / 3.5 (a bias of the reference)
    r = sc.sample_light(1, p, n, 1000, frame=3)
    np.testing.assert_allclose(r["col"] / r["pdf"][:, None], np.tile(point * 5 / 3.5, (1000, 1)), rtol=1e-9)


# ---------------------------------------------------------------------------
# closest hit: the surface record
# ---------------------------------------------------------------------------
def test_hit_position_and_normal_under_instance_transform(oracle_lib):
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float64)
    nor = unit(np.array([[0.2, 0.0, 1.0], [0.0, 0.3, 1.0], [-0.1, -0.2, 1.0]]))
    ang = 0.4
    rot = np.array([[np.cos(ang), -np.sin(ang), 0], [np.sin(ang), np.cos(ang), 0], [0, 0, 1]])
    m3 = rot @ np.diag([2.0, 0.5, 1.5])
    xf = np.eye(4)
    xf[:3, :3] = m3
    xf[:3, 3] = [0.3, -0.2, 1.0]
    d = Desc([dict(positions=pos, normals=nor, indices=[0, 1, 2], transform=xf)])
    sc = oracle.OracleScene(d.desc)
    ninv = np.linalg.inv(m3).T
    for bary in ([0.2, 0.3, 0.5], [0.6, 0.2, 0.2], [0.1, 0.1, 0.8]):
        b = np.asarray(bary)
        target = (xf @ np.append(b @ pos, 1.0))[:3]
        origin = target + np.array([0.1, 0.2, 3.0])
        s = sc.shade(origin, unit(target - origin))
        assert s is not None
        np.testing.assert_allclose(s["position"], target, atol=1e-5)
        want_n = unit(ninv @ unit(b @ nor))
        np.testing.assert_allclose(s["normal"], want_n, atol=1e-5)
        basis = s["basis"]
        np.testing.assert_allclose(basis @ basis.T, np.eye(3), atol=1e-5)
        np.testing.assert_allclose(basis[2], s["normal"], atol=1e-6)


def quad_mesh(**kw):
    """The unit square at z = 0 facing +z, uv = (x, y)."""
    return dict(positions=[[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], normals=[[0, 0, 1]] * 4,
                uvs=[[0, 0], [1, 0], [1, 1], [0, 1]], indices=[0, 1, 2, 0, 2, 3], **kw)


def test_tangent_follows_u(oracle_lib):
    d = Desc([quad_mesh()])
    sc = oracle.OracleScene(d.desc)
    s = sc.shade([0.3, 0.6, 2.0], [0, 0, -1.0])
    np.testing.assert_allclose(s["basis"][0], [1, 0, 0], atol=1e-6)
    np.testing.assert_allclose(np.abs(s["basis"][1]), [0, 1, 0], atol=1e-6)


def srgb_to_linear(c):
    return np.asarray(c, np.float64) ** 2.2


def test_bilinear_texture_fetch(oracle_lib):
    tex = np.array([[[255, 0, 0, 255], [0, 255, 0, 255]],
                    [[0, 0, 255, 255], [255, 255, 255, 255]]], np.uint8)
    d = Desc([quad_mesh(diffuse=(1.0, 1.0, 1.0), diffuse_texture=0)], textures=[tex])
    sc = oracle.OracleScene(d.desc)
    texel = tex[..., :3].astype(np.float64) / 255

    def diffuse_at(u, v):
        return sc.shade([u, v, 1.0], [0, 0, -1.0])["diffuse"]

    # texel centres of a 2x2 texture sit at 0.25 and 0.75; row 0 is v in [0, 0.5)
    for (u, v, want) in ((0.25, 0.25, texel[0, 0]), (0.75, 0.25, texel[0, 1]), (0.25, 0.75, texel[1, 0]),
                         (0.75, 0.75, texel[1, 1])):
        np.testing.assert_allclose(diffuse_at(u, v), srgb_to_linear(want), atol=1e-5)
    # halfway between two centres: their average, linearised after filtering
    np.testing.assert_allclose(diffuse_at(0.5, 0.25), srgb_to_linear((texel[0, 0] + texel[0, 1]) / 2), atol=1e-5)
    # repeat addressing: at u = 0.05 the left neighbour is the rightmost texel
    w = 0.05 * 2 - 0.5 + 1.0  # weight of texel x = 0 (x0 = -1 wraps to 1)
    np.testing.assert_allclose(diffuse_at(0.05, 0.25),
                               srgb_to_linear((1 - w) * texel[0, 1] + w * texel[0, 0]), atol=1e-5)


def test_metallic_conversion_closed_forms(oracle_lib):
    # a non-specular material keeps its albedo, gets the dielectric f0 = 0.04 and r90 = 1
    alb = np.array([0.6, 0.3, 0.1])
    d = Desc([quad_mesh(diffuse=alb, specular=(0.0, 0.0, 0.0), roughness=0.7)])
    s = oracle.OracleScene(d.desc).shade([0.4, 0.4, 1.0], [0, 0, -1.0])
    np.testing.assert_allclose(s["diffuse"], alb, rtol=1e-6)
    np.testing.assert_allclose(s["r0"], [0.04] * 3, rtol=1e-6)
    assert s["metalness"] == 0.0 and abs(s["r90"] - 1.0) < 1e-12
    assert abs(s["perceptual_roughness"] - 0.7) < 1e-7 and abs(s["alpha_roughness"] - 0.49) < 1e-6
    # a white specular material without albedo is a metal reflecting white
    d = Desc([quad_mesh(diffuse=(0.0, 0.0, 0.0), specular=(1.0, 1.0, 1.0))])
    s = oracle.OracleScene(d.desc).shade([0.4, 0.4, 1.0], [0, 0, -1.0])
    assert abs(s["metalness"] - 1.0) < 1e-6
    np.testing.assert_allclose(s["r0"], [1.0] * 3, rtol=1e-6)
    np.testing.assert_allclose(s["diffuse"], [0.0] * 3, atol=1e-9)


def test_emission_only_from_the_front(oracle_lib):
    d = Desc([quad_mesh(emission=(2.0, 1.0, 0.5))])
    sc = oracle.OracleScene(d.desc)
    front = sc.shade([0.5, 0.5, 1.0], [0, 0, -1.0])
    back = sc.shade([0.5, 0.5, -1.0], [0, 0, 1.0])
    np.testing.assert_allclose(front["emissive"], [2.0, 1.0, 0.5], rtol=1e-7)
    np.testing.assert_array_equal(back["emissive"], [0.0, 0.0, 0.0])


# ---------------------------------------------------------------------------
# the path: next-event estimation with MIS and the first bounce (R10, R11)
# ---------------------------------------------------------------------------
def sky(d):
    """The sky colour of lighting.glsl:231-240 (the environment a missed ray sees)."""
    up = np.array([0.3, 0.5, 0.92]) ** 2.2
    up = 1.0 + (up - 1.0) * max(d[2], 0.0)
    t = np.clip((d[2] + 0.02) / 0.04, 0, 1)
    w = t * t * (3 - 2 * t)
    return 0.2 + (up - 0.2) * w


def sky_many(d):
    return np.stack([sky(x) for x in d]) if len(d) else np.zeros((0, 3))


class FloorUnderSun:
    """A large rough floor (z = 0) lit by one directional light; every pixel of the
    frame traces the same camera ray (the projection ignores the pixel position), so
    each pixel of each frame is one independent sample of the same path."""

    light_dir = unit([0.3, 0.2, 1.0])  # towards the light
    light_col = np.array([6.0, 4.8, 3.6])  # bright enough that NEE outweighs any negative bounce term
    cam = np.array([0.3, -0.2, 3.0])
    view_dir = unit([0.1, 0.2, -1.0])

    def __init__(self, occluder=False):
        floor = dict(positions=[[-50, -50, 0], [50, -50, 0], [50, 50, 0], [-50, 50, 0]], normals=[[0, 0, 1]] * 4,
                     uvs=[[0, 0], [1, 0], [1, 1], [0, 1]], indices=[0, 1, 2, 0, 2, 3], diffuse=(0.6, 0.5, 0.4),
                     roughness=0.6)
        meshes = [floor]
        self.hit = self.cam + self.view_dir * (-self.cam[2] / self.view_dir[2])
        if occluder:
            c = self.hit + self.light_dir * 1.0
            q = 0.1
            meshes.append(dict(positions=[c + [-q, -q, 0], c + [q, -q, 0], c + [q, q, 0], c + [-q, q, 0]],
                               indices=[0, 1, 2, 0, 2, 3]))
        sun = capi.Light()
        sun.v0_type[:] = [0.0, 0.0, 0.0, 1.0]
        sun.dir_angle2[:] = list(-self.light_dir) + [0.0]
        sun.col_diffuse[:] = list(self.light_col) + [0.0]
        sun.strengths[:] = [1.0, 0.0, 0.0, 0.0]
        self.d = Desc(meshes, lights=[sun])
        self.sc = oracle.OracleScene(self.d.desc)
        z = self.view_dir
        x = unit(np.cross(z, [0.0, 0.0, 1.0]))
        y = np.cross(z, x)
        m = np.eye(4)
        m[:3, 0], m[:3, 1], m[:3, 2], m[:3, 3] = x, y, z, self.cam
        self.inv_view = m.T.ravel().astype(np.float32)  # column-major
        p = np.zeros((4, 4))
        p[2, 2] = p[3, 3] = 1.0  # every pixel maps to view direction +z
        self.inv_proj = p.T.ravel().astype(np.float32)

    def surface(self):
        s = self.sc.shade(self.cam, self.view_dir)
        np.testing.assert_allclose(s["position"], self.hit, atol=1e-5)
        rec = oracle.surface(perceptual_roughness=s["perceptual_roughness"], metalness=s["metalness"],
                             alpha_roughness=s["alpha_roughness"], r0=s["r0"], r90=(s["r90"],) * 3,
                             diffuse=s["diffuse"], specular=s["specular"], normal=s["normal"], basis=s["basis"])
        return rec, s

    def trace(self, frames, max_depth, light_mode=2, size=32):
        out = []
        for f in frames:
            r = self.sc.trace(size, size, self.inv_view, self.inv_proj, frame=f, max_depth=max_depth,
                              light_mode=light_mode, precision=64, threads=8)
            out.append(r["radiance"].reshape(-1, 3))
        return np.concatenate(out)


@pytest.mark.parametrize("light_mode", [0, 1, 2])
def test_nee_with_mis_closed_form(oracle_lib, light_mode):
    """No bounce (max depth 0): the pixel is the light's term times the BRDF times the
    power-heuristic weight of the light's selection probability (1, one light) against
    the BRDF's pdf: L cos B / (1 + pdf_B^2); zero where an occluder blocks the light."""
    scene = FloorUnderSun()
    rec, _ = scene.surface()
    v = -scene.view_dir
    b, pdf_b, _ = oracle.brdf(rec, v, scene.light_dir)
    want = scene.light_col * scene.light_dir[2] * b[0] / (1.0 + pdf_b[0] ** 2)
    got = scene.trace([1], 0, light_mode, size=2)
    # the light, camera and vertices reach the oracle as fp32 values
    np.testing.assert_allclose(got, np.tile(want, (4, 1)), rtol=1e-6)
    shadowed = FloorUnderSun(occluder=True).trace([1], 0, light_mode, size=2)
    np.testing.assert_array_equal(shadowed, 0.0)


def test_first_bounce_expectation(oracle_lib):
    """One bounce to the sky (max depth 1): a path's radiance is NEE + sky(l) B(l) cos(l) /
    pdf(l), clamped per channel to [0, 10]; its mean over many paths equals the
    expectation of that under sampleBRDF's mixture - the specular lobe's half vectors
    distributed as D cos (integrated over h, with l the mirror of v) weighted by the
    specular weight, the diffuse lobe's cosine-weighted directions by the rest; draws
    whose pdf falls under 1e-6 end the path after NEE.  (Mirrored directions below the
    horizon get a negative cosine and the BRDF pdf can be tiny there, so the clamp
    matters for a few paths in 10^4.)"""
    scene = FloorUnderSun()
    rec, s = scene.surface()
    v = -scene.view_dir
    b, pdf_b, sw = oracle.brdf(rec, v, scene.light_dir)
    nee = scene.light_col * scene.light_dir[2] * b[0] / (1.0 + pdf_b[0] ** 2)
    sw = sw[0]
    basis = s["basis"]  # rows t, b, n
    alpha = s["alpha_roughness"]

    def grid(nt, nphi, power):
        st = (np.arange(nt) + 0.5) / nt
        theta = st ** power * (np.pi / 2)
        dtheta = power * st ** (power - 1) * (np.pi / 2) / nt
        phi = (np.arange(nphi) + 0.5) / nphi * 2 * np.pi
        T, P = np.meshgrid(theta, phi, indexing="ij")
        local = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
        dw = (np.sin(T) * dtheta[:, None] * (2 * np.pi / nphi)).reshape(-1)
        return local @ basis, np.cos(T).reshape(-1), dw

    def path_radiance(l, h):
        f, pdf, _ = oracle.brdf(rec, np.tile(v, (len(l), 1)), l, h)
        cos_l = l @ basis[2]
        keep = pdf >= 1e-6
        val = np.tile(nee, (len(l), 1))
        val[keep] += sky_many(l[keep]) * f[keep] * cos_l[keep, None] / pdf[keep, None]
        return np.clip(val, 0.0, 10.0)

    h, cos_h, dw = grid(1200, 256, 2)
    w_spec = ggx_d(cos_h, alpha) * cos_h * dw
    assert abs(w_spec.sum() - 1) < 1e-4
    l = 2 * (h @ v)[:, None] * h - v
    spec = path_radiance(l, h).T @ w_spec
    l, cos_l, dw = grid(600, 256, 1)
    w_diff = cos_l / np.pi * dw
    assert abs(w_diff.sum() - 1) < 1e-4
    diff = path_radiance(l, unit(v + l)).T @ w_diff
    want = sw * spec + (1 - sw) * diff
    got = scene.trace(range(1, 65), 1)  # 64 frames x 1024 pixels
    mean, se = got.mean(0), got.std(0) / np.sqrt(len(got))
    print("first bounce: mean", mean, "+-", se, "expected", want, "(NEE", nee, "sw", sw, ")")
    assert np.all(np.abs(mean - want) < 5 * se + 2e-4 * want)
