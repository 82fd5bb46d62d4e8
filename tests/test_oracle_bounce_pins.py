"""Pins of the oracle's bounce loop and Russian roulette (ptRaygen.rgen:62-81,
lighting.glsl:263-287) against a closed form, on the CPU.

The scene is a closed box of six inward-facing walls, all a smooth metal (diffuse 0,
specular 1, perceptual roughness 0.01) whose emission is (e, -e, 0).  Two facts of
the reference make this box a closed-form case:

* a mesh is an emitter for light sampling only when its emission components sum to
  non-zero (RayTracingVisitor.cpp:289), so the box emits on every hit but holds no
  light; with no light the scene gets the default sun, whose shadow rays all end on
  a wall: next-event estimation adds exactly 0;
* with F = 1 (reflectance0 = reflectance90 = 1) and the specular lobe always chosen
  (the diffuse luminance is 0), one bounce weights the path by
  brdf * cos / pdf = F G D / (4 NL NV) * NL / (D NH) = G / (4 NV NH); a camera with a
  0.5 degree field of view looks along x at normal incidence, the GGX half vectors of
  a roughness-0.01 lobe stay within ~1e-4 of the normal and G = 1 - O(1e-8), so every
  bounce weights the path by 1/4 to within ~1e-6.

The red radiance of every pixel is then e * sum_{k=0..M} 4^-k (emission at the
primary hit and after each of the M bounces), the green channel (-e ...) clamps to 0,
and the path never ends early: with Russian roulette after bounce m the expectation is
unchanged (throughput / 0.95) and the bounce-i ray is traced with probability
0.95^max(0, i - m).  A dropped bounce, a missing or doubled throughput factor, the
roulette's 0.05 or its compensation, or its start depth each move one of these.
"""
import numpy as np
import pytest

from tests.test_oracle_shading import Desc
from vulkanpbrt_amd import camera

E = 3.0
W = H = 32


def box_walls():
    """Six quads of [-1, 1]^3, each with its own vertices and winding facing inwards."""
    meshes = []
    for axis in range(3):
        for sign in (-1.0, 1.0):
            u, v = (axis + 1) % 3, (axis + 2) % 3
            quad = np.zeros((4, 3))
            quad[:, axis] = sign
            for k, (a, b) in enumerate([(-1, -1), (1, -1), (1, 1), (-1, 1)]):
                quad[k, u], quad[k, v] = a, b
            # (p1 - p0) x (p2 - p0) points along +axis for this corner order; inwards is -sign
            idx = [0, 1, 2, 0, 2, 3] if sign < 0 else [0, 2, 1, 0, 3, 2]
            meshes.append(dict(positions=quad, indices=idx, diffuse=(0.0, 0.0, 0.0), specular=(1.0, 1.0, 1.0),
                               roughness=0.01, emission=(E, -E, 0.0)))
    return meshes


@pytest.fixture(scope="module")
def mirror_box(oracle_lib):
    d = Desc(box_walls())
    osc = oracle_lib.OracleScene(d.desc)
    pose = dict(eye=(0.0, 0.0, 0.0), center=(-1.0, 0.0, 0.0), up=(0.0, 0.0, 1.0))
    view = camera.look_at(pose["eye"], pose["center"], pose["up"])
    proj = camera.perspective(0.5, 1.0, 0.01, 100.0)
    iv = camera.colmajor(np.linalg.inv(view))
    ip = camera.colmajor(np.linalg.inv(proj))
    return d, osc, iv, ip


def series(m):
    return E * sum(0.25 ** k for k in range(m + 1))


def test_box_holds_no_light_but_the_sun(mirror_box):
    _, osc, _, _ = mirror_box
    li = osc.lights()
    assert li.shape[0] == 1 and li[0, 3] == 1.0  # the default directional light
    assert osc.triangle_count == 12


@pytest.mark.parametrize("max_depth", [1, 2, 3, 5])
def test_radiance_is_the_truncated_series(mirror_box, max_depth):
    """No Russian roulette (min depth = max depth): every path adds e 4^-k at bounce k."""
    _, osc, iv, ip = mirror_box
    r = osc.trace(W, H, iv, ip, frame=3, max_depth=max_depth, min_depth=max_depth, precision=64)
    red = r["radiance"][..., 0]
    assert np.allclose(red, series(max_depth), rtol=2e-5, atol=0)
    assert np.all(r["radiance"][..., 1] == 0.0) and np.all(r["radiance"][..., 2] == 0.0)
    assert r["rays"] == W * H * (1 + max_depth)  # primary + one extension ray per bounce


def test_russian_roulette_keeps_the_expectation(mirror_box):
    """minRecursionDepth m = 1 of 6 bounces: the mean over many paths matches the series
    to its statistical error, and the traced rays per bounce follow 0.95^(i - m)."""
    _, osc, iv, ip = mirror_box
    max_depth, min_depth, frames = 6, 1, 24
    red, rays = [], 0
    for f in range(frames):
        r = osc.trace(W, H, iv, ip, frame=f + 1, max_depth=max_depth, min_depth=min_depth, precision=64)
        red.append(r["radiance"][..., 0].ravel())
        rays += r["rays"]
    red = np.concatenate(red)
    n = red.size
    want = series(max_depth)
    assert abs(red.mean() - want) < 4.0 * red.std() / np.sqrt(n) + 1e-5 * want
    # the estimate is not the series on each path: roulette made some paths end early
    assert red.std() > 1e-3
    survive = [0.95 ** max(0, i - min_depth) for i in range(max_depth)]
    expect_rays = n * (1 + sum(survive))
    # variance of the number of extension rays per path: sum of a stopped chain of
    # Bernoulli(0.95) survivals; bounded by the variance of a geometric count
    sigma = np.sqrt(n * sum(survive) * 0.05 / 0.95 ** 2) * 3
    assert abs(rays - expect_rays) < 4.0 * sigma


def test_roulette_starts_after_min_depth(mirror_box):
    """Bounce i (0-based) plays roulette only when i > m (lighting.glsl:275): with
    m = max depth - 2 only the last bounce ray can be dropped, and each path either adds
    the full series (its last term over 0.95) or stops one short; with m = max depth - 1
    no path is dropped."""
    _, osc, iv, ip = mirror_box
    max_depth = 4
    r = osc.trace(W, H, iv, ip, frame=9, max_depth=max_depth, min_depth=max_depth - 1, precision=64)
    assert np.allclose(r["radiance"][..., 0], series(max_depth), rtol=2e-5)
    r = osc.trace(W, H, iv, ip, frame=9, max_depth=max_depth, min_depth=max_depth - 2, precision=64)
    red = r["radiance"][..., 0]
    full = E * sum(0.25 ** k for k in range(max_depth)) + E * 0.25 ** max_depth / 0.95
    short = series(max_depth - 1)
    ok = np.isclose(red, full, rtol=2e-5) | np.isclose(red, short, rtol=2e-5)
    assert ok.all()
    frac_short = np.isclose(red, short, rtol=2e-5).mean()
    assert abs(frac_short - 0.05) < 4 * np.sqrt(0.05 * 0.95 / red.size)
