"""Pins of the oracle's scene packing (RayTracingVisitor.cpp:52-118, 170-221, 414-441)
against hand computations, on the CPU.

* An emissive quad under a known instance transform becomes two area lights whose
  vertices are the transform applied in double precision to the mesh vertices,
  whose three colours are the emission, attenuation (0, 0, 1), and whose inclusive
  strengths are the running sum of the nine colour components (explicit lights
  continue the sum after them).
* A scene with no emitter and no explicit light gets the default sun.
* A mesh without normals gets area-weighted vertex normals: the shading normal at a
  hit with known barycentrics is the normalised interpolation of the hand-computed
  vertex normals, carried by the instance's inverse-transpose matrix.
"""
import math

import numpy as np
import pytest

from oracle import oracle
from tests.test_oracle_shading import Desc
from vulkanpbrt_amd import capi

QUAD = dict(positions=[[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.0, 2.0, 0.0], [0.0, 2.0, 0.0]],
            indices=[0, 1, 2, 0, 2, 3])

# PackedLight (Light.h:23-37) as 32 floats
V0, TYPE, V1, RADIUS, V2, ANGLE, DIR, ANGLE2 = slice(0, 3), 3, slice(4, 7), 7, slice(8, 11), 11, slice(12, 15), 15
AMB, DIFF, SPEC, STRENGTHS, INCLUSIVE = slice(16, 20), slice(20, 24), slice(24, 28), slice(28, 31), 31
LST_DIRECTIONAL, LST_POINT, LST_AREA = 1, 2, 5


def transform(scale, angle_z, translate):
    c, s = math.cos(angle_z), math.sin(angle_z)
    m = np.eye(4)
    m[:3, :3] = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]]) @ np.diag(scale)
    m[:3, 3] = translate
    return m.astype(np.float32)


def test_emissive_quad_lights_under_a_transform(oracle_lib):
    emission = (0.5, 1.0, 2.0)
    m = transform((2.0, 0.5, 3.0), 0.7, (1.0, -2.0, 3.0))
    lamp = dict(QUAD, emission=emission, transform=m)
    floor = dict(QUAD, transform=transform((5.0, 5.0, 1.0), 0.0, (0.0, 0.0, -1.0)))  # not an emitter
    point = capi.Light()
    point.v0_type[:] = [4.0, 5.0, 6.0, float(LST_POINT)]
    point.col_ambient[:] = [0.1, 0.2, 0.3, 0.0]
    point.col_diffuse[:] = [1.0, 1.0, 1.0, 0.0]
    point.col_specular[:] = [0.25, 0.0, 0.0, 0.0]
    point.strengths[:] = [1.0, 0.0, 0.5, 0.0]
    d = Desc([floor, lamp], lights=[point])
    osc = oracle_lib.OracleScene(d.desc)
    li = osc.lights()
    assert li.shape[0] == 3  # two triangle lights, then the explicit one
    m64 = m.astype(np.float64)
    pos = np.asarray(QUAD["positions"], np.float64)
    world = (np.c_[pos, np.ones(4)] @ m64.T)[:, :3].astype(np.float32)
    for k, tri in enumerate(([0, 1, 2], [0, 2, 3])):
        assert np.array_equal(li[k, V0], world[tri[0]])
        assert np.array_equal(li[k, V1], world[tri[1]])
        assert np.array_equal(li[k, V2], world[tri[2]])
        assert li[k, TYPE] == LST_AREA and li[k, RADIUS] == 0.0
        for sl in (AMB, DIFF, SPEC):
            assert np.array_equal(li[k, sl], np.array(emission + (0.0,), np.float32))
        assert np.array_equal(li[k, STRENGTHS], np.array([0.0, 0.0, 1.0], np.float32))
    per_light = np.float32(3 * sum(emission))
    assert li[0, INCLUSIVE] == per_light and li[1, INCLUSIVE] == np.float32(2 * per_light)
    point_strength = np.float32(0.1 + 0.2 + 0.3 + 3.0 + 0.25)
    assert np.isclose(li[2, INCLUSIVE], 2 * per_light + point_strength, rtol=1e-6)
    assert np.array_equal(li[2, V0], np.array([4.0, 5.0, 6.0], np.float32)) and li[2, TYPE] == LST_POINT
    assert osc.strength_sum == li[2, INCLUSIVE]


def test_no_lights_gives_the_default_sun(oracle_lib):
    d = Desc([QUAD])
    osc = oracle_lib.OracleScene(d.desc)
    li = osc.lights()
    assert li.shape[0] == 1
    sun = li[0]
    assert sun[TYPE] == LST_DIRECTIONAL
    want = np.array([0.1, 1.0, -5.1]) / np.linalg.norm([0.1, 1.0, -5.1])
    assert np.allclose(sun[DIR], want, rtol=0, atol=1e-7)
    assert np.array_equal(sun[AMB], np.float32([2, 2, 2, 0])) and np.array_equal(sun[DIFF], np.float32([2, 2, 2, 0]))
    assert np.array_equal(sun[SPEC], np.zeros(4, np.float32))
    assert np.array_equal(sun[STRENGTHS], np.float32([0.5, 0.0, 0.0]))
    assert sun[INCLUSIVE] == 12.0 and osc.strength_sum == 12.0


ROOF = dict(positions=[[0.0, 0.0, 0.0], [2.0, 0.0, 0.0], [0.0, 1.0, 0.0], [-1.0, 0.0, 1.0]],
            indices=[0, 1, 2, 0, 2, 3], normals=None)


def roof_vertex_normals():
    """Area-weighted average of unit face normals (RayTracingVisitor.cpp:52-118).  Face
    (A, B, C): cross = (0, 0, 2), weight 2; face (A, C, D): cross = (1, 0, 1), weight
    sqrt(2); A and C are shared."""
    n1, w1 = np.array([0.0, 0.0, 1.0]), 2.0
    n2, w2 = np.array([1.0, 0.0, 1.0]) / math.sqrt(2.0), math.sqrt(2.0)
    shared = (w1 * n1 + w2 * n2) / (w1 + w2)
    return np.array([shared, n1, shared, n2])


@pytest.mark.parametrize("m", [np.eye(4, dtype=np.float32), transform((1.0, 2.0, 0.5), 0.4, (3.0, 1.0, -2.0))])
def test_area_weighted_normals(oracle_lib, m):
    d = Desc([dict(ROOF, transform=m)])
    osc = oracle_lib.OracleScene(d.desc)
    bary = np.array([0.2, 0.3, 0.5])  # on face (A, B, C)
    pos = np.asarray(ROOF["positions"], np.float64)
    p_obj = bary @ pos[[0, 1, 2]]
    m64 = m.astype(np.float64)
    p_world = m64[:3, :3] @ p_obj + m64[:3, 3]
    d_world = m64[:3, :3] @ np.array([0.0, 0.0, -1.0])
    d_world /= np.linalg.norm(d_world)
    s = osc.shade(p_world - 3.0 * d_world, d_world)
    assert s is not None and s["tri"] >= 0
    assert np.allclose([s["u"], s["v"]], bary[1:], atol=1e-6)
    vn = roof_vertex_normals()
    n_obj = bary @ vn[[0, 1, 2]]
    n_world = np.linalg.inv(m64[:3, :3]).T @ (n_obj / np.linalg.norm(n_obj))
    n_world /= np.linalg.norm(n_world)
    assert np.allclose(s["normal"], n_world, atol=1e-6)
    # the geometric normal of face (A, B, C) differs: the shared vertices lean towards +x
    assert abs(s["normal"] @ (np.linalg.inv(m64[:3, :3]).T @ [0, 0, 1.0])) < 0.999 * np.linalg.norm(
        np.linalg.inv(m64[:3, :3]).T @ [0, 0, 1.0])
