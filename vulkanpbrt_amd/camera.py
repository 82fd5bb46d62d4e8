"""Camera matrices with the reference's conventions.

vsg builds ``perspective`` (Vulkan 0..1 depth, -f Y flip) and ``lookAt`` in
double precision and inverts them in double before they are cast to float
push constants (external/vsg/include/vsg/maths/transform.h:148-183,
source/VulkanPBRT.cpp:314-328).  Matrices here are numpy float64 arrays
indexed [row, col]; ``colmajor`` flattens them the way GLSL/vsg store them.
"""
from __future__ import annotations

import math

import numpy as np

from . import capi


def perspective(fovy_degrees: float, aspect: float, near: float, far: float) -> np.ndarray:
    f = 1.0 / math.tan(math.radians(fovy_degrees) * 0.5)
    r = 1.0 / (near - far)
    m = np.zeros((4, 4))
    m[0, 0] = f / aspect
    m[1, 1] = -f
    m[2, 2] = far * r
    m[3, 2] = -1.0
    m[2, 3] = far * near * r
    return m


def _normalize(v):
    return v / math.sqrt(float(np.dot(v, v)))


def look_at(eye, center, up) -> np.ndarray:
    eye = np.asarray(eye, dtype=np.float64)
    center = np.asarray(center, dtype=np.float64)
    up = np.asarray(up, dtype=np.float64)
    forward = _normalize(center - eye)
    up_n = _normalize(up)
    side = _normalize(np.cross(forward, up_n))
    u = _normalize(np.cross(side, forward))
    rot = np.eye(4)
    rot[0, :3] = side
    rot[1, :3] = u
    rot[2, :3] = -forward
    tr = np.eye(4)
    tr[:3, 3] = -eye
    return rot @ tr


def colmajor(m: np.ndarray) -> np.ndarray:
    return np.asarray(m, dtype=np.float64).flatten(order="F").astype(np.float32)


def pose_matrices(pose, width: int, height: int):
    """(view, proj) in double for a CameraPose-like object."""
    view = look_at(list(pose.eye), list(pose.center), list(pose.up))
    proj = perspective(float(pose.fovy_degrees), width / height, float(pose.near_distance),
                       float(pose.far_distance))
    return view, proj


def push_constants(view: np.ndarray, proj: np.ndarray, prev_view: np.ndarray, frame_number: int,
                   sample_number: int = 0) -> capi.PushConstants:
    pc = capi.PushConstants()
    pc.view_inverse[:] = colmajor(np.linalg.inv(view)).tolist()
    pc.proj_inverse[:] = colmajor(np.linalg.inv(proj)).tolist()
    pc.prev_view[:] = colmajor(prev_view).tolist()
    pc.frame_number = frame_number
    pc.sample_number = sample_number
    return pc
