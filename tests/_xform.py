"""Instance transforms for test scenes: the reference's Transform::matrix = T * Rz * Ry * Rx * S and its
normal matrix (src/scene/scene.cpp:126-139), evaluated in FP64, and the SceneBvh flattening of an
instance into world-space FP32 triangles (src/render/bvh.cpp:53-89): world positions are
float(M * (p, 1)), corner normals are the normal matrix applied to the object-space normal (the
oracle and the kernels normalize them), and an instance is mirrored when det(M) < 0."""
from __future__ import annotations

import math

import numpy as np


def _axis_rotation(axis, degrees):
    a = math.radians(degrees)
    c, s = math.cos(a), math.sin(a)
    x, y, z = axis
    return np.array([[c + x * x * (1 - c), x * y * (1 - c) - z * s, x * z * (1 - c) + y * s],
                     [y * x * (1 - c) + z * s, c + y * y * (1 - c), y * z * (1 - c) - x * s],
                     [z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, c + z * z * (1 - c)]])


def matrix(translation=(0, 0, 0), rotation_degrees=(0, 0, 0), scale=(1, 1, 1)):
    """4x4 object-to-world matrix T * Rz * Ry * Rx * S (column vectors)."""
    r = _axis_rotation((0, 0, 1), rotation_degrees[2]) @ _axis_rotation((0, 1, 0), rotation_degrees[1]) @ \
        _axis_rotation((1, 0, 0), rotation_degrees[0])
    m = np.eye(4)
    m[:3, :3] = r @ np.diag(np.asarray(scale, np.float64))
    m[:3, 3] = translation
    return m


def normal_matrix(m):
    return np.linalg.inv(m[:3, :3]).T


def point(m, p):
    return (m[:3, :3] @ np.asarray(p, np.float64)) + m[:3, 3]


def flatten(m, positions, normals):
    """Object-space triangles (n, 3, 3) and corner normals (n, 3, 3) -> world (positions, normals, mirrored)."""
    p = np.asarray(positions, np.float64).reshape(-1, 3)
    n = np.asarray(normals, np.float64).reshape(-1, 3)
    wp = (p @ m[:3, :3].T + m[:3, 3]).astype(np.float32).reshape(-1, 3, 3)
    wn = (n @ normal_matrix(m).T).astype(np.float32).reshape(-1, 3, 3)
    return wp, wn, bool(np.linalg.det(m[:3, :3]) < 0)
