"""Flattened world-space scene and camera records passed across the C ABI.

The reference renders an immutable ``SceneSnapshot`` (include/minirender/scene/scene.hpp:220-230)
whose instances are flattened by ``SceneBvh`` into world-space FP32 triangles in scene order
(instance-major, triangle-minor; src/render/bvh.cpp:53-89). This module holds exactly that
flattened form: every triangle's three world positions and per-corner world normals, its instance
and in-instance triangle index, one material per instance, the constant environment and the
directional lights. The camera is per render job (render_cli.cpp builds it from the scene file;
here a job carries its own camera so one resident scene serves many views).
"""
from __future__ import annotations

import dataclasses
import hashlib

import numpy as np

# MaterialKind (scene.hpp:105). Only the four kinds on the offline hot path are supported.
DIFFUSE, CONDUCTOR, MIRROR, DIELECTRIC = 0, 1, 2, 3
KIND_NAMES = {"diffuse": DIFFUSE, "conductor": CONDUCTOR, "mirror": MIRROR, "dielectric": DIELECTRIC}

# TextureFilter (texture.hpp:13).
FILTERS = {"nearest": 0, "bilinear": 1, "trilinear": 2}

# Layout shared with `mrg_material` (include/mrg.h): 68 bytes, 4-byte aligned.
MATERIAL_DTYPE = np.dtype(
    [
        ("kind", "<u4"),
        ("base_color", "<f4", (3,)),
        ("emission", "<f4", (3,)),
        ("roughness", "<f4"),
        ("ior", "<f4"),
        ("double_sided", "<u4"),
        ("mirrored", "<u4"),
        ("compensate_energy", "<u4"),
        ("base_color_texture", "<u4"),  # 1 + texture index, 0 = none
        ("texture_filter", "<u4"),
        ("normal_texture", "<u4"),      # 1 + texture index, 0 = none
        ("normal_scale", "<f4"),
        ("normal_flip_y", "<u4"),
    ]
)

# `mrg_directional_light`: direction is the propagation direction (scene.hpp:180-186).
DIRECTIONAL_DTYPE = np.dtype([("direction", "<f4", (3,)), ("color", "<f4", (3,)), ("intensity", "<f4")])

# `mrg_camera` (Camera, scene.hpp:85-103).
CAMERA_DTYPE = np.dtype(
    [
        ("position", "<f4", (3,)),
        ("target", "<f4", (3,)),
        ("up", "<f4", (3,)),
        ("vertical_fov_degrees", "<f4"),
        ("near_plane", "<f4"),
        ("far_plane", "<f4"),
        ("aperture_radius", "<f4"),
        ("focus_distance", "<f4"),
    ]
)


def material(kind="diffuse", base_color=(0.7, 0.7, 0.7), emission=(0, 0, 0), roughness=0.35, ior=1.5,
             double_sided=True, mirrored=False, compensate_energy=False, base_color_texture=-1,
             texture_filter="trilinear", normal_texture=-1, normal_scale=1.0, normal_flip_y=False):
    """One material record with the reference defaults (Material, scene.hpp:108-129).
    base_color_texture and normal_texture index Scene.textures (-1 = none)."""
    m = np.zeros((), MATERIAL_DTYPE)
    m["kind"] = KIND_NAMES[kind] if isinstance(kind, str) else int(kind)
    m["base_color"] = base_color
    m["emission"] = emission
    m["roughness"] = roughness
    m["ior"] = ior
    m["double_sided"] = 1 if double_sided else 0
    m["mirrored"] = 1 if mirrored else 0
    m["compensate_energy"] = 1 if compensate_energy else 0
    m["base_color_texture"] = int(base_color_texture) + 1
    m["texture_filter"] = FILTERS[texture_filter] if isinstance(texture_filter, str) else int(texture_filter)
    m["normal_texture"] = int(normal_texture) + 1
    m["normal_scale"] = normal_scale
    m["normal_flip_y"] = 1 if normal_flip_y else 0
    return m


@dataclasses.dataclass
class Camera:
    """Right-handed camera; uv origin at the top-left (scene.hpp:82-103)."""

    position: tuple = (0.0, 1.5, 4.0)
    target: tuple = (0.0, 0.5, 0.0)
    up: tuple = (0.0, 1.0, 0.0)
    vertical_fov_degrees: float = 45.0
    near_plane: float = 0.05
    far_plane: float = 100.0
    aperture_radius: float = 0.0
    focus_distance: float = 4.0

    def record(self) -> np.ndarray:
        c = np.zeros((), CAMERA_DTYPE)
        for f in CAMERA_DTYPE.names:
            c[f] = getattr(self, f)
        return c

    def validate(self):
        """Camera::validate (scene.cpp:141-158)."""
        p, t, u = (np.asarray(v, np.float32) for v in (self.position, self.target, self.up))
        if not (np.all(np.isfinite(p)) and np.all(np.isfinite(t)) and np.all(np.isfinite(u))):
            raise ValueError("Camera vectors must be finite")
        if np.linalg.norm(t - p) <= 1e-6:
            raise ValueError("Camera position and target must differ")
        if np.linalg.norm(np.cross(t - p, u)) <= 1e-6:
            raise ValueError("Camera up must not be parallel to view direction")
        if not (0 < self.vertical_fov_degrees < 179):
            raise ValueError("Camera FOV must be between 0 and 179 degrees")
        if not (0 < self.near_plane < self.far_plane):
            raise ValueError("Camera clipping planes must satisfy 0 < near < far")
        if not (self.aperture_radius >= 0 and self.focus_distance > 0):
            raise ValueError("Camera lens requires nonnegative aperture and positive focus distance")


@dataclasses.dataclass
class Scene:
    """Flattened snapshot: world-space FP32 triangles in scene order plus per-instance materials."""

    positions: np.ndarray            # (n, 3, 3) float32
    normals: np.ndarray              # (n, 3, 3) float32 per-corner world normals
    instance: np.ndarray             # (n,) uint32
    triangle: np.ndarray             # (n,) uint32
    materials: np.ndarray            # (instances,) MATERIAL_DTYPE
    environment: tuple = (0.03, 0.03, 0.03)
    directional: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, DIRECTIONAL_DTYPE))
    camera: Camera = dataclasses.field(default_factory=Camera)
    name: str = "scene"
    uvs: np.ndarray = None           # (n, 3, 2) float32 per-corner texture coordinates, or None
    textures: list = dataclasses.field(default_factory=list)  # linear RGB float32 (h, w, 3), row 0 = top
    environment_texture: int = -1    # EnvironmentMap (scene.hpp:188-193): index into textures
    environment_strength: float = 1.0
    environment_rotation_degrees: float = 0.0

    def __post_init__(self):
        self.positions = np.ascontiguousarray(self.positions, np.float32).reshape(-1, 3, 3)
        self.normals = np.ascontiguousarray(self.normals, np.float32).reshape(-1, 3, 3)
        self.instance = np.ascontiguousarray(self.instance, np.uint32).reshape(-1)
        self.triangle = np.ascontiguousarray(self.triangle, np.uint32).reshape(-1)
        self.materials = np.ascontiguousarray(np.atleast_1d(self.materials), MATERIAL_DTYPE)
        self.directional = np.ascontiguousarray(np.atleast_1d(self.directional), DIRECTIONAL_DTYPE)
        self.environment = tuple(float(np.float32(v)) for v in self.environment)
        n = self.positions.shape[0]
        if self.normals.shape[0] != n or self.instance.shape[0] != n or self.triangle.shape[0] != n:
            raise ValueError("per-triangle arrays must have the same length")
        if self.uvs is not None:
            self.uvs = np.ascontiguousarray(self.uvs, np.float32).reshape(-1, 3, 2)
            if self.uvs.shape[0] != n:
                raise ValueError("per-triangle arrays must have the same length")
        self.textures = [np.ascontiguousarray(t, np.float32).reshape(np.shape(t)[0], np.shape(t)[1], 3) for t in self.textures]
        self.environment_strength = float(np.float32(self.environment_strength))
        self.environment_rotation_degrees = float(np.float32(self.environment_rotation_degrees))
        if n and int(self.instance.max()) >= len(self.materials):
            raise ValueError("instance index out of range")

    @property
    def triangle_count(self) -> int:
        return int(self.positions.shape[0])

    def content_hash(self) -> str:
        """Hash of everything the device copy depends on (scene dedup key)."""
        h = hashlib.blake2b(digest_size=16)
        for a in (self.positions, self.normals, self.instance, self.triangle, self.materials, self.directional):
            h.update(np.ascontiguousarray(a).tobytes())
        h.update(np.asarray(self.environment, np.float32).tobytes())
        if self.uvs is not None:
            h.update(self.uvs.tobytes())
        for t in self.textures:
            h.update(np.asarray(t.shape, np.uint32).tobytes())
            h.update(t.tobytes())
        h.update(np.asarray([self.environment_texture], np.int32).tobytes())
        h.update(np.asarray([self.environment_strength, self.environment_rotation_degrees], np.float32).tobytes())
        return h.hexdigest()

    def device_bytes(self) -> int:
        """Approximate HBM footprint of the uploaded scene (BVH + triangles + shading data)."""
        n = self.triangle_count
        return n * (48 + 36 + 36 + 8 + 4) + 2 * n * 64


def concat(parts, materials, **kw) -> Scene:
    """Build a scene from (positions, normals) per instance, in instance order."""
    pos, nrm, inst, tri = [], [], [], []
    for i, (p, n) in enumerate(parts):
        p = np.asarray(p, np.float32).reshape(-1, 3, 3)
        n = np.asarray(n, np.float32).reshape(-1, 3, 3)
        pos.append(p)
        nrm.append(n)
        inst.append(np.full(len(p), i, np.uint32))
        tri.append(np.arange(len(p), dtype=np.uint32))
    if not pos:
        pos, nrm = [np.zeros((0, 3, 3), np.float32)], [np.zeros((0, 3, 3), np.float32)]
        inst, tri = [np.zeros(0, np.uint32)], [np.zeros(0, np.uint32)]
    mats = np.array([np.asarray(m, MATERIAL_DTYPE) for m in materials], MATERIAL_DTYPE) if materials else np.zeros(0, MATERIAL_DTYPE)
    return Scene(np.concatenate(pos), np.concatenate(nrm), np.concatenate(inst), np.concatenate(tri), mats, **kw)


def directional_lights(*lights) -> np.ndarray:
    out = np.zeros(len(lights), DIRECTIONAL_DTYPE)
    for i, (d, c, s) in enumerate(lights):
        out[i]["direction"] = d
        out[i]["color"] = c
        out[i]["intensity"] = s
    return out
