"""Seeded scene and ray generators for the benchmark configurations and the tests.

Every scene is generated directly in world space (identity instance transforms), so no matrix
arithmetic touches vertex bits: the flattening in src/render/bvh.cpp:53-89 then reduces to a copy.

* ``cornell_box``       Cornell box: 5 walls + 2 rotated boxes + 1 quad lamp (36 triangles).
* ``icosphere``         icosahedron subdivided ``level`` times, radially displaced by value noise.
* ``intersection_rays`` the seeded ray set of the intersection microbench.
* ``mixed_terrain``     fBm heightfield terrain + 64 spheres of 4 materials, sun + sky, thin lens.
* ``job_stream``        the uneven multi-job distribution for batching and serving.
"""
from __future__ import annotations

import math

import numpy as np

from . import rng
from .scene import Camera, Scene, concat, directional_lights, material

# ----------------------------------------------------------------------------------------------
# Small helpers.


def _quad(a, b, c, d):
    """Two triangles (a,b,c), (a,c,d); winding gives normal cross(b-a, c-a)."""
    return np.array([[a, b, c], [a, c, d]], np.float64)


def _flat_normals(tris):
    n = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    return np.repeat(n[:, None, :], 3, axis=1)


def _box(center, size, yaw_degrees):
    """Closed box, outward winding, rotated about +Y (12 triangles)."""
    hx, hy, hz = (s / 2 for s in size)
    c = [(-hx, -hy, -hz), (hx, -hy, -hz), (hx, hy, -hz), (-hx, hy, -hz),
         (-hx, -hy, hz), (hx, -hy, hz), (hx, hy, hz), (-hx, hy, hz)]
    faces = [(4, 5, 6, 7), (1, 0, 3, 2), (5, 1, 2, 6), (0, 4, 7, 3), (7, 6, 2, 3), (0, 1, 5, 4)]
    tris = np.concatenate([_quad(*[c[i] for i in f]) for f in faces])
    a = math.radians(yaw_degrees)
    rot = np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])
    return tris @ rot.T + np.asarray(center)


# ----------------------------------------------------------------------------------------------
# Cornell box.


def cornell_box(lamp_radiance=(17.0, 12.0, 4.0)) -> Scene:
    """Cornell box in world space; single-sided inward walls, as scenes/cornell-box.json."""
    walls = [
        ("floor", _quad((-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1)), (0.73, 0.73, 0.73)),
        ("ceiling", _quad((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1)), (0.73, 0.73, 0.73)),
        ("back", _quad((-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1)), (0.73, 0.73, 0.73)),
        ("left", _quad((-1, 0, -1), (-1, 2, -1), (-1, 2, 1), (-1, 0, 1)), (0.65, 0.05, 0.05)),
        ("right", _quad((1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1)), (0.12, 0.45, 0.15)),
    ]
    parts, mats = [], []
    for _, tris, color in walls:
        parts.append((tris, _flat_normals(tris)))
        mats.append(material("diffuse", color, double_sided=False))
    for center, size, yaw in (((-0.36, 0.3, 0.32), (0.6, 0.6, 0.6), -17.0),
                              ((0.34, 0.6, -0.3), (0.6, 1.2, 0.6), 17.0)):
        tris = _box(center, size, yaw)
        parts.append((tris, _flat_normals(tris)))
        mats.append(material("diffuse", (0.73, 0.73, 0.73), double_sided=False))
    lamp = _quad((-0.25, 1.985, -0.35), (0.25, 1.985, -0.35), (0.25, 1.985, 0.05), (-0.25, 1.985, 0.05))
    parts.append((lamp, _flat_normals(lamp)))
    mats.append(material("diffuse", (0, 0, 0), emission=lamp_radiance, double_sided=False))
    cam = Camera(position=(0, 1, 3.4), target=(0, 1, 0), up=(0, 1, 0), vertical_fov_degrees=40,
                 near_plane=0.03, far_plane=100)
    return concat(parts, mats, environment=(0, 0, 0), camera=cam, name="cornell")


# ----------------------------------------------------------------------------------------------
# Value noise and the displaced icosphere.


def _hash3(ix, iy, iz, seed):
    salt = np.uint64((int(seed) * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF)
    with np.errstate(over="ignore"):
        h = (ix.astype(np.uint64) * np.uint64(0x8DA6B343) ^ iy.astype(np.uint64) * np.uint64(0xD8163841)
             ^ iz.astype(np.uint64) * np.uint64(0xCB1AB31F) ^ salt)
    return (rng.mix64(h) >> np.uint64(11)).astype(np.float64) / float(1 << 53)


def value_noise3(p, seed, frequency=4.0):
    """Trilinear lattice value noise in [0,1) with smoothstep weights."""
    q = np.asarray(p, np.float64) * frequency
    i = np.floor(q)
    f = q - i
    w = f * f * (3 - 2 * f)
    i = i.astype(np.int64)
    out = np.zeros(q.shape[:-1])
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                v = _hash3(i[..., 0] + dx, i[..., 1] + dy, i[..., 2] + dz, seed)
                wx = w[..., 0] if dx else 1 - w[..., 0]
                wy = w[..., 1] if dy else 1 - w[..., 1]
                wz = w[..., 2] if dz else 1 - w[..., 2]
                out += v * wx * wy * wz
    return out


def _icosahedron():
    t = (1 + 5 ** 0.5) / 2
    v = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
                  (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    return v[np.array(f)]


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def subdivided_sphere(level):
    """Unit-sphere triangles (20*4^level, outward winding). Shared edges get identical midpoints."""
    tris = _icosahedron()
    for _ in range(level):
        a, b, c = tris[:, 0], tris[:, 1], tris[:, 2]
        ab, bc, ca = _unit(a + b), _unit(b + c), _unit(c + a)
        tris = np.stack([np.stack([a, ab, ca], 1), np.stack([ab, b, bc], 1), np.stack([ca, bc, c], 1),
                         np.stack([ab, bc, ca], 1)], 1).reshape(-1, 3, 3)
    return tris


def icosphere(level=8, amplitude=0.05, seed=2, radius=1.0) -> Scene:
    """Icosphere radially displaced by r*(1 + amplitude*(2*noise-1)); all triangles double-sided."""
    dirs = subdivided_sphere(level)
    noise = value_noise3(dirs, seed)
    r = radius * (1 + amplitude * (2 * noise - 1))
    pos = (dirs * r[..., None]).astype(np.float32)
    nrm = dirs.astype(np.float32)
    mats = [material("diffuse", (0.7, 0.7, 0.7), double_sided=True)]
    return concat([(pos, nrm)], mats, environment=(0.5, 0.5, 0.5), name=f"icosphere{level}")


def intersection_rays(scene: Scene, count, seed=2, start=0, scale=1.2, indices=None):
    """Seeded ray set: ray i uses Sampler::for_sample(seed, i, 0); origins uniform in the `scale`x
    AABB about its centre (3 dims), directions uniform on the sphere (2 dims). FP32 rays, as
    consumed by the GPU; the FP64 oracle is given the same float values widened."""
    p = scene.positions.reshape(-1, 3).astype(np.float64)
    lo, hi = p.min(0), p.max(0)
    c, h = (lo + hi) / 2, (hi - lo) / 2 * scale
    lo, hi = c - h, c + h
    idx = np.arange(start, start + count, dtype=np.uint64) if indices is None else np.asarray(indices, np.uint64)
    s = rng.Streams.for_sample(np.uint64(seed), idx, np.uint64(0))
    u = np.stack([s.next_1d() for _ in range(5)], axis=1)
    origin = lo + (hi - lo) * u[:, :3]
    d = rng.sample_sphere(u[:, 3], u[:, 4])
    rays = np.concatenate([origin, d], axis=1).astype(np.float32)
    return rays


# ----------------------------------------------------------------------------------------------
# Mixed-material terrain scene.


def _fbm_height(x, z, seed, octaves=6):
    h = np.zeros_like(x)
    amp, freq = 1.0, 1.0 / 24.0
    for o in range(octaves):
        p = np.stack([x * freq, z * freq, np.full_like(x, 0.37 * o)], -1)
        h += amp * (2 * value_noise3(p, seed + 101 * o, 1.0) - 1)
        amp *= 0.5
        freq *= 2.0
    return h


def terrain(cells=1024, extent=40.0, height=4.0, seed=3):
    """fBm heightfield over [-extent, extent]^2, (cells+1)^2 vertices, 2*cells^2 triangles; per-vertex
    normals from central differences of the height grid."""
    g = np.linspace(-extent, extent, cells + 1)
    x, z = np.meshgrid(g, g, indexing="xy")
    y = height * _fbm_height(x, z, seed)
    step = 2 * extent / cells
    dz, dx = np.gradient(y, step, step)
    n = _unit(np.stack([-dx, np.ones_like(dx), -dz], -1))
    v = np.stack([x, y, z], -1)
    a, b, c, d = v[:-1, :-1], v[1:, :-1], v[1:, 1:], v[:-1, 1:]
    na, nb, nc, nd = n[:-1, :-1], n[1:, :-1], n[1:, 1:], n[:-1, 1:]
    pos = np.stack([np.stack([a, b, c], -2), np.stack([a, c, d], -2)], 2).reshape(-1, 3, 3)
    nrm = np.stack([np.stack([na, nb, nc], -2), np.stack([na, nc, nd], -2)], 2).reshape(-1, 3, 3)
    return pos, nrm, (lambda px, pz: height * _fbm_height(np.asarray(px, float), np.asarray(pz, float), seed))


def mixed_terrain(cells=1024, spheres=64, sphere_level=3, seed=3, width=1920, height=1080) -> Scene:
    """Terrain + `spheres` spheres of diffuse / GGX conductor / mirror / glass, sun + sky, f/2.8."""
    r = np.random.default_rng(seed)
    tpos, tnrm, h = terrain(cells=cells, seed=seed)
    parts = [(tpos, tnrm)]
    mats = [material("diffuse", (0.42, 0.40, 0.33), double_sided=True)]
    unit = subdivided_sphere(sphere_level)
    kinds = ["diffuse", "conductor", "mirror", "dielectric"]
    for i in range(spheres):
        radius = r.uniform(0.6, 1.6)
        cx, cz = r.uniform(-14, 14), r.uniform(-16, 10)
        cy = float(h(cx, cz)) + radius
        parts.append((unit * radius + (cx, cy, cz), unit))
        kind = kinds[i % 4]
        if kind == "diffuse":
            mats.append(material(kind, r.uniform(0.15, 0.85, 3)))
        elif kind == "conductor":
            mats.append(material(kind, r.uniform(0.5, 0.95, 3), roughness=r.uniform(0.05, 0.6)))
        elif kind == "mirror":
            mats.append(material(kind, (0.9, 0.9, 0.9)))
        else:
            mats.append(material(kind, (1.0, 1.0, 1.0), ior=1.5))
    y0 = float(h(0.0, 0.0))
    fov = 40.0
    focal = 0.012 / math.tan(math.radians(fov) / 2)          # 24 mm film height
    position = np.array([0.0, y0 + 5.0, 26.0])
    target = np.array([0.0, y0 + 0.5, -2.0])
    cam = Camera(position=tuple(position), target=tuple(target), up=(0, 1, 0), vertical_fov_degrees=fov,
                 near_plane=0.05, far_plane=1000.0, aperture_radius=focal / (2 * 2.8),
                 focus_distance=float(np.linalg.norm(target - position)))
    sun = directional_lights(((-1.0, -2.0, -0.6), (1.0, 0.95, 0.85), 3.0))
    return concat(parts, mats, environment=(0.35, 0.45, 0.65), directional=sun, camera=cam, name="mixed")


def textured_scene(sphere_level=3, seed=6) -> Scene:
    """Base-color textures on a floor (uv tiled 3x, so repeat addressing wraps), a back wall and three
    spheres (latitude-longitude uv) of diffuse, rough conductor and glass material, one per filter,
    with a tangent-space normal map on the wall (scale 0.8) and two spheres (one with y flipped),
    under a sky and a sun."""
    r = np.random.default_rng(seed)
    checker = np.zeros((8, 8, 3), np.float32)
    checker[(np.arange(8)[:, None] + np.arange(8)[None, :]) % 2 == 0] = (0.9, 0.2, 0.1)
    checker[(np.arange(8)[:, None] + np.arange(8)[None, :]) % 2 == 1] = (0.1, 0.6, 0.9)
    noise = r.random((12, 20, 3)).astype(np.float32) * 1.2 - 0.1  # values outside [0, 1] exercise the clamp
    gy, gx = np.mgrid[0:16, 0:16] / 16.0
    n = np.stack([0.6 * np.sin(2 * math.pi * 3 * gx), 0.6 * np.cos(2 * math.pi * 2 * gy), np.ones_like(gx)], -1)
    bumps = (n / np.linalg.norm(n, axis=-1, keepdims=True) * 0.5 + 0.5).astype(np.float32)  # linear normal map
    textures = [checker, noise, bumps]
    parts, mats, uvs = [], [], []
    floor = np.array([[(-4, 0, -4), (-4, 0, 4), (4, 0, 4)], [(-4, 0, -4), (4, 0, 4), (4, 0, -4)]], np.float64)
    parts.append((floor, np.tile(np.array([0, 1, 0], np.float64), (2, 3, 1))))
    uvs.append((floor[..., [0, 2]] + 4) / 8 * 3 - 0.3)
    mats.append(material("diffuse", (0.9, 0.9, 0.9), base_color_texture=0, texture_filter="nearest"))
    wall = np.array([[(-4, 0, -3), (4, 0, -3), (4, 4, -3)], [(-4, 0, -3), (4, 4, -3), (-4, 4, -3)]], np.float64)
    parts.append((wall, np.tile(np.array([0, 0, 1], np.float64), (2, 3, 1))))
    uvs.append((wall[..., [0, 1]] + (4, 0)) / 8)
    mats.append(material("diffuse", (1.0, 1.0, 1.0), base_color_texture=1, texture_filter="bilinear", normal_texture=2,
                         normal_scale=0.8))
    unit = subdivided_sphere(sphere_level)
    lon = np.arctan2(unit[..., 2], unit[..., 0]) / (2 * math.pi) + 0.5
    lat = np.arccos(np.clip(unit[..., 1], -1, 1)) / math.pi
    for i, (kind, tex, filt) in enumerate((("diffuse", 1, "trilinear"), ("conductor", 0, "bilinear"),
                                            ("dielectric", 1, "nearest"))):
        parts.append((unit * 0.7 + (-1.6 + 1.6 * i, 0.7, 0.0), unit))
        uvs.append(np.stack([lon * 2, 1 - lat], -1))
        mats.append(material(kind, (0.95, 0.9, 0.85), roughness=0.3, base_color_texture=tex, texture_filter=filt,
                             normal_texture=2 if i < 2 else -1, normal_flip_y=i == 1))
    cam = Camera(position=(0, 2.0, 5.5), target=(0, 0.7, 0), up=(0, 1, 0), vertical_fov_degrees=45)
    sun = directional_lights(((-0.4, -1.0, -0.6), (1.0, 0.95, 0.9), 2.5))
    sc = concat(parts, mats, environment=(0.3, 0.35, 0.45), directional=sun, camera=cam, name="textured")
    sc.uvs = np.concatenate([np.asarray(u, np.float32).reshape(-1, 3, 2) for u in uvs])
    sc.textures = textures
    sc.__post_init__()
    return sc


def sky_map(width=64, height=32, seed=7, sun=(0.3, 0.28), sun_radiance=60.0):
    """A procedural HDR latitude-longitude sky: a vertical gradient, seeded clouds and a bright
    sun disk at (u, v) = `sun`; linear RGB float32 [height][width][3]."""
    r = np.random.default_rng(seed)
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    sky = np.zeros((height, width, 3), np.float32)
    sky[:] = np.stack([0.25 + 0.5 * v, 0.35 + 0.4 * v, 0.6 + 0.2 * v], -1)[:, None, :]
    sky *= (0.7 + 0.6 * r.random((height, width, 1))).astype(np.float32)
    sky[v > 0.5] *= 0.3  # darker ground half
    du = np.minimum(np.abs(u[None, :] - sun[0]), 1 - np.abs(u[None, :] - sun[0]))
    disk = (du ** 2 + (v[:, None] - sun[1]) ** 2) < (1.5 / height) ** 2
    sky[disk] = (sun_radiance, 0.9 * sun_radiance, 0.75 * sun_radiance)
    return sky


def envmap_scene(map_width=64, map_height=32, strength=1.5, rotation=30.0, sphere_level=3) -> Scene:
    """A floor and three spheres (diffuse, rough conductor, mirror) lit by an importance-sampled
    HDR environment map (with a sun disk), rotated, scaled and added to a small constant term."""
    unit = subdivided_sphere(sphere_level)
    parts, mats = [], []
    floor = np.array([[(-5, 0, -5), (-5, 0, 5), (5, 0, 5)], [(-5, 0, -5), (5, 0, 5), (5, 0, -5)]], np.float64)
    parts.append((floor, np.tile(np.array([0, 1, 0], np.float64), (2, 3, 1))))
    mats.append(material("diffuse", (0.6, 0.6, 0.6)))
    for i, (kind, rough) in enumerate((("diffuse", 0.3), ("conductor", 0.25), ("mirror", 0.0))):
        parts.append((unit * 0.6 + (-1.4 + 1.4 * i, 0.6, 0.0), unit))
        mats.append(material(kind, (0.85, 0.7, 0.5), roughness=rough))
    cam = Camera(position=(0, 1.6, 4.8), target=(0, 0.5, 0), up=(0, 1, 0), vertical_fov_degrees=45)
    return concat(parts, mats, environment=(0.02, 0.02, 0.03), camera=cam, name="envmap",
                  textures=[sky_map(map_width, map_height)], environment_texture=0, environment_strength=strength,
                  environment_rotation_degrees=rotation)


def furnace_spheres(sphere_level=3, roughness=(0.25, 0.65, 1.0), environment=(1.0, 1.0, 1.0), sun=True) -> Scene:
    """Two rows of rough conductor spheres, single-scattering above and Kulla-Conty compensated
    below (the layout of the reference's scenes/white-furnace.json), on a diffuse floor under a
    constant environment and, optionally, a sun."""
    unit = subdivided_sphere(sphere_level)
    parts, mats = [], []
    for row, compensate in enumerate((False, True)):
        for col, r in enumerate(roughness):
            center = (1.1 * (col - (len(roughness) - 1) / 2), 1.6 - 1.1 * row, 0.0)
            parts.append((unit * 0.5 + center, unit))
            mats.append(material("conductor", (0.95, 0.64, 0.54) if row else (1.0, 1.0, 1.0), roughness=r,
                                 compensate_energy=compensate))
    quad = np.array([[(-4, -0.5, -4), (-4, -0.5, 4), (4, -0.5, 4)], [(-4, -0.5, -4), (4, -0.5, 4), (4, -0.5, -4)]])
    parts.append((quad, np.tile(np.array([0, 1, 0], np.float64), (2, 3, 1))))
    mats.append(material("diffuse", (0.5, 0.5, 0.5)))
    cam = Camera(position=(0, 1, 4.4), target=(0, 1, 0), up=(0, 1, 0), vertical_fov_degrees=32, near_plane=0.05,
                 far_plane=100.0)
    lights = directional_lights(((-0.5, -1.0, -0.4), (1.0, 0.9, 0.8), 2.0)) if sun else None
    kw = dict(directional=lights) if sun else {}
    return concat(parts, mats, environment=environment, camera=cam, name="furnace", **kw)


# ----------------------------------------------------------------------------------------------
# Small analytic scenes used by the reference's own tests (path_tests.cpp:23-68).


def floor_scene(kind="diffuse", color=0.6, environment=(1, 1, 1), scale=4.0, translate=(0, 0, 0), y=0.0):
    s = scale
    t = np.asarray(translate, np.float64)
    quad = np.array([[(-s, y, -s), (-s, y, s), (s, y, s)], [(-s, y, -s), (s, y, s), (s, y, -s)]]) + t
    nrm = np.tile(np.array([0, 1, 0], np.float64), (2, 3, 1))
    cam = Camera(position=(0, 1, 0), target=(0, 0, 0), up=(0, 0, -1))
    return concat([(quad, nrm)], [material(kind, (color,) * 3)], environment=environment, camera=cam, name="floor")


# ----------------------------------------------------------------------------------------------
# Uneven multi-job distribution (batching / serving).


def job_stream(count=64, seed=4, side=(128, 1024), spp=(16, 1024), depth=(8, 16)):
    """Per-job (scene kind, width, height, spp, max depth); log-uniform sizes and sample counts."""
    r = np.random.default_rng(seed)
    jobs = []
    for i in range(count):
        s = int(round(math.exp(r.uniform(math.log(side[0]), math.log(side[1])))))
        n = int(round(math.exp(r.uniform(math.log(spp[0]), math.log(spp[1])))))
        d = int(r.integers(depth[0], depth[1] + 1))
        kind = ("cornell", "icosphere", "mixed")[int(r.integers(0, 3))]
        jobs.append(dict(index=i, scene=kind, width=s, height=s, spp=n, max_depth=d, seed=int(r.integers(0, 2**31))))
    return jobs
