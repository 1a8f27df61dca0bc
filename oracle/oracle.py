"""TEST INFRASTRUCTURE ONLY — ctypes wrapper around the CPU oracle (oracle/build/liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline may import this module. It is the
checker, never the product: the product path (minirender_amd) never loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")


NATIVE_PATH = os.path.join(HERE, "build", "liboracle_native.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def use_native():
    """Builds the oracle with -march=native for the host this process runs on and loads that build
    instead of the portable one (the timed CPU baseline). Call before the first use of the oracle."""
    global LIB_PATH
    if _lib is not None:
        raise RuntimeError("the oracle library is already loaded")
    subprocess.run(["make", "-s", "-B", "-C", HERE, "native"], check=True)
    LIB_PATH = NATIVE_PATH


class _SceneDesc(C.Structure):
    _fields_ = [
        ("triangle_count", C.c_uint64),
        ("positions", C.c_void_p),
        ("normals", C.c_void_p),
        ("instance", C.c_void_p),
        ("triangle", C.c_void_p),
        ("instance_count", C.c_uint32),
        ("materials", C.c_void_p),
        ("environment", C.c_float * 3),
        ("directional_count", C.c_uint32),
        ("directional", C.c_void_p),
        ("uvs", C.c_void_p),
        ("texture_count", C.c_uint32),
        ("textures", C.c_void_p),
        ("environment_texture", C.c_uint32),
        ("environment_strength", C.c_float),
        ("environment_rotation_degrees", C.c_float),
    ]


class _Texture(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgb", C.c_void_p)]


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("paths", "shadow_rays", "scattering_events", "roulette_terminations",
                                          "bounds_tests", "triangle_tests", "extension_rays",
                                          "shadow_bounds_tests", "shadow_triangle_tests")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class _Render(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32), ("seed", C.c_uint64),
                ("mode", C.c_uint32), ("maximum_depth", C.c_uint32), ("russian_roulette", C.c_uint32),
                ("roulette_depth", C.c_uint32), ("x0", C.c_uint32), ("y0", C.c_uint32), ("x1", C.c_uint32),
                ("y1", C.c_uint32)]


MODES = {"bsdf": 0, "light": 1, "mis": 2}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_last_error.restype = C.c_char_p
        L.orc_scene_new.restype = P
        L.orc_scene_new.argtypes = [C.POINTER(_SceneDesc)]
        L.orc_scene_free.argtypes = [P]
        L.orc_scene_build_stats.argtypes = [P, P]
        L.orc_intersect.argtypes = [P, C.c_int, P, C.c_uint64, C.c_double, C.c_double, C.c_int, P, P, P, P, P, P]
        L.orc_occluded.argtypes = [P, C.c_int, P, C.c_uint64, C.c_double, C.c_double, C.c_int, P, P]
        L.orc_render.argtypes = [P, C.c_int, P, P, C.c_int, C.c_uint32, C.c_uint32, P, P, P]
        L.orc_sample_pixel.argtypes = [P, C.c_int, P, P, C.c_uint32, C.c_uint32, C.c_uint64, P, P]
        L.orc_trace.argtypes = [P, C.c_int, P, P, C.c_double, C.c_double, C.c_int, C.c_uint64, C.c_uint64,
                                C.c_uint64, P, P]
        L.orc_pcg.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, P]
        L.orc_for_sample.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, P]
        L.orc_next_1d.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, P, P]
        L.orc_warp.argtypes = [C.c_int, C.c_int, P, C.c_uint64, P]
        L.orc_sincos_turns.argtypes = [P, C.c_uint64, P, P]
        L.orc_discrete.argtypes = [P, C.c_uint64, C.c_double, P, P, P, P]
        L.orc_power_heuristic.restype = C.c_double
        L.orc_power_heuristic.argtypes = [C.c_double, C.c_double]
        L.orc_fresnel.restype = C.c_double
        L.orc_fresnel.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_bsdf_sample.argtypes = [P, C.c_int, C.c_int, P, C.c_double, P, P]
        L.orc_bsdf_eval.argtypes = [P, C.c_int, C.c_int, P, P, P, P]
        L.orc_light_sample.argtypes = [P, C.c_int, P, C.c_double, P, P]
        L.orc_light_selection.restype = C.c_double
        L.orc_light_selection.argtypes = [P, C.c_uint64]
        L.orc_light_count.restype = C.c_uint64
        L.orc_light_count.argtypes = [P]
        L.orc_camera_ray.argtypes = [P, C.c_int, P, C.c_double, P, P]
        L.orc_energy_table.argtypes = [P]
        L.orc_env_eval.argtypes = [P, C.c_int, P, C.c_uint64, P, P]
        L.orc_env_sample.argtypes = [P, C.c_int, P, C.c_uint64, P]
        L.orc_env_integral.restype = C.c_double
        L.orc_env_integral.argtypes = [P, C.c_int]
        L.orc_atan2_turns.argtypes = [P, P, C.c_uint64, P]
        L.orc_surface_hit.argtypes = [P, C.c_int, P, P]
        L.orc_shadow_segment.argtypes = [P, C.c_int, P, C.c_double, P, P]
        L.orc_texture_sample.argtypes = [C.c_uint32, C.c_uint32, P, P, C.c_uint64, C.c_float, C.c_uint32, P]
        L.orc_texture_level.argtypes = [C.c_uint32, C.c_uint32, P, C.c_uint32, P, P, P]
        L.orc_integrate_ibl_brdf.argtypes = [C.c_double, C.c_double, C.c_uint32, C.c_uint64, P]
        L.orc_average_energy_of_ramp.restype = C.c_double
        L.orc_average_energy_of_ramp.argtypes = [P, C.c_uint32]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())


def _prec(precision):
    if precision not in (32, 64):
        raise ValueError("precision must be 32 or 64")
    return precision


class OracleScene:
    """The reference SceneBvh + LightSampler restated on the CPU, for one flattened scene."""

    def __init__(self, scene):
        self._keep = [np.ascontiguousarray(scene.positions, np.float32), np.ascontiguousarray(scene.normals, np.float32),
                      np.ascontiguousarray(scene.instance, np.uint32), np.ascontiguousarray(scene.triangle, np.uint32),
                      np.ascontiguousarray(scene.materials), np.ascontiguousarray(scene.directional)]
        d = _SceneDesc()
        d.triangle_count = scene.triangle_count
        d.positions, d.normals, d.instance, d.triangle = (_ptr(a) for a in self._keep[:4])
        d.instance_count = len(scene.materials)
        d.materials = _ptr(self._keep[4])
        d.environment[:] = list(scene.environment)
        d.directional_count = len(scene.directional)
        d.directional = _ptr(self._keep[5])
        if getattr(scene, "uvs", None) is not None:
            uv = np.ascontiguousarray(scene.uvs, np.float32)
            self._keep.append(uv)
            d.uvs = _ptr(uv)
        textures = getattr(scene, "textures", [])
        if textures:
            tex = (_Texture * len(textures))()
            for i, t in enumerate(textures):
                t = np.ascontiguousarray(t, np.float32)
                self._keep.append(t)
                tex[i].height, tex[i].width = t.shape[0], t.shape[1]
                tex[i].rgb = _ptr(t)
            self._keep.append(tex)
            d.texture_count = len(textures)
            d.textures = C.cast(tex, C.c_void_p)
        d.environment_texture = getattr(scene, "environment_texture", -1) + 1
        d.environment_strength = getattr(scene, "environment_strength", 1.0)
        d.environment_rotation_degrees = getattr(scene, "environment_rotation_degrees", 0.0)
        self.handle = lib().orc_scene_new(C.byref(d))
        if not self.handle:
            raise ValueError(lib().orc_last_error().decode())
        self.scene = scene

    def __del__(self):
        if getattr(self, "handle", None):
            lib().orc_scene_free(self.handle)
            self.handle = None

    def build_stats(self):
        out = np.zeros(4, np.uint64)
        lib().orc_scene_build_stats(self.handle, _ptr(out))
        return dict(triangles=int(out[0]), nodes=int(out[1]), leaves=int(out[2]), maximum_depth=int(out[3]))

    def intersect(self, rays, tmin=0.0, tmax=np.inf, precision=64, threads=None, second=False):
        rays = np.ascontiguousarray(rays, np.float64).reshape(-1, 6)
        n = len(rays)
        prim = np.empty(n, np.int32)
        t = np.empty(n)
        bary = np.empty((n, 3))
        t2 = np.empty(n) if second else None
        p2 = np.empty(n, np.int32) if second else None
        st = Stats()
        _check(lib().orc_intersect(self.handle, _prec(precision), _ptr(rays), n, tmin, tmax,
                                   threads or os.cpu_count(), _ptr(prim), _ptr(t), _ptr(t2), _ptr(p2), _ptr(bary),
                                   C.byref(st)))
        out = dict(prim=prim, t=t, bary=bary, stats=st.as_dict())
        if second:
            out.update(t2=t2, prim2=p2)
        return out

    def occluded(self, rays, tmin=0.0, tmax=np.inf, precision=64, threads=None):
        rays = np.ascontiguousarray(rays, np.float64).reshape(-1, 6)
        out = np.empty(len(rays), np.uint8)
        st = Stats()
        _check(lib().orc_occluded(self.handle, _prec(precision), _ptr(rays), len(rays), tmin, tmax,
                                  threads or os.cpu_count(), _ptr(out), C.byref(st)))
        return out.astype(bool), st.as_dict()

    @staticmethod
    def _render_settings(width, height, spp, seed, mode, max_depth, rr, rr_depth, crop=None):
        r = _Render()
        r.width, r.height, r.spp, r.seed = width, height, spp, seed
        r.mode = MODES[mode] if isinstance(mode, str) else mode
        r.maximum_depth, r.russian_roulette, r.roulette_depth = max_depth, int(bool(rr)), rr_depth
        if crop is not None:
            r.x0, r.y0, r.x1, r.y1 = crop
        return r

    def render(self, camera, width, height, spp, seed=1, mode="mis", max_depth=8, rr=True, rr_depth=3,
               precision=64, threads=None, crop=None, tile=16, batch=4, sums=False):
        """Session-order render (tiles of `tile`^2, batches of `batch` samples) of an optional crop."""
        rs = self._render_settings(width, height, spp, seed, mode, max_depth, rr, rr_depth, crop)
        x0, y0, x1, y1 = crop if crop is not None else (0, 0, width, height)
        img = np.empty((y1 - y0, x1 - x0, 3), np.float32)
        s = np.empty((y1 - y0, x1 - x0, 3), np.float64) if sums else None
        st = Stats()
        cam = camera.record()
        _check(lib().orc_render(self.handle, _prec(precision), _ptr(cam), C.byref(rs), threads or os.cpu_count(),
                                tile, batch, _ptr(img), _ptr(s), C.byref(st)))
        return (img, s, st.as_dict()) if sums else (img, st.as_dict())

    def sample_pixel(self, camera, width, height, x, y, sample, seed=1, mode="mis", max_depth=8, rr=True,
                     rr_depth=3, precision=64):
        rs = self._render_settings(width, height, 1, seed, mode, max_depth, rr, rr_depth)
        out = np.empty(3)
        st = Stats()
        cam = camera.record()
        _check(lib().orc_sample_pixel(self.handle, _prec(precision), _ptr(cam), C.byref(rs), x, y, sample, _ptr(out),
                                      C.byref(st)))
        return out, st.as_dict()

    def trace(self, origin, direction, tmin=0.0, tmax=np.inf, sampler=("seed", 0, 1), mode="mis", max_depth=8,
              rr=True, rr_depth=3, precision=64):
        """PathTracer::trace with Sampler(seed, seq) (("seed", s, q)) or for_sample (("sample", s, p, i))."""
        rs = self._render_settings(1, 1, 1, 0, mode, max_depth, rr, rr_depth)
        ray = np.array(list(origin) + list(direction), np.float64)
        out = np.empty(3)
        st = Stats()
        kind = 1 if sampler[0] == "sample" else 0
        a, b = sampler[1], sampler[2]
        c = sampler[3] if len(sampler) > 3 else 0
        _check(lib().orc_trace(self.handle, _prec(precision), C.byref(rs), _ptr(ray), tmin, tmax, kind, a, b, c,
                               _ptr(out), C.byref(st)))
        return out, st.as_dict()

    def light_sample(self, reference, light_u, u, precision=64):
        out = np.empty(11)
        ref = np.asarray(reference, np.float64)
        uu = np.asarray(u, np.float64)
        rc = lib().orc_light_sample(self.handle, _prec(precision), _ptr(ref), light_u, _ptr(uu), _ptr(out))
        if rc < 0:
            raise ValueError(lib().orc_last_error().decode())
        if rc == 0:
            return None
        return dict(direction=out[0:3], radiance=out[3:6], distance=out[6], probability=out[7],
                    discrete=bool(out[8]), kind=int(out[9]), emitter=int(out[10]))

    def surface_hit(self, origin, direction, precision=64):
        """SceneBvh::intersect + surface_hit: dict(position, geometric_normal, shading_normal, uv, distance,
        front_face, primitive) or None."""
        ray = np.ascontiguousarray(np.concatenate([origin, direction]), np.float64)
        out = np.zeros(14)
        rc = lib().orc_surface_hit(self.handle, _prec(precision), _ptr(ray), _ptr(out))
        if rc < 0:
            raise ValueError(lib().orc_last_error().decode())
        if rc == 0:
            return None
        return dict(position=out[0:3], geometric_normal=out[3:6], shading_normal=out[6:9], uv=out[9:11],
                    distance=out[11], front_face=bool(out[12]), primitive=int(out[13]))

    def shadow_segment(self, origin, light_u, u, precision=64):
        """The shadow segment the path tracer traces toward a light sampled from `origin`
        (path_tracer.cpp:126-151): dict(ray (6,), limit, light_distance, occluded) or None."""
        o = np.asarray(origin, np.float64)
        uu = np.asarray(u, np.float64)
        out = np.zeros(9)
        rc = lib().orc_shadow_segment(self.handle, _prec(precision), _ptr(o), light_u, _ptr(uu), _ptr(out))
        if rc < 0:
            raise ValueError(lib().orc_last_error().decode())
        if rc == 0:
            return None
        return dict(ray=out[0:6].copy(), limit=float(out[6]), light_distance=float(out[7]), occluded=bool(out[8]))

    def env_eval(self, directions, precision=64):
        """EnvironmentLight::evaluate and ::pdf of unit directions [n][3]."""
        d = np.ascontiguousarray(directions, np.float64).reshape(-1, 3)
        rad, pdf = np.zeros((len(d), 3)), np.zeros(len(d))
        _check(lib().orc_env_eval(self.handle, _prec(precision), _ptr(d), len(d), _ptr(rad), _ptr(pdf)))
        return rad, pdf

    def env_sample(self, u, precision=64):
        """EnvironmentLight::sample of u [n][2] -> (directions [n][3], radiance [n][3], pdf [n])."""
        uu = np.ascontiguousarray(u, np.float64).reshape(-1, 2)
        out = np.zeros((len(uu), 7))
        _check(lib().orc_env_sample(self.handle, _prec(precision), _ptr(uu), len(uu), _ptr(out)))
        return out[:, :3], out[:, 3:6], out[:, 6]

    def env_integral(self, precision=64):
        return lib().orc_env_integral(self.handle, _prec(precision))

    def light_selection(self):
        return np.array([lib().orc_light_selection(self.handle, i) for i in range(lib().orc_light_count(self.handle))])


# ----------------------------------------------------------------------------------------------
# Free functions.


def pcg(seed, sequence, n):
    out = np.empty(n, np.uint32)
    lib().orc_pcg(seed, sequence, n, _ptr(out))
    return out


def for_sample_u32(seed, pixel, sample, n):
    out = np.empty(n, np.uint32)
    lib().orc_for_sample(seed, pixel, sample, n, _ptr(out))
    return out


def next_1d(seed, pixel, sample, n):
    a = np.empty(n)
    b = np.empty(n, np.float32)
    lib().orc_next_1d(seed, pixel, sample, n, _ptr(a), _ptr(b))
    return a, b


WARPS = {"disk": 0, "sphere": 1, "hemisphere": 2, "cosine": 3, "triangle": 4}


def warp(which, u, precision=64):
    u = np.ascontiguousarray(u, np.float64).reshape(-1, 2)
    out = np.empty((len(u), 3))
    _check(lib().orc_warp(WARPS[which], _prec(precision), _ptr(u), len(u), _ptr(out)))
    return out


def sincos_turns(t):
    t = np.ascontiguousarray(t, np.float32)
    s = np.empty_like(t)
    c = np.empty_like(t)
    lib().orc_sincos_turns(_ptr(t), t.size, _ptr(s), _ptr(c))
    return s, c


def discrete(weights, u):
    w = np.ascontiguousarray(weights, np.float64)
    p = np.empty(len(w))
    idx = C.c_uint64()
    mass = C.c_double()
    rem = C.c_double()
    _check(lib().orc_discrete(_ptr(w), len(w), u, _ptr(p), C.byref(idx), C.byref(mass), C.byref(rem)))
    return p, int(idx.value), mass.value, rem.value


def power_heuristic(a, b):
    v = lib().orc_power_heuristic(a, b)
    if np.isnan(v):
        raise ValueError(lib().orc_last_error().decode())
    return v


def fresnel(cosine, eta_i, eta_t):
    v = lib().orc_fresnel(cosine, eta_i, eta_t)
    if np.isnan(v):
        raise ValueError(lib().orc_last_error().decode())
    return v


def bsdf_sample(mat, wo, component, u, front_face=True, precision=64):
    m = np.ascontiguousarray(np.asarray(mat).reshape(()))
    out = np.empty(10)
    wo = np.asarray(wo, np.float64)
    uu = np.asarray(u, np.float64)
    rc = lib().orc_bsdf_sample(_ptr(m), int(front_face), _prec(precision), _ptr(wo), component, _ptr(uu), _ptr(out))
    if rc < 0:
        raise ValueError(lib().orc_last_error().decode())
    if rc == 0:
        return None
    return dict(direction=out[0:3], weight=out[3:6], probability=out[6], discrete=bool(out[7]), event=int(out[8]),
                eta=out[9])


def bsdf_eval(mat, wo, wi, front_face=True, precision=64):
    m = np.ascontiguousarray(np.asarray(mat).reshape(()))
    v = np.empty(3)
    pdf = C.c_double()
    wo = np.asarray(wo, np.float64)
    wi = np.asarray(wi, np.float64)
    _check(lib().orc_bsdf_eval(_ptr(m), int(front_face), _prec(precision), _ptr(wo), _ptr(wi), _ptr(v), C.byref(pdf)))
    return v, pdf.value


def camera_ray(camera, uv, aspect, lens=(0.5, 0.5), precision=64):
    out = np.empty(6)
    cam = camera.record()
    uvv = np.asarray(uv, np.float64)
    ln = np.asarray(lens, np.float64)
    _check(lib().orc_camera_ray(_ptr(cam), _prec(precision), _ptr(uvv), aspect, _ptr(ln), _ptr(out)))
    return out[:3], out[3:]


def energy_table():
    """The Kulla-Conty table [64][64][3]: (A, B, row average) per (roughness row, cos column)."""
    out = np.zeros((64, 64, 3), np.float32)
    lib().orc_energy_table(_ptr(out))
    return out


def integrate_ibl_brdf(mu, roughness, samples, seed=101):
    out = np.zeros(2, np.float64)
    lib().orc_integrate_ibl_brdf(float(mu), float(roughness), int(samples), int(seed), _ptr(out))
    return out


def average_energy_of_ramp(e):
    e = np.ascontiguousarray(e, np.float32)
    return lib().orc_average_energy_of_ramp(_ptr(e), len(e))


FILTERS = {"nearest": 0, "bilinear": 1, "trilinear": 2}


def texture_sample(image, uv, lod=0.0, filter="trilinear"):
    """Texture::sample (texture.cpp:115-137) of a linear [h][w][3] image at uvs [n][2] (float)."""
    im = np.ascontiguousarray(image, np.float32)
    uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
    out = np.zeros((len(uv), 3), np.float32)
    _check(lib().orc_texture_sample(im.shape[1], im.shape[0], _ptr(im), _ptr(uv), len(uv), float(lod), FILTERS[filter],
                                    _ptr(out)))
    return out


def texture_level(image, level):
    """Mip level `level` of the texture made from a linear [h][w][3] image."""
    im = np.ascontiguousarray(image, np.float32)
    w, h = C.c_uint32(), C.c_uint32()
    _check(lib().orc_texture_level(im.shape[1], im.shape[0], _ptr(im), level, C.byref(w), C.byref(h), None))
    out = np.zeros((h.value, w.value, 3), np.float32)
    _check(lib().orc_texture_level(im.shape[1], im.shape[0], _ptr(im), level, C.byref(w), C.byref(h), _ptr(out)))
    return out


def atan2_turns(y, x):
    """The FP32 twin's atan2(y, x) / (2 pi) (float32 in, float32 out)."""
    y = np.ascontiguousarray(y, np.float32).reshape(-1)
    x = np.ascontiguousarray(x, np.float32).reshape(-1)
    out = np.zeros(len(y), np.float32)
    lib().orc_atan2_turns(_ptr(y), _ptr(x), len(y), _ptr(out))
    return out
