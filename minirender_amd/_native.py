"""ctypes binding of libmrg.so (include/mrg.h). The library is the only compute path: importing a
GPU entry point without the built library raises, there is no CPU fallback."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB_PATH = os.environ.get("MRG_LIBRARY") or os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libmrg.so")

MRG_OK, MRG_INVALID_ARGUMENT, MRG_CANCELLED, MRG_NOT_ADMITTED, MRG_NONFINITE, MRG_DEVICE_ERROR, MRG_TIMEOUT = range(7)

STATE_NAMES = ["idle", "preparing", "running", "pausing", "paused", "finalizing", "cancelling", "cancelled",
               "completed", "failed"]
TERMINAL = {"idle", "cancelled", "completed", "failed"}


class MrgError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"[{code}] {message}")
        self.code = code


class InvalidArgument(MrgError, ValueError):
    pass


class NotAdmitted(MrgError):
    pass


class Material(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("base_color", C.c_float * 3), ("emission", C.c_float * 3),
                ("roughness", C.c_float), ("ior", C.c_float), ("double_sided", C.c_uint32), ("mirrored", C.c_uint32),
                ("compensate_energy", C.c_uint32), ("base_color_texture", C.c_uint32), ("texture_filter", C.c_uint32),
                ("normal_texture", C.c_uint32), ("normal_scale", C.c_float), ("normal_flip_y", C.c_uint32)]


class Texture(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgb", C.c_void_p)]


class Camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("target", C.c_float * 3), ("up", C.c_float * 3),
                ("vertical_fov_degrees", C.c_float), ("near_plane", C.c_float), ("far_plane", C.c_float),
                ("aperture_radius", C.c_float), ("focus_distance", C.c_float)]


class SceneDesc(C.Structure):
    _fields_ = [("triangle_count", C.c_uint64), ("positions", C.c_void_p), ("normals", C.c_void_p),
                ("instance", C.c_void_p), ("triangle", C.c_void_p), ("instance_count", C.c_uint32),
                ("materials", C.c_void_p), ("environment", C.c_float * 3), ("directional_count", C.c_uint32),
                ("directional", C.c_void_p), ("uvs", C.c_void_p), ("texture_count", C.c_uint32),
                ("textures", C.c_void_p), ("environment_texture", C.c_uint32), ("environment_strength", C.c_float),
                ("environment_rotation_degrees", C.c_float)]


class SceneInfo(C.Structure):
    _fields_ = [("triangles", C.c_uint64), ("nodes", C.c_uint64), ("leaves", C.c_uint64),
                ("maximum_depth", C.c_uint64), ("device_bytes", C.c_uint64), ("reference_count", C.c_uint64),
                ("build_seconds", C.c_double), ("content_hash", C.c_char * 33)]


class JobDesc(C.Structure):
    _fields_ = [("camera", Camera), ("width", C.c_uint32), ("height", C.c_uint32),
                ("samples_per_pixel", C.c_uint32), ("mode", C.c_uint32), ("seed", C.c_uint64),
                ("maximum_depth", C.c_uint32), ("russian_roulette", C.c_uint32), ("roulette_depth", C.c_uint32),
                ("pass_samples", C.c_uint32)]


class Progress(C.Structure):
    _fields_ = [("state", C.c_uint32), ("status", C.c_int32), ("job_id", C.c_uint64),
                ("recorded_samples", C.c_uint64), ("total_samples", C.c_uint64), ("completed_pixels", C.c_uint64),
                ("elapsed_seconds", C.c_double), ("queued_seconds", C.c_double),
                ("first_pass_seconds", C.c_double), ("completion_seconds", C.c_double),
                ("paths", C.c_uint64), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("scattering_events", C.c_uint64), ("roulette_terminations", C.c_uint64), ("error", C.c_char * 256)]


class DeviceOptions(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("path_capacity", C.c_uint64), ("schedule", C.c_uint32),
                ("max_jobs_in_flight", C.c_uint32), ("use_graph", C.c_uint32), ("block_size", C.c_uint32),
                ("persistent_blocks_per_cu", C.c_uint32), ("traversal", C.c_uint32), ("batches_in_flight", C.c_uint32),
                ("query_precision", C.c_uint32)]


class Hit(C.Structure):
    _fields_ = [("t", C.c_float), ("primitive", C.c_int32)]


class TraceStats(C.Structure):
    _fields_ = [("bounds_tests", C.c_uint64), ("triangle_tests", C.c_uint64)]


class KernelTimes(C.Structure):
    _fields_ = [("generate_ms", C.c_double), ("extend_ms", C.c_double), ("shade_ms", C.c_double),
                ("shadow_ms", C.c_double), ("accumulate_ms", C.c_double), ("batches", C.c_uint64),
                ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("extend_launches", C.c_uint64),
                ("shadow_launches", C.c_uint64), ("refill_launches", C.c_uint64)]


# Every entry point declared in include/mrg.h: name -> (restype, argtypes)
P = C.c_void_p
SIGNATURES = {
    "mrg_last_error": (C.c_char_p, []),
    "mrg_version": (C.c_char_p, []),
    "mrg_device_open": (C.c_int, [C.c_int, P, C.POINTER(P)]),
    "mrg_device_close": (C.c_int, [P]),
    "mrg_device_query": (C.c_int, [P, C.c_char_p, C.c_size_t, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                                   C.POINTER(C.c_uint64)]),
    "mrg_scene_upload": (C.c_int, [P, C.POINTER(SceneDesc), C.POINTER(P)]),
    "mrg_scene_release": (C.c_int, [P]),
    "mrg_scene_info_get": (C.c_int, [P, C.POINTER(SceneInfo)]),
    "mrg_intersect": (C.c_int, [P, P, C.c_uint64, C.c_float, C.c_float, P, P, P]),
    "mrg_occluded": (C.c_int, [P, P, C.c_uint64, C.c_float, C.c_float, P, P]),
    "mrg_intersect_device": (C.c_int, [P, P, C.c_uint64, C.c_float, C.c_float, P, P, C.POINTER(C.c_float)]),
    "mrg_intersect_noalloc": (C.c_int, [P, P, C.c_uint64, C.c_float, C.c_float, P, P, P]),
    "mrg_job_submit": (C.c_int, [P, P, C.POINTER(JobDesc), C.POINTER(P)]),
    "mrg_job_pause": (C.c_int, [P, C.POINTER(C.c_int)]),
    "mrg_job_resume": (C.c_int, [P, C.POINTER(C.c_int)]),
    "mrg_job_cancel": (C.c_int, [P]),
    "mrg_job_wait": (C.c_int, [P, C.c_int]),
    "mrg_job_progress": (C.c_int, [P, C.POINTER(Progress)]),
    "mrg_job_read_film": (C.c_int, [P, P, P]),
    "mrg_job_read_image": (C.c_int, [P, P]),
    "mrg_job_preview": (C.c_int, [P, C.c_uint32, P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                  C.POINTER(C.c_uint64)]),
    "mrg_job_release": (C.c_int, [P]),
    "mrg_render": (C.c_int, [P, P, C.POINTER(JobDesc), P, C.POINTER(Progress)]),
    "mrg_device_kernel_times": (C.c_int, [P, C.POINTER(KernelTimes), C.c_int]),
    "mrg_bvh_build_host": (C.c_int, [P, C.c_uint64, P]),
    "mrg_device_memory": (C.c_int, [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "mrg_conductor_energy_table": (C.c_int, [P]),
}

_lib = None


def lib():
    """Load libmrg.so; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m minirender_amd.build` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc == MRG_OK:
        return
    msg = lib().mrg_last_error().decode(errors="replace")
    if rc == MRG_INVALID_ARGUMENT:
        raise InvalidArgument(rc, msg)
    if rc == MRG_NOT_ADMITTED:
        raise NotAdmitted(rc, msg)
    raise MrgError(rc, msg)


def fill_scene_desc(d, scene, keep):
    """Fill a scene descriptor (SceneDesc here, or the oracle's layout-identical one) from a Scene;
    `keep` collects the arrays whose memory the descriptor points into."""
    arrays = [np.ascontiguousarray(scene.positions, np.float32), np.ascontiguousarray(scene.normals, np.float32),
              np.ascontiguousarray(scene.instance, np.uint32), np.ascontiguousarray(scene.triangle, np.uint32),
              np.ascontiguousarray(scene.materials), np.ascontiguousarray(scene.directional)]
    keep.extend(arrays)
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    d.triangle_count = scene.triangle_count
    d.positions, d.normals, d.instance, d.triangle = (ptr(a) for a in arrays[:4])
    d.instance_count = len(scene.materials)
    d.materials = ptr(arrays[4])
    d.environment[:] = list(scene.environment)
    d.directional_count = len(scene.directional)
    d.directional = ptr(arrays[5])
    if scene.uvs is not None:
        uv = np.ascontiguousarray(scene.uvs, np.float32)
        keep.append(uv)
        d.uvs = ptr(uv)
    if scene.textures:
        tex = (Texture * len(scene.textures))()
        for i, t in enumerate(scene.textures):
            t = np.ascontiguousarray(t, np.float32)
            keep.append(t)
            tex[i].height, tex[i].width = t.shape[0], t.shape[1]
            tex[i].rgb = ptr(t)
        keep.append(tex)
        d.texture_count = len(scene.textures)
        d.textures = C.cast(tex, C.c_void_p)
    d.environment_texture = scene.environment_texture + 1
    d.environment_strength = scene.environment_strength
    d.environment_rotation_degrees = scene.environment_rotation_degrees
