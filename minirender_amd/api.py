"""Host-side mirror of the reference's render interface over the C ABI.

* ``Device``            one GPU (one process per GPU); owns the job loop of libmrg.
* ``SceneBvh``          a resident scene: ``intersect`` / ``occluded`` as SceneBvh (bvh.hpp:45-108).
* ``RenderSettings``, ``IntegratorSettings`` (path_tracer.hpp:21-50) and ``RenderSession``
  (render_session.hpp:82-115): start / pause / resume / cancel / wait / wait_for / progress /
  preview / result, with the reference's states and error behaviour.
* ``render``            synchronous PathTracer::render equivalent (path_tracer.cpp:388-405).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import threading

import numpy as np

from . import _native as N
from .scene import Camera, Scene

MODES = {"bsdf": 0, "light": 1, "mis": 2}


@dataclasses.dataclass
class RenderSettings:
    width: int = 256
    height: int = 256
    samples_per_pixel: int = 64
    seed: int = 1

    def validate(self):
        if not (0 < self.width <= 16384 and 0 < self.height <= 16384 and self.width * self.height <= 64 * 1024 * 1024
                and self.samples_per_pixel > 0):
            raise ValueError("Render dimensions or sample count exceed the image budget")


@dataclasses.dataclass
class IntegratorSettings:
    mode: str = "mis"
    maximum_depth: int = 8
    russian_roulette: bool = True
    roulette_depth: int = 3

    def validate(self):
        if self.mode not in MODES:
            raise ValueError("Sampling mode must be bsdf, light or mis")
        if self.maximum_depth > 128 or self.roulette_depth > 128:
            raise ValueError("Path and roulette depths must not exceed 128")


@dataclasses.dataclass
class Progress:
    state: str
    job_id: int
    recorded_samples: int
    total_samples: int
    completed_pixels: int
    elapsed_seconds: float
    queued_seconds: float
    first_pass_seconds: float
    completion_seconds: float
    statistics: dict
    error: str


def _progress(p: N.Progress) -> Progress:
    return Progress(N.STATE_NAMES[p.state], int(p.job_id), int(p.recorded_samples), int(p.total_samples),
                    int(p.completed_pixels), p.elapsed_seconds, p.queued_seconds, p.first_pass_seconds,
                    p.completion_seconds,
                    dict(paths=int(p.paths), extension_rays=int(p.extension_rays), shadow_rays=int(p.shadow_rays),
                         scattering_events=int(p.scattering_events),
                         roulette_terminations=int(p.roulette_terminations)),
                    p.error.decode(errors="replace"))


class Device:
    """One MI355X (or any gfx950 device) with its own scheduler thread inside libmrg."""

    def __init__(self, ordinal=0, schedule="plain", hbm_budget_bytes=0, path_capacity=0, max_jobs_in_flight=0,
                 use_graph=False, persistent_blocks_per_cu=0, traversal="auto", batches_in_flight=0, query_precision=64):
        L = N.lib()
        o = N.DeviceOptions()
        o.hbm_budget_bytes = hbm_budget_bytes
        o.path_capacity = path_capacity
        o.schedule = {"plain": 0, "batched": 1, "plain_fill": 2}[schedule]
        o.max_jobs_in_flight = max_jobs_in_flight
        o.use_graph = int(bool(use_graph))
        o.persistent_blocks_per_cu = persistent_blocks_per_cu
        o.traversal = {"auto": 0, "split": 1, "fused": 2, "refill": 3}[traversal]
        o.batches_in_flight = batches_in_flight
        o.query_precision = query_precision
        h = C.c_void_p()
        N.check(L.mrg_device_open(ordinal, C.byref(o), C.byref(h)))
        self.handle = h
        self.schedule = schedule

    def close(self):
        if getattr(self, "handle", None):
            N.check(N.lib().mrg_device_close(self.handle))
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def query(self):
        name = C.create_string_buffer(256)
        cus, hbm, l2 = C.c_uint32(), C.c_uint64(), C.c_uint64()
        N.check(N.lib().mrg_device_query(self.handle, name, 256, C.byref(cus), C.byref(hbm), C.byref(l2)))
        return dict(name=name.value.decode(), compute_units=cus.value, hbm_bytes=hbm.value, l2_bytes=l2.value)

    def memory(self):
        b, r, p = C.c_uint64(), C.c_uint64(), C.c_uint64()
        N.check(N.lib().mrg_device_memory(self.handle, C.byref(b), C.byref(r), C.byref(p)))
        return dict(budget_bytes=b.value, resident_bytes=r.value, pool_bytes=p.value)

    def upload(self, scene: Scene) -> "SceneBvh":
        return SceneBvh(self, scene)

    def kernel_times(self, reset=False):
        t = N.KernelTimes()
        N.check(N.lib().mrg_device_kernel_times(self.handle, C.byref(t), int(reset)))
        return {f: getattr(t, f) for f, _ in N.KernelTimes._fields_}


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class SceneBvh:
    """A scene resident on the device (deduplicated by content) with the SceneBvh query interface."""

    def __init__(self, device: Device, scene: Scene):
        self.device = device
        self.scene = scene
        keep = []
        d = N.SceneDesc()
        N.fill_scene_desc(d, scene, keep)
        h = C.c_void_p()
        N.check(N.lib().mrg_scene_upload(device.handle, C.byref(d), C.byref(h)))
        self.handle = h

    def release(self):
        # valid after Device.close too: the library keeps a closed device's bookkeeping until its last
        # scene and job are released
        if getattr(self, "handle", None):
            N.check(N.lib().mrg_scene_release(self.handle))
        self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    def info(self):
        i = N.SceneInfo()
        N.check(N.lib().mrg_scene_info_get(self.handle, C.byref(i)))
        return {f: (getattr(i, f).decode() if f == "content_hash" else getattr(i, f)) for f, _ in N.SceneInfo._fields_}

    def intersect(self, rays, tmin=0.0, tmax=np.inf, barycentric=False, stats=False):
        """Closest hits for [n][6] FP32 rays: (t, primitive) with primitive -1 for a miss."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        n = len(rays)
        hits = np.empty(n, np.dtype([("t", "<f4"), ("primitive", "<i4")]))
        bary = np.empty((n, 3), np.float32) if barycentric else None
        st = N.TraceStats()
        N.check(N.lib().mrg_intersect(self.handle, _ptr(rays), n, tmin, tmax, _ptr(hits), _ptr(bary),
                                      C.byref(st) if stats else None))
        out = dict(t=hits["t"].copy(), primitive=hits["primitive"].copy())
        if barycentric:
            out["barycentric"] = bary
        if stats:
            out["stats"] = dict(bounds_tests=st.bounds_tests, triangle_tests=st.triangle_tests)
        return out

    def occluded(self, rays, tmin=0.0, tmax=np.inf):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        out = np.empty(len(rays), np.uint8)
        N.check(N.lib().mrg_occluded(self.handle, _ptr(rays), len(rays), tmin, tmax, _ptr(out), None))
        return out.astype(bool)

    def intersect_device(self, rays_ptr, n, hits_ptr, tmin=0.0, tmax=np.inf, stats=False):
        """Closest hits for rays already resident in HBM; returns the kernel time in ms."""
        ms = C.c_float()
        st = N.TraceStats()
        N.check(N.lib().mrg_intersect_device(self.handle, C.c_void_p(rays_ptr), n, tmin, tmax, C.c_void_p(hits_ptr),
                                             C.byref(st) if stats else None, C.byref(ms)))
        return (ms.value, dict(bounds_tests=st.bounds_tests, triangle_tests=st.triangle_tests)) if stats else ms.value


def job_desc(camera: Camera, render: RenderSettings, integrator: IntegratorSettings, pass_samples=0) -> N.JobDesc:
    render.validate()
    integrator.validate()
    camera.validate()
    d = N.JobDesc()
    c = d.camera
    c.position[:] = list(camera.position)
    c.target[:] = list(camera.target)
    c.up[:] = list(camera.up)
    c.vertical_fov_degrees = camera.vertical_fov_degrees
    c.near_plane = camera.near_plane
    c.far_plane = camera.far_plane
    c.aperture_radius = camera.aperture_radius
    c.focus_distance = camera.focus_distance
    d.width, d.height = render.width, render.height
    d.samples_per_pixel = render.samples_per_pixel
    d.seed = render.seed
    d.mode = MODES[integrator.mode]
    d.maximum_depth = integrator.maximum_depth
    d.russian_roulette = int(bool(integrator.russian_roulette))
    d.roulette_depth = integrator.roulette_depth
    d.pass_samples = pass_samples
    return d


class Job:
    """One submitted render job (C handle wrapper)."""

    def __init__(self, bvh: SceneBvh, camera, render, integrator, pass_samples=0):
        self.bvh = bvh
        self.render = render
        d = job_desc(camera, render, integrator, pass_samples)
        h = C.c_void_p()
        N.check(N.lib().mrg_job_submit(bvh.device.handle, bvh.handle, C.byref(d), C.byref(h)))
        self.handle = h

    def progress(self) -> Progress:
        p = N.Progress()
        N.check(N.lib().mrg_job_progress(self.handle, C.byref(p)))
        return _progress(p)

    def wait(self, timeout_ms=-1) -> bool:
        rc = N.lib().mrg_job_wait(self.handle, int(timeout_ms))
        if rc == N.MRG_TIMEOUT:
            return False
        N.check(rc)
        return True

    def pause(self) -> bool:
        a = C.c_int()
        N.check(N.lib().mrg_job_pause(self.handle, C.byref(a)))
        return bool(a.value)

    def resume(self) -> bool:
        a = C.c_int()
        N.check(N.lib().mrg_job_resume(self.handle, C.byref(a)))
        return bool(a.value)

    def cancel(self):
        N.check(N.lib().mrg_job_cancel(self.handle))

    def film(self):
        w, h = self.render.width, self.render.height
        sums = np.empty((h, w, 3), np.float64)
        counts = np.empty((h, w), np.uint32)
        N.check(N.lib().mrg_job_read_film(self.handle, _ptr(sums), _ptr(counts)))
        return sums, counts

    def image(self):
        img = np.empty((self.render.height, self.render.width, 3), np.float32)
        N.check(N.lib().mrg_job_read_image(self.handle, _ptr(img)))
        return img

    def preview(self, maximum_dimension=1024):
        w, h, rec = C.c_uint32(), C.c_uint32(), C.c_uint64()
        N.check(N.lib().mrg_job_preview(self.handle, maximum_dimension, None, C.byref(w), C.byref(h), None))
        img = np.empty((h.value, w.value, 3), np.float32)
        N.check(N.lib().mrg_job_preview(self.handle, maximum_dimension, _ptr(img), C.byref(w), C.byref(h),
                                        C.byref(rec)))
        return img, int(rec.value)

    def release(self):
        if getattr(self, "handle", None):
            N.check(N.lib().mrg_job_release(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class RenderSession:
    """The reference RenderSession contract (render_session.hpp:82-115) on the GPU backend.

    Lifecycle calls come from one control thread; progress/preview may be called concurrently."""

    def __init__(self, device: Device):
        self.device = device
        self._job = None
        self._bvh = None
        self._result = None
        self._lock = threading.Lock()

    def start(self, scene: Scene, render: RenderSettings, integrator: IntegratorSettings = None, camera: Camera = None,
              pass_samples=0):
        if scene is None:
            raise ValueError("Render session requires a scene snapshot")
        integrator = integrator or IntegratorSettings()
        render.validate()
        integrator.validate()
        if self._job is not None and self.progress().state not in N.TERMINAL:
            raise RuntimeError("Render session is still active")
        if camera is None:
            scene.camera.validate()
        if self._job is not None:
            self._job.release()
            self._job = None
        self._result = None
        self._failure = None
        self._jobs_started = getattr(self, "_jobs_started", 0) + 1
        try:
            # the reference builds the BVH in the controller thread, so scene-content errors surface
            # as an asynchronous `failed` state rather than from start() (render_session.cpp:494-505)
            self._bvh = self.device.upload(scene)
        except N.MrgError as e:
            self._failure = Progress("failed", 0, 0, render.width * render.height * render.samples_per_pixel, 0,
                                     0.0, 0.0, 0.0, 0.0, {}, str(e))
            return
        self._job = Job(self._bvh, camera or scene.camera, render, integrator, pass_samples)

    def pause(self) -> bool:
        return self._job.pause() if self._job else False

    def resume(self) -> bool:
        return self._job.resume() if self._job else False

    def cancel(self):
        if self._job:
            self._job.cancel()

    def wait(self):
        if self._job:
            self._job.wait(-1)

    def wait_for(self, milliseconds) -> bool:
        return self._job.wait(milliseconds) if self._job else True

    def progress(self) -> Progress:
        if getattr(self, "_failure", None) is not None:
            return self._failure
        if self._job is None:
            return Progress("idle", 0, 0, 0, 0, 0.0, 0.0, 0.0, 0.0, {}, "")
        return self._job.progress()

    def preview(self, maximum_dimension=1024):
        if self._job is None:
            return None, 0
        return self._job.preview(maximum_dimension)

    def result(self):
        """The full linear image, only for a completed job (render_session.cpp:652-656)."""
        if self._job is None or self.progress().state != "completed":
            return None
        if self._result is None:
            self._result = self._job.image()
        return self._result

    def film(self):
        return self._job.film()


def render(device: Device, scene: Scene, render: RenderSettings, integrator: IntegratorSettings = None,
           camera: Camera = None, pass_samples=0):
    """Synchronous render (PathTracer::render): returns (image [h][w][3] float32, progress)."""
    integrator = integrator or IntegratorSettings()
    bvh = device.upload(scene)
    job = Job(bvh, camera or scene.camera, render, integrator, pass_samples)
    try:
        job.wait(-1)
        p = job.progress()
        if p.state != "completed":
            raise N.MrgError(N.MRG_NONFINITE if "nonfinite" in p.error or "range" in p.error else N.MRG_DEVICE_ERROR,
                             p.error or f"render ended {p.state}")
        return job.image(), p
    finally:
        job.release()
