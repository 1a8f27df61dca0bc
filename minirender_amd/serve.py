"""Render-request serving loop over one device: Poisson arrivals of uneven jobs against a set of
resident (deduplicated) scenes, HBM-budget admission and progressive passes inside libmrg, random
cancellation, and per-job latency (time to first pass, time to completion).

The same request stream can be replayed on a device with the plain schedule (one job at a time)
or the batched schedule (up to `max_jobs_in_flight` jobs share every wavefront batch); both see
identical submission times, so throughput and latency compare directly.
"""
from __future__ import annotations

import dataclasses
import math
import threading
import time

import numpy as np

from . import api, scenegen


def scene_pool(seed=4):
    """Eight distinct resident scenes from the three generators, about 1k to 500k triangles."""
    return [
        scenegen.mixed_terrain(cells=24, spheres=4, sphere_level=1, seed=seed),          # ~1.5k
        scenegen.cornell_box(),                                                         # 36
        scenegen.icosphere(level=4, seed=seed),                                         # 5k
        scenegen.mixed_terrain(cells=96, spheres=8, sphere_level=2, seed=seed + 1),     # ~20k
        scenegen.icosphere(level=6, seed=seed + 1),                                     # 82k
        scenegen.mixed_terrain(cells=256, spheres=32, sphere_level=2, seed=seed + 2),   # ~140k
        scenegen.icosphere(level=7, seed=seed + 2),                                     # 328k
        scenegen.mixed_terrain(cells=440, spheres=16, sphere_level=3, seed=seed + 3),   # ~410k
    ]


@dataclasses.dataclass
class Request:
    index: int
    scene: int
    width: int
    height: int
    spp: int
    max_depth: int
    seed: int
    arrival: float          # seconds after the stream starts
    cancel_after_passes: int  # cancelled once this many progressive passes are recorded, or -1


def request_stream(count=64, rate=40.0, seed=5, cancel_fraction=0.05, scenes=8, side=(128, 1024), spp=(16, 1024),
                   depth=(8, 16)):
    """Poisson arrivals at `rate` jobs/s (rate = inf: every request at 0); sizes and sample counts
    log-uniform; a cancelled request is cancelled after 1-3 recorded passes, so every schedule renders
    about the same samples of it. Seeded: the same arguments give the same requests at every rate."""
    r = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(count):
        gap = r.exponential(1.0)
        t += gap / rate if math.isfinite(rate) else 0.0
        s = int(round(math.exp(r.uniform(math.log(side[0]), math.log(side[1])))))
        n = int(round(math.exp(r.uniform(math.log(spp[0]), math.log(spp[1])))))
        cancel = int(r.integers(1, 4)) if r.random() < cancel_fraction else -1
        reqs.append(Request(i, int(r.integers(0, scenes)), s, s, n, int(r.integers(depth[0], depth[1] + 1)),
                            int(r.integers(0, 2**31)), t, cancel))
    return reqs


def batch_jobs(count=64, seed=4, spp_divisor=1, **kw):
    """Uneven jobs submitted together (no arrivals, no cancels): the request distribution of
    request_stream with every arrival at 0; `spp_divisor` shrinks sample counts for calibration."""
    reqs = request_stream(count=count, seed=seed, cancel_fraction=0.0, **kw)
    for r in reqs:
        r.arrival = 0.0
        r.spp = max(2, r.spp // spp_divisor)
    return reqs


def _percentile(v, q):
    return float(np.percentile(np.asarray(v, np.float64), q)) if len(v) else None


TERMINAL = ("cancelled", "completed", "failed", "idle")


def run_stream(device: api.Device, bvhs, scenes, requests, pass_samples=16, poll=0.001, keep_films=False):
    """Submits each request at its arrival time, cancels the selected ones once they have recorded
    their pass count, waits for all of them and returns (summary, per-job records, films)."""
    jobs = []
    watch, lock, done_submitting = {}, threading.Lock(), threading.Event()

    def canceller():
        while True:
            with lock:
                items = list(watch.items())
            if not items and done_submitting.is_set():
                return
            for index, (job, threshold) in items:
                p = job.progress()
                if p.state in TERMINAL or p.recorded_samples >= threshold:
                    if p.state not in TERMINAL:
                        job.cancel()
                    with lock:
                        del watch[index]
            time.sleep(poll)

    watcher = threading.Thread(target=canceller, daemon=True)
    watcher.start()
    start = time.perf_counter()
    try:
        for req in requests:
            delay = req.arrival - (time.perf_counter() - start)
            if delay > 0:
                time.sleep(delay)
            sc = scenes[req.scene]
            job = api.Job(bvhs[req.scene], sc.camera, api.RenderSettings(req.width, req.height, req.spp, req.seed),
                          api.IntegratorSettings("mis", req.max_depth, True, 3), pass_samples)
            jobs.append(job)
            if req.cancel_after_passes >= 0:
                with lock:
                    watch[req.index] = (job, req.cancel_after_passes * pass_samples * req.width * req.height)
    finally:
        done_submitting.set()
    for j in jobs:
        j.wait(-1)
    end = time.perf_counter()
    watcher.join()
    records, films = [], {}
    for req, j in zip(requests, jobs):
        p = j.progress()
        records.append(dict(index=req.index, state=p.state, samples=p.recorded_samples, total=p.total_samples,
                            first_pass=p.first_pass_seconds, completion=p.completion_seconds, queued=p.queued_seconds))
        if keep_films and p.state == "completed":
            films[req.index] = j.film()
        j.release()
    done = [r for r in records if r["state"] == "completed"]
    samples = sum(r["samples"] for r in records)
    makespan = end - start
    summary = dict(
        jobs=len(records), completed=len(done), cancelled=sum(r["state"] == "cancelled" for r in records),
        failed=sum(r["state"] == "failed" for r in records), samples=samples, makespan_s=makespan,
        msamples_per_s=samples / makespan / 1e6,
        first_pass_p50_s=_percentile([r["first_pass"] for r in records if r["first_pass"] > 0], 50),
        first_pass_p99_s=_percentile([r["first_pass"] for r in records if r["first_pass"] > 0], 99),
        completion_p50_s=_percentile([r["completion"] for r in done], 50),
        completion_p99_s=_percentile([r["completion"] for r in done], 99),
    )
    return summary, records, films


SCHEDULES = ("plain", "plain_fill", "batched")


def open_devices(scenes, hbm_budget=64 << 30, max_jobs=64, use_graph=True, path_capacity=1 << 27, batches_in_flight=0,
                 warm=True):
    """One device per schedule with every scene resident, warmed by a short stream (graph capture, first
    touch of the path pool). The serving pool is 2^27 paths (24 GB): a batch is the unit of progress and
    pause latency. Measured on the saturated stream (batched schedule): 2^26 paths 2,900 Msamples/s at a
    first-pass p50 of 0.098 s, 2^27 3,095 at 0.154 s, 2^28 3,136 at 0.294 s."""
    devices = {}
    for mode in SCHEDULES:
        dev = api.Device(0, schedule=mode, hbm_budget_bytes=hbm_budget, max_jobs_in_flight=max_jobs,
                         use_graph=use_graph, path_capacity=path_capacity, batches_in_flight=batches_in_flight)
        bvhs = [dev.upload(s) for s in scenes]
        if warm:
            run_stream(dev, bvhs, scenes, request_stream(count=8, rate=math.inf, seed=99, spp=(16, 64),
                                                         scenes=len(scenes), cancel_fraction=0.0))
        devices[mode] = (dev, bvhs)
    return devices


def close_devices(devices):
    for dev, bvhs in devices.values():
        for b in bvhs:
            b.release()
        dev.close()


def compare(requests=None, scenes=None, devices=None, keep_films=False, **device_options):
    """Replays one request stream on the plain, plain-fill and batched schedules."""
    scenes = scenes or scene_pool()
    requests = requests or request_stream()
    own = devices is None
    devices = devices or open_devices(scenes, **device_options)
    out, films = {}, {}
    try:
        for mode in SCHEDULES:
            dev, bvhs = devices[mode]
            summary, records, f = run_stream(dev, bvhs, scenes, requests, keep_films=keep_films)
            out[mode] = summary
            films[mode] = f
    finally:
        if own:
            close_devices(devices)
    out["speedup"] = out["batched"]["msamples_per_s"] / out["plain"]["msamples_per_s"]
    out["speedup_vs_plain_fill"] = out["batched"]["msamples_per_s"] / out["plain_fill"]["msamples_per_s"]
    # No schedule can finish before the last request arrives: at a fixed arrival rate the largest
    # speedup over plain that this stream can show is plain's makespan over the last arrival time.
    out["last_arrival_s"] = max(r.arrival for r in requests)
    out["speedup_ceiling"] = out["plain"]["makespan_s"] / max(out["last_arrival_s"], 1e-9)
    return out, films


def sweep(rates=(40.0, 80.0, 160.0, 320.0, math.inf), count=64, seed=5, scenes=None, **device_options):
    """Latency against throughput: the same request stream (same jobs and cancels) at each arrival rate,
    on every schedule; rate inf submits every request at once (the saturated stream)."""
    scenes = scenes or scene_pool()
    devices = open_devices(scenes, **device_options)
    rows = []
    try:
        for rate in rates:
            out, _ = compare(request_stream(count=count, rate=rate, seed=seed, scenes=len(scenes)), scenes, devices)
            rows.append(dict(rate=rate, **out))
    finally:
        close_devices(devices)
    return rows
