"""Autotuner for libmrg's device options.

The knobs are the launch and queue choices of the wavefront integrator:

* ``path_capacity``             paths in flight per batch (the extension / shadow queue capacity),
* ``persistent_blocks_per_cu``  resident 256-thread blocks per CU of the persistent traversal grids
                                (0: the library's choice, 4 for fused / split and 7 for refill),
* ``traversal``                 ``split`` (closest-hit and any-hit kernels per bounce), ``fused``
                                (one traversal launch per bounce for both kinds of ray), ``refill``
                                (the fused launch where lanes that finish take the next ray at once)
                                or ``auto`` (refill for batches of >= 2^23 paths holding a
                                scene of >= 2^14 triangles, fused otherwise),
* ``use_graph``                 the bounce loop captured once in a hipGraph and replayed,
* ``max_jobs_in_flight``        (batched schedule) jobs co-scheduled in one batch,
* ``batches_in_flight``         1, or 2 (a second path pool renders the next batch while the previous
                                one drains its tail of late bounces).

A calibration workload (a list of scenes and job requests, all submitted at once) is rendered with
each candidate; the score is samples per second over the whole workload, best of ``repeats`` runs
after one warm-up. A candidate the device refuses (a pool that leaves no room for any job in the
HBM budget, such as two 2^28-path pools in 64 GiB) scores zero and is recorded as infeasible.
Coordinate descent: starting from the library defaults, each knob in turn is
swept with the others fixed and the best value kept; ``rounds`` passes over all knobs. Results are
plain JSON (``save`` / ``load``), so a C++ caller can read the same file into ``mrg_device_options``.
"""
from __future__ import annotations

import json
import time

from . import _native as N
from . import api

KNOBS = {
    "path_capacity": [1 << 22, 1 << 23, 1 << 24, 1 << 25, 1 << 26, 1 << 27, 1 << 28],
    "persistent_blocks_per_cu": [0, 2, 3, 4, 5, 6],
    "traversal": ["split", "fused", "refill", "auto"],
    "use_graph": [False, True],
    "max_jobs_in_flight": [8, 16, 32, 64],
    "batches_in_flight": [1, 2],
}

DEFAULTS = {"path_capacity": 1 << 26, "persistent_blocks_per_cu": 0, "traversal": "auto", "use_graph": False,
            "max_jobs_in_flight": 64, "batches_in_flight": 1}


def run_workload(options, schedule, scenes, requests, repeats=2, pass_samples=16):
    """Renders every request (all submitted together) on a fresh device with `options`; returns
    (best samples/s, per-run seconds). Scene uploads happen before the timed region."""
    dev = api.Device(0, schedule=schedule, **options)
    try:
        bvhs = [dev.upload(s) for s in scenes]
        runs = []
        for rep in range(repeats + 1):
            t = time.perf_counter()
            jobs = [api.Job(bvhs[r.scene], scenes[r.scene].camera, api.RenderSettings(r.width, r.height, r.spp, r.seed),
                            api.IntegratorSettings("mis", r.max_depth, True, 3), pass_samples) for r in requests]
            samples = 0
            for j in jobs:
                j.wait(-1)
                p = j.progress()
                if p.state != "completed":
                    raise RuntimeError(f"calibration job ended {p.state}: {p.error}")
                samples += p.recorded_samples
                j.release()
            dt = time.perf_counter() - t
            if rep > 0:  # the first run warms up (graph capture, page faults of the pool)
                runs.append((samples / dt, dt))
        for b in bvhs:
            b.release()
    finally:
        dev.close()
    best = max(runs)
    return best[0], [r[1] for r in runs]


def tune(scenes, requests, schedule="batched", knobs=None, rounds=1, repeats=2, log=None, scorer=None):
    """Coordinate descent over `knobs` (name -> candidate values; default KNOBS, without
    max_jobs_in_flight for the plain schedules). Returns the best options and every measurement.
    `scorer(options) -> (samples/s, per-run seconds)` replaces the device measurement (run_workload on
    `scenes` and `requests`) when given."""
    knobs = dict(knobs or KNOBS)
    if schedule != "batched":
        knobs.pop("max_jobs_in_flight", None)
    best = {k: DEFAULTS[k] for k in knobs}
    cache = {}
    trials = []
    measure = scorer or (lambda opts: run_workload(opts, schedule, scenes, requests, repeats))

    def score(opts):
        key = json.dumps(opts, sort_keys=True)
        if key not in cache:
            try:
                value, secs = measure(opts)
                trial = {"options": dict(opts), "samples_per_s": value, "seconds": secs}
                note = f"{value / 1e6:.1f} Msamples/s"
            except N.NotAdmitted as e:
                value = 0.0
                trial = {"options": dict(opts), "samples_per_s": 0.0, "infeasible": str(e)}
                note = f"infeasible ({e})"
            cache[key] = value
            trials.append(trial)
            if log:
                log(f"[autotune] {schedule} {opts} -> {note}")
        return cache[key]

    best_value = score(best)
    for _ in range(rounds):
        for name, values in knobs.items():
            for v in values:
                cand = dict(best, **{name: v})
                s = score(cand)
                if s > best_value:
                    best, best_value = cand, s
    return {"schedule": schedule, "options": best, "samples_per_s": best_value, "trials": trials}


def save(result, path):
    with open(path, "w") as f:
        json.dump(result, f, indent=1)


def load(path):
    """Options of a saved result, ready for ``api.Device(**options)``."""
    with open(path) as f:
        return dict(json.load(f)["options"])
