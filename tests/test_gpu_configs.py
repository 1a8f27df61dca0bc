"""GPU parity at the benchmark configurations and for the FP32 shadow-segment hazard.

* The reference's emitter case (tests/path_tests.cpp:103-120): the segments the FP32 twin traces,
  run through mrg_occluded, give the twin's false-occlusion count; the render path's shadow rays
  (which skip their emitter triangle) match the twin pixel for pixel on a receiver above the emitter.
* The 1080p / 256-spp mixed scene (2,179,072 triangles) in its benchmark batch shape (default device
  options: automatic pool, refill traversal, one-scene kernel) against the twin on a band of rows, and
  the same scene at 5 spp over the whole image with the statistics counters equal.
* 64 uneven jobs in shared batches: every film equals its plain (one job at a time) film bit for bit,
  and the small-scene jobs match the twin.
* A served request stream with cancels and HBM-budget admission: resident scenes deduplicated, every
  completed film equal to the same job rendered alone.
"""
import math

import numpy as np
import pytest

from minirender_amd import api, scenegen, serve
from minirender_amd.scene import Camera, concat, material

pytestmark = pytest.mark.gpu

REL = 1e-4
TRI_N = np.array([[[0, 0, 1]] * 3], np.float64)


def _check(img, twin, label):
    diff = np.abs(img.astype(np.float64) - twin.astype(np.float64))
    bad = diff > REL * np.maximum(np.abs(twin.astype(np.float64)), 1e-6)
    print(f"  {label}: pixels={img.shape[0] * img.shape[1]} exact={np.mean(img == twin):.6f} "
          f"over_1e-4={bad.any(axis=-1).sum()}", flush=True)
    assert not bad.any()


# ---------------------------------------------------------------------------------------------
# FP32 shadow segments (path_tests.cpp:103-120)


def _emitter():
    tri = np.array([[[-1, -1, 0], [1, -1, 0], [-1, 1, 0]]], np.float64)
    return tri, material(emission=(1, 1, 1))


def _samples(count=1000):
    from minirender_amd.rng import Streams
    s = Streams(np.uint64(171), np.uint64(1))
    return [(float(s.next_1d()), float(s.next_1d())) for _ in range(count)]


def test_emitter_segments_through_mrg_occluded_match_the_twin(oracle):
    """The FP32 twin's shadow segments of the reference's emitter case, traced by mrg_occluded with the
    render kernels' FP32 hit decisions (query_precision 32; the plain any-hit query, which skips nothing):
    the GPU reports the same false endpoint occlusions as the twin's any-hit query, segment for segment
    (37 of 1000: the reason the render path's shadow rays skip their emitter triangle)."""
    tri, mat = _emitter()
    scene = concat([(tri, TRI_N)], [mat], environment=(0, 0, 0))
    o = oracle.OracleScene(scene)
    device = api.Device(0, query_precision=32, path_capacity=1 << 16)  # queries only: no render pool
    bvh = device.upload(scene)
    origin = np.array([0.5, 0.25, 4.0])
    gpu, twin = [], []
    print("[shadow] emitter case, 1000 FP32 segments", flush=True)
    for u in _samples():
        seg = o.shadow_segment(origin, 0.5, u, precision=32)
        ray = seg["ray"][None].astype(np.float32)
        gpu.append(bool(bvh.occluded(ray, 0.0, seg["limit"])[0]))
        twin.append(bool(o.occluded(ray.astype(np.float64), 0.0, seg["limit"], precision=32)[0][0]))
        assert not seg["occluded"]  # the render path's segment (emitter skipped)
    bvh.release()
    device.close()
    print(f"  false occlusions without the skip: gpu={sum(gpu)} twin={sum(twin)}", flush=True)
    assert gpu == twin and sum(twin) == 37


def test_receiver_above_emitter_matches_the_twin(device, oracle):
    """A diffuse receiver at z = 4 lit only through light sampling by the emitter on z = 0 (the geometry of
    the reference's emitter case): every pixel equals the twin's, whose segments have no false occlusions."""
    tri, lamp = _emitter()
    quad = np.array([[(0.3, 0.05, 4), (0.7, 0.45, 4), (0.7, 0.05, 4)], [(0.3, 0.05, 4), (0.3, 0.45, 4), (0.7, 0.45, 4)]],
                    np.float64)
    qn = np.tile(np.array([0, 0, -1.0]), (2, 3, 1))
    cam = Camera(position=(0.5, 0.25, 3.0), target=(0.5, 0.25, 4.0), up=(0, 1, 0), vertical_fov_degrees=10.0)
    scene = concat([(tri, TRI_N), (quad, qn)], [lamp, material("diffuse", (0.8, 0.8, 0.8))], environment=(0, 0, 0),
                   camera=cam)
    rs, ins = api.RenderSettings(16, 16, 16, 3), api.IntegratorSettings("light", 1, False, 3)
    print("[shadow] receiver above the emitter 16x16x16 light sampling", flush=True)
    img, prog = api.render(device, scene, rs, ins)
    twin, _, st = oracle.OracleScene(scene).render(cam, 16, 16, 16, seed=3, mode="light", max_depth=1, rr=False,
                                                   precision=32, sums=True)
    _check(img, twin, "receiver")
    assert prog.statistics["shadow_rays"] == st["shadow_rays"] == 16 * 16 * 16
    assert (twin > 0).all()


# ---------------------------------------------------------------------------------------------
# The 1080p / 256-spp mixed scene


@pytest.fixture(scope="module")
def headline():
    from oracle import oracle as O
    scene = scenegen.mixed_terrain()
    assert scene.triangle_count == 2_179_072
    return scene, O.OracleScene(scene)


def test_headline_scene_in_its_benchmark_shape_matches_the_twin(headline):
    """1920 x 1080 x 256 spp, depth 16, default device options (the benchmark's job): rows 516-519 against
    the FP32 twin within 1e-4; the kernel times show the refill traversal in the automatic pool (the whole
    job in one 2^29-path batch on a 288 GB device)."""
    scene, o = headline
    w, h, spp, depth, rows = 1920, 1080, 256, 16, (516, 520)
    print(f"[config 3] {w}x{h} spp={spp} depth={depth}, default options, rows {rows}", flush=True)
    with api.Device(0) as dev:
        dev.kernel_times(reset=True)
        img, prog = api.render(dev, scene, api.RenderSettings(w, h, spp, 0), api.IntegratorSettings("mis", depth, True, 3))
        kt = dev.kernel_times()
    assert prog.recorded_samples == w * h * spp
    print(f"  batches={kt['batches']} traversal launches={kt['extend_launches']} refill={kt['refill_launches']}", flush=True)
    assert kt["refill_launches"] == kt["extend_launches"] > 0
    twin, _ = o.render(scene.camera, w, h, spp, seed=0, mode="mis", max_depth=depth, rr=True, precision=32,
                       crop=(0, rows[0], w, rows[1]))
    _check(img[rows[0]:rows[1]], twin, "config 3 rows")


def test_headline_scene_counters_equal_the_twin(headline):
    """The same scene over the whole image at 5 spp (10.4M paths, so the default options still pick the
    refill traversal): every pixel within 1e-4 of the twin and every statistics counter equal."""
    scene, o = headline
    w, h, spp, depth = 1920, 1080, 5, 16
    print(f"[config 3] {w}x{h} spp={spp} depth={depth}, full image, counters", flush=True)
    with api.Device(0) as dev:
        dev.kernel_times(reset=True)
        img, prog = api.render(dev, scene, api.RenderSettings(w, h, spp, 7), api.IntegratorSettings("mis", depth, True, 3))
        kt = dev.kernel_times()
    assert kt["refill_launches"] == kt["extend_launches"] > 0
    twin, _, st = o.render(scene.camera, w, h, spp, seed=7, mode="mis", max_depth=depth, rr=True, precision=32, sums=True)
    _check(img, twin, "config 3 full image")
    for k in ("paths", "extension_rays", "shadow_rays", "scattering_events", "roulette_terminations"):
        print(f"  {k}: gpu={prog.statistics[k]} twin={st[k]}", flush=True)
        assert prog.statistics[k] == st[k], k


# ---------------------------------------------------------------------------------------------
# 64 uneven jobs in shared batches


def test_sixty_four_uneven_jobs_batched_equal_plain_and_the_twin(oracle):
    """The batch workload's 64 jobs (sides 128-1024, 8 scenes of 36 to ~410k triangles) at 1/16 of their
    samples: rendered together on the batched schedule with automatic traversal, every job's FP64 film
    equals the one the plain schedule renders for it alone, bit for bit; the jobs of the three smallest
    scenes match the FP32 twin."""
    scenes = serve.scene_pool(seed=4)
    jobs = serve.batch_jobs(seed=4, spp_divisor=16)
    print(f"[config 4] {len(jobs)} jobs, {sum(r.width * r.height * r.spp for r in jobs)} samples", flush=True)

    def render_all(schedule):
        films = {}
        with api.Device(0, schedule=schedule, traversal="auto") as dev:
            bvhs = [dev.upload(s) for s in scenes]
            if schedule == "batched":
                handles = [api.Job(bvhs[r.scene], scenes[r.scene].camera, api.RenderSettings(r.width, r.height, r.spp, r.seed),
                                   api.IntegratorSettings("mis", r.max_depth, True, 3), 16) for r in jobs]
                for r, j in zip(jobs, handles):
                    j.wait(-1)
                    assert j.progress().state == "completed"
                    films[r.index] = j.film()
                    j.release()
            else:
                for r in jobs:
                    j = api.Job(bvhs[r.scene], scenes[r.scene].camera, api.RenderSettings(r.width, r.height, r.spp, r.seed),
                                api.IntegratorSettings("mis", r.max_depth, True, 3), 16)
                    j.wait(-1)
                    assert j.progress().state == "completed"
                    films[r.index] = j.film()
                    j.release()
            kt = dev.kernel_times()
            for b in bvhs:
                b.release()
        print(f"  {schedule}: {kt['batches']} batches", flush=True)
        return films, kt

    batched, kb = render_all("batched")
    plain, kp = render_all("plain")
    assert kb["batches"] < kp["batches"]
    for r in jobs:
        np.testing.assert_array_equal(batched[r.index][1], plain[r.index][1])
        np.testing.assert_array_equal(batched[r.index][0], plain[r.index][0])
    small = sorted((r for r in jobs if r.scene in (0, 1, 2)), key=lambda r: r.width * r.height * r.spp)[:3]
    for r in small:
        sc = scenes[r.scene]
        twin, _ = oracle.OracleScene(sc).render(sc.camera, r.width, r.height, r.spp, seed=r.seed, mode="mis",
                                                max_depth=r.max_depth, rr=True, precision=32)
        sums, counts = batched[r.index]
        _check((sums / counts[..., None]).astype(np.float32), twin, f"job {r.index} ({sc.name} {r.width}^2 x {r.spp})")


# ---------------------------------------------------------------------------------------------
# Served request stream


def test_served_stream_admits_dedups_cancels_and_matches_alone(oracle):
    """serve.run_stream on a batched device: 16 requests at once over four scenes, two of them the same
    content (one resident copy, reference-counted), an HBM budget that holds the films of about three jobs
    (later jobs wait for admission), and cancels after 1-3 passes. Every completed job's film equals the
    same job rendered alone on the plain schedule, bit for bit; cancelled jobs keep whole recorded passes."""
    scenes = [scenegen.cornell_box(), scenegen.mixed_terrain(cells=24, spheres=4, sphere_level=1, seed=4),
              scenegen.icosphere(level=3, seed=4), scenegen.cornell_box()]
    reqs = serve.request_stream(count=16, rate=math.inf, seed=7, cancel_fraction=0.3, scenes=4, side=(48, 128),
                                spp=(16, 96), depth=(4, 8))
    assert any(r.cancel_after_passes > 0 for r in reqs)
    film_bytes = max(r.width * r.height * 28 for r in reqs)
    with api.Device(0, schedule="batched", path_capacity=1 << 20, hbm_budget_bytes=1 << 40) as probe:
        pool = probe.memory()["pool_bytes"]
        bvhs = [probe.upload(s) for s in scenes]
        scene_bytes = probe.memory()["resident_bytes"]
        for b in bvhs:
            b.release()
    budget = pool + scene_bytes + 3 * film_bytes
    print(f"[config 5] 16 requests, budget = pool {pool} + scenes {scene_bytes} + 3 films of {film_bytes}", flush=True)
    with api.Device(0, schedule="batched", path_capacity=1 << 20, hbm_budget_bytes=budget, use_graph=True) as dev:
        bvhs = [dev.upload(s) for s in scenes]
        assert bvhs[0].info()["content_hash"] == bvhs[3].info()["content_hash"]
        assert bvhs[0].info()["reference_count"] == 2
        assert dev.memory()["resident_bytes"] == scene_bytes
        summary, records, films = serve.run_stream(dev, bvhs, scenes, reqs, keep_films=True)
        assert dev.memory()["resident_bytes"] == scene_bytes  # every finished film left HBM
        for b in bvhs:
            b.release()
    print(f"  completed={summary['completed']} cancelled={summary['cancelled']} "
          f"queued_max={max(r['queued'] for r in records):.4f}s", flush=True)
    assert summary["failed"] == 0 and summary["completed"] + summary["cancelled"] == 16
    assert summary["cancelled"] >= 1
    assert max(r["queued"] for r in records) > 0  # admission made jobs wait
    by_index = {r.index: r for r in reqs}
    for rec in records:
        req = by_index[rec["index"]]
        if rec["state"] == "cancelled":
            assert rec["samples"] % (req.width * req.height) == 0  # whole passes only
            assert rec["samples"] >= min(req.cancel_after_passes * 16, req.spp) * req.width * req.height
        else:
            assert rec["first_pass"] <= rec["completion"]
    with api.Device(0, schedule="plain") as alone:
        resident = [alone.upload(sc) for sc in scenes]
        for index, (sums, counts) in films.items():
            req = by_index[index]
            j = api.Job(resident[req.scene], scenes[req.scene].camera, api.RenderSettings(req.width, req.height, req.spp, req.seed),
                        api.IntegratorSettings("mis", req.max_depth, True, 3), 16)
            j.wait(-1)
            s2, c2 = j.film()
            j.release()
            np.testing.assert_array_equal(c2, counts)
            np.testing.assert_array_equal(s2, sums)
        for b in resident:
            b.release()
    # one completed job also against the twin
    first = min(films)
    req = by_index[first]
    sc = scenes[req.scene]
    twin, _ = oracle.OracleScene(sc).render(sc.camera, req.width, req.height, req.spp, seed=req.seed, mode="mis",
                                            max_depth=req.max_depth, rr=True, precision=32)
    sums, counts = films[first]
    _check((sums / counts[..., None]).astype(np.float32), twin, f"served job {first}")
