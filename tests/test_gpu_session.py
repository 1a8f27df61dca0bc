"""RenderSession contract on the GPU backend (tests/session_tests.cpp): bitwise equality with the
serial render across pass sizes and schedules, pause/resume storms, cancellation that publishes only
recorded samples, failure without partial publication, recovery, HBM-budget admission, and each job's
result identical alone and inside a batched mix."""
import time

import numpy as np
import pytest

from minirender_amd import _native as N, api, scenegen
from minirender_amd.scene import Camera, concat

pytestmark = pytest.mark.gpu


def sky():
    return concat([], [], environment=(0.2, 0.4, 0.6))


def wait_until(pred, timeout=20.0):
    end = time.time() + timeout
    while not pred():
        assert time.time() < end, "timed out waiting for a render state"
        time.sleep(0.001)


def test_session_equals_serial_across_pass_sizes(device):
    scene = scenegen.cornell_box()
    rs, ins = api.RenderSettings(17, 13, 9, 413), api.IntegratorSettings("mis", 4, False, 3)
    ref, _ = api.render(device, scene, rs, ins)
    s = api.RenderSession(device)
    for pass_samples in (1, 4, 8):
        print(f"[session] pass_samples={pass_samples}", flush=True)
        s.start(scene, rs, ins, pass_samples=pass_samples)
        assert s.wait_for(20000)
        p = s.progress()
        assert p.state == "completed", p.error
        assert p.recorded_samples == 17 * 13 * 9 and p.completed_pixels == 17 * 13
        assert p.statistics["paths"] == p.recorded_samples
        np.testing.assert_array_equal(s.result(), ref)
        full, rec = s.preview(0)
        assert rec == p.total_samples
        np.testing.assert_array_equal(full, ref)
        small, _ = s.preview(8)
        assert small.shape[0] <= 8 and small.shape[1] <= 8


def test_pause_resume_storm_and_cancel(device):
    s = api.RenderSession(device)
    s.start(sky(), api.RenderSettings(64, 40, 1000000, 31), api.IntegratorSettings(), pass_samples=4)
    wait_until(lambda: s.progress().recorded_samples > 0)
    with pytest.raises(RuntimeError):
        s.start(sky(), api.RenderSettings(8, 8, 1, 1))
    assert s.pause()
    wait_until(lambda: s.progress().state == "paused")
    count = s.progress().recorded_samples
    time.sleep(0.05)
    assert s.progress().recorded_samples == count
    img, rec = s.preview(32)
    assert rec == count
    vals = {tuple(v) for v in img.reshape(-1, 3)}
    assert vals <= {(0.0, 0.0, 0.0), tuple(np.float32([0.2, 0.4, 0.6]))}
    assert any(v[0] > 0 for v in vals)
    for _ in range(12):
        assert s.resume()
        assert s.pause()
        wait_until(lambda: s.progress().state == "paused")
    assert s.resume()
    wait_until(lambda: s.progress().recorded_samples > count)
    s.cancel()
    assert s.wait_for(10000)
    p = s.progress()
    assert p.state == "cancelled"
    assert count <= p.recorded_samples < p.total_samples
    assert s.result() is None
    sums, counts = s.film()
    assert int(counts.sum()) == p.recorded_samples


def test_cancel_restart_cycles_and_job_ids(device):
    s = api.RenderSession(device)
    first = None
    for _ in range(8):
        s.start(sky(), api.RenderSettings(64, 64, 1000000, 1))
        first = first or s.progress().job_id
        s.cancel()
        assert s.wait_for(10000)
        assert s.progress().state == "cancelled"
    s.start(sky(), api.RenderSettings(8, 6, 5, 11))
    assert s.wait_for(10000)
    p = s.progress()
    assert p.state == "completed" and p.recorded_samples == 240 and p.job_id == first + 8


def test_overflow_fails_without_publishing(device):
    """session_tests.cpp:203-242: a sample above FLT_MAX fails the job and publishes nothing."""
    from minirender_amd.scene import Camera, material
    q = np.array([[(-1, 0, -1), (-1, 0, 1), (1, 0, 1)], [(-1, 0, -1), (1, 0, 1), (1, 0, -1)]], np.float64)
    scene = concat([(q, np.tile([0, 1.0, 0], (2, 3, 1)))], [material("dielectric", (1, 1, 1))],
                   environment=(2e38, 2e38, 2e38),
                   camera=Camera(position=(0, -1, 0), target=(0, 0, 0), up=(0, 0, 1)))
    s = api.RenderSession(device)
    s.start(scene, api.RenderSettings(4, 4, 64, 0), api.IntegratorSettings("mis", 2, False, 3))
    assert s.wait_for(10000)
    p = s.progress()
    assert p.state == "failed" and p.error and p.recorded_samples == 0
    img, _ = s.preview(0)
    assert img is None or (img == 0).all()
    # recovery
    s.start(sky(), api.RenderSettings(8, 8, 1, 1))
    assert s.wait_for(10000) and s.progress().state == "completed"


def _overflow_scene():
    """A film overflow that only the film check catches (the scene itself uploads): a single-sided
    area light of radiance 1.8e38 (power 2.8e38, finite in float) seen from below through a glass
    sheet; leaving the glass scales radiance by eta^2 = 2.25, past FLT_MAX (session_tests.cpp:203-242
    builds the same overflow from a 2e38 environment, whose selection power 4 pi 2e38 float cannot
    hold)."""
    from minirender_amd.scene import Camera, material
    g = np.array([[(-1, 0, -1), (-1, 0, 1), (1, 0, 1)], [(-1, 0, -1), (1, 0, 1), (1, 0, -1)]], np.float64)
    a, y = 0.35, 0.3
    lamp = np.array([[(-a, y, -a), (a, y, -a), (a, y, a)], [(-a, y, -a), (a, y, a), (-a, y, a)]], np.float64)  # faces -y
    return concat([(g, np.tile([0, 1.0, 0], (2, 3, 1))), (lamp, np.tile([0, -1.0, 0], (2, 3, 1)))],
                  [material("dielectric", (1, 1, 1)), material("diffuse", (0, 0, 0), emission=(1.8e38, 1.8e38, 1.8e38),
                                                            double_sided=False)],
                  camera=Camera(position=(0, -1, 0), target=(0, 0, 0), up=(0, 0, 1)))


def test_failing_job_in_a_batch_leaves_the_other_films():
    """A job that fails the film check inside a co-scheduled batch publishes nothing, and the job beside
    it commits the film it gets alone (the commit stages each pixel's sums per batch pixel)."""
    good_scene, bad_scene = scenegen.cornell_box(), _overflow_scene()
    rs, ins = api.RenderSettings(40, 24, 16, 5), api.IntegratorSettings("mis", 8)
    films = {}
    for mode in ("plain", "batched"):
        print(f"[commit] {mode}", flush=True)
        with api.Device(0, schedule=mode, path_capacity=1 << 16) as d:
            good = api.Job(d.upload(good_scene), good_scene.camera, rs, ins, pass_samples=4)
            if mode == "batched":
                bads = [api.Job(d.upload(bad_scene), bad_scene.camera, api.RenderSettings(4, 4, 64, k),
                                api.IntegratorSettings("mis", 2, False, 3), pass_samples=16) for k in range(3)]
                for b in bads:
                    b.wait()
                    p = b.progress()
                    assert p.state == "failed" and p.error and p.recorded_samples == 0
                    sums, counts = b.film()
                    assert (counts == 0).all() and (sums == 0).all()
                    b.release()
            good.wait()
            assert good.progress().state == "completed"
            films[mode] = good.film()
            good.release()
    np.testing.assert_array_equal(films["plain"][1], films["batched"][1])
    np.testing.assert_array_equal(films["plain"][0], films["batched"][0])


def test_collapsed_scene_fails_then_recovers(device):
    scene = scenegen.floor_scene()
    scene.positions[:] = np.float32(1e10) + scene.positions * np.float32(1e-4)
    s = api.RenderSession(device)
    s.start(scene, api.RenderSettings(8, 8, 1, 1))
    assert s.wait_for(1000)
    p = s.progress()
    assert p.state == "failed" and p.error
    assert s.result() is None
    s.start(sky(), api.RenderSettings(8, 8, 1, 1))
    assert s.wait_for(10000) and s.progress().state == "completed" and not s.progress().error


def test_hbm_budget_admission():
    """A job that does not fit the budget is refused; two that do not fit together run one after the other."""
    scene = scenegen.cornell_box()
    film = 256 * 256 * 28
    with api.Device(0, path_capacity=1 << 18) as probe:
        b0 = probe.upload(scene)
        m = probe.memory()
        b0.release()
    budget = m["pool_bytes"] + m["resident_bytes"] + 3 * film // 2
    with api.Device(0, schedule="batched", path_capacity=1 << 18, hbm_budget_bytes=budget) as d:
        bvh = d.upload(scene)
        with pytest.raises(N.NotAdmitted):
            api.Job(bvh, scene.camera, api.RenderSettings(1024, 1024, 1, 0), api.IntegratorSettings())
        a = api.Job(bvh, scene.camera, api.RenderSettings(256, 256, 64, 0), api.IntegratorSettings())
        b = api.Job(bvh, scene.camera, api.RenderSettings(256, 256, 64, 1), api.IntegratorSettings())
        a.wait()
        b.wait()
        pa, pb = a.progress(), b.progress()
        assert pa.state == pb.state == "completed"
        assert pb.queued_seconds >= pa.queued_seconds
        a.release()
        b.release()


def test_results_identical_alone_and_in_the_mix():
    """Each job's film is bit-identical when rendered alone (plain) and co-scheduled (batched). The
    mix includes scenes with textures, an environment map and energy-compensated conductors, so the
    plain scenes' paths are shaded by the extended kernel in the batches and by the plain one alone."""
    scenes = [scenegen.cornell_box(), scenegen.mixed_terrain(cells=48, spheres=8, sphere_level=2),
              scenegen.icosphere(level=5), scenegen.furnace_spheres(sphere_level=2), scenegen.textured_scene(sphere_level=2),
              scenegen.envmap_scene(sphere_level=2)]
    specs = [(0, 96, 64, 24, 8), (1, 128, 72, 16, 16), (2, 64, 64, 40, 10), (1, 33, 17, 7, 12), (0, 200, 150, 5, 8),
             (3, 48, 40, 12, 6), (4, 56, 40, 9, 7), (5, 50, 30, 11, 6)]
    films = {}
    # the third run traces the co-scheduled batches with the refill driver, whose lanes then hold
    # rays of different scenes side by side
    for mode, traversal in (("plain", "auto"), ("batched", "auto"), ("batched", "refill")):
        print(f"[mix] {mode} traversal={traversal}", flush=True)
        with api.Device(0, schedule=mode, path_capacity=1 << 17, use_graph=(mode == "batched"), traversal=traversal) as d:
            bvhs = [d.upload(s) for s in scenes]
            jobs = [api.Job(bvhs[k], scenes[k].camera, api.RenderSettings(w, h, n, 7 + i), api.IntegratorSettings("mis", dep),
                            pass_samples=4) for i, (k, w, h, n, dep) in enumerate(specs)]
            for j in jobs:
                j.wait()
                assert j.progress().state == "completed"
            films[mode, traversal] = [j.film() for j in jobs]
            for j in jobs:
                j.release()
    for other in (("batched", "auto"), ("batched", "refill")):
        for (s0, c0), (s1, c1) in zip(films["plain", "auto"], films[other]):
            np.testing.assert_array_equal(c0, c1)
            np.testing.assert_array_equal(s0, s1)


def test_close_with_a_paused_job_outstanding():
    """Closing a device cancels its unfinished jobs and wakes their waiters; the caller's job and scene
    handles stay valid: progress and the recorded film are still readable, release works, and ray
    queries on the closed device fail cleanly."""
    d = api.Device(0, path_capacity=1 << 18)
    bvh = d.upload(sky())
    job = api.Job(bvh, Camera(), api.RenderSettings(64, 40, 1000000, 5), api.IntegratorSettings(), pass_samples=4)
    wait_until(lambda: job.progress().recorded_samples > 0)
    assert job.pause()
    wait_until(lambda: job.progress().state == "paused")
    recorded = job.progress().recorded_samples
    print(f"[close] paused at {recorded} samples; closing the device", flush=True)
    d.close()
    assert job.wait(1000)
    p = job.progress()
    assert p.state == "cancelled" and p.recorded_samples == recorded
    sums, counts = job.film()
    assert int(counts.sum()) == recorded
    with pytest.raises(N.InvalidArgument):
        bvh.occluded(np.array([[0, 0, 0, 0, 0, 1]], np.float32))
    job.release()
    bvh.release()


def test_finished_films_leave_the_admission_budget():
    """A budget that holds two films at a time: six jobs submitted together, every handle kept until all
    finish. Admission moves finished films out of HBM when a waiting job needs the room (their results
    stay readable), so the later jobs are admitted as the earlier ones complete instead of waiting on
    handles the caller still holds; the budget is never exceeded."""
    scene = scenegen.cornell_box()
    film = 128 * 128 * 28
    with api.Device(0, path_capacity=1 << 16) as probe:
        b0 = probe.upload(scene)
        m = probe.memory()
        b0.release()
    budget = m["pool_bytes"] + m["resident_bytes"] + 2 * film + film // 2
    with api.Device(0, schedule="batched", path_capacity=1 << 16, hbm_budget_bytes=budget) as d:
        bvh = d.upload(scene)
        jobs = [api.Job(bvh, scene.camera, api.RenderSettings(128, 128, 8, i), api.IntegratorSettings("mis", 4), 4)
                for i in range(6)]
        for j in jobs:
            assert j.wait(20000)
        ps = [j.progress() for j in jobs]
        assert all(p.state == "completed" for p in ps)
        assert max(p.queued_seconds for p in ps) > 0
        assert d.memory()["resident_bytes"] + m["pool_bytes"] <= budget
        ref = api.render(d, scene, api.RenderSettings(128, 128, 8, 3), api.IntegratorSettings("mis", 4))[0]
        np.testing.assert_array_equal(jobs[3].image(), ref)  # a film moved to host memory
        np.testing.assert_array_equal(jobs[0].image(), api.render(d, scene, api.RenderSettings(128, 128, 8, 0),
                                                                  api.IntegratorSettings("mis", 4))[0])
        for j in jobs:
            j.release()
        assert d.memory()["resident_bytes"] == m["resident_bytes"]
        bvh.release()


@pytest.mark.parametrize("schedule", ["plain_fill", "batched"])
def test_two_batches_in_flight_keep_films_and_session_semantics(schedule):
    """Two path pools: the next batch renders while the previous one drains. Films still commit in batch
    order, so every job equals its plain single-pool render bit for bit; pause parks the job once both
    batches have landed, previews read whole batches only, cancel keeps the recorded samples."""
    scenes = [scenegen.cornell_box(), scenegen.mixed_terrain(cells=48, spheres=8, sphere_level=2)]
    specs = [(0, 96, 64, 24, 8), (1, 128, 72, 16, 16), (0, 33, 17, 40, 6), (1, 64, 48, 12, 10)]
    films = {}
    for mode, tracks in (("plain", 1), (schedule, 2)):
        print(f"[two tracks] {mode} batches_in_flight={tracks}", flush=True)
        with api.Device(0, schedule=mode, path_capacity=1 << 15, batches_in_flight=tracks) as d:
            bvhs = [d.upload(s) for s in scenes]
            jobs = [api.Job(bvhs[k], scenes[k].camera, api.RenderSettings(w, h, n, 3 + i), api.IntegratorSettings("mis", dep),
                            pass_samples=2) for i, (k, w, h, n, dep) in enumerate(specs)]
            for j in jobs:
                assert j.wait(60000)
                assert j.progress().state == "completed"
            films[mode] = [j.film() for j in jobs]
            for j in jobs:
                j.release()
            for b in bvhs:
                b.release()
    for (s0, c0), (s1, c1) in zip(films["plain"], films[schedule]):
        np.testing.assert_array_equal(c0, c1)
        np.testing.assert_array_equal(s0, s1)
    with api.Device(0, schedule=schedule, path_capacity=1 << 14, batches_in_flight=2) as d:
        s = api.RenderSession(d)
        s.start(sky(), api.RenderSettings(64, 40, 1000000, 31), api.IntegratorSettings(), pass_samples=1)
        wait_until(lambda: s.progress().recorded_samples > 0)
        for _ in range(6):
            img, rec = s.preview(0)
            sums, counts = s.film()
            assert int(counts.sum()) % (64 * 40) == 0  # whole passes only
        assert s.pause()
        wait_until(lambda: s.progress().state == "paused")
        count = s.progress().recorded_samples
        time.sleep(0.05)
        assert s.progress().recorded_samples == count
        sums, counts = s.film()
        assert int(counts.sum()) == count
        assert s.resume()
        wait_until(lambda: s.progress().recorded_samples > count)
        s.cancel()
        assert s.wait_for(10000)
        p = s.progress()
        assert p.state == "cancelled"
        sums, counts = s.film()
        assert int(counts.sum()) == p.recorded_samples


def test_cancel_lands_within_a_bounce_not_a_batch():
    """Cancellation reaches a batch in flight: its shading ends the job's paths at the next bounce and its
    film commit skips the job, so cancel-to-terminal takes a fraction of a batch (the reference checks
    cancellation per BVH node, bvh.cpp:237, and per bounce, path_tracer.cpp:165). The film keeps exactly
    the recorded samples."""
    scene = scenegen.mixed_terrain(cells=256, spheres=16, sphere_level=2)
    with api.Device(0, schedule="plain_fill", path_capacity=1 << 27) as d:
        bvh = d.upload(scene)
        job = api.Job(bvh, scene.camera, api.RenderSettings(2048, 2048, 4096, 1), api.IntegratorSettings("mis", 16), 32)
        wait_until(lambda: job.progress().recorded_samples > 0, timeout=60)
        t0, r0 = time.perf_counter(), job.progress().recorded_samples
        wait_until(lambda: job.progress().recorded_samples > r0, timeout=60)
        batch = time.perf_counter() - t0  # one batch of 2^27 paths
        r1 = job.progress().recorded_samples
        time.sleep(batch * 0.3)
        t1 = time.perf_counter()
        job.cancel()
        assert job.wait(60000)
        latency = time.perf_counter() - t1
        p = job.progress()
        print(f"[cancel] batch {batch * 1e3:.1f} ms, cancel to terminal {latency * 1e3:.1f} ms, "
              f"recorded {r1} -> {p.recorded_samples}", flush=True)
        assert p.state == "cancelled"
        assert latency < 0.5 * batch
        sums, counts = job.film()
        assert int(counts.sum()) == p.recorded_samples
        job.release()
        bvh.release()
