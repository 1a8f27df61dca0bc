"""GPU closest-hit / any-hit parity against the oracle.

Ray queries decide hits in FP64 by default (query_precision 64: the reference intersector's
arithmetic, RayQuery in double, intersection.cpp:108-152): their ids, distances and barycentrics equal
the FP64 oracle's, rounded to float, on every ray, near-ties included, and any-hit equals the FP64
any-hit.

With query_precision 32 the queries run the render kernels' FP32 watertight test: ids, t and
barycentrics equal the FP32 twin's bit for bit; against FP64 the id rule then has measured exceptions
outside near-ties: rays that cross the edge shared by two triangles, where the FP32 test puts the
crossing on the other side of the edge than FP64 does. They are counted as mismatches, every one is
checked to be such a crossing, and their rate is bounded by SHARED_EDGE_RATE (41 of the 16.7M
microbench rays = 2.4e-6 measured)."""
import numpy as np
import pytest

from minirender_amd import api, scenegen

pytestmark = pytest.mark.gpu

NEAR_TIE = 1e-6   # |t2 - t1| <= 1e-6 * t1 counts as a near-tie
T_REL = 1e-5
SHARED_EDGE_RATE = 4e-6  # bound on id mismatches outside near-ties (all shared-edge crossings)


@pytest.fixture(scope="module")
def device32():
    d = api.Device(0, query_precision=32, path_capacity=1 << 16)  # queries only: no render pool
    yield d
    d.close()


def shared_edge_crossing(scene, prim_a, t_a, prim_b, t_b):
    """Mismatches where the two precisions hit adjacent triangles (sharing a vertex) at t within
    NEAR_TIE relative: the ray crosses the common edge, and each precision assigns it to one side."""
    out = np.zeros(len(prim_a), bool)
    for i in np.where((prim_a != prim_b) & (prim_a >= 0) & (prim_b >= 0))[0]:
        if abs(float(t_a[i]) - float(t_b[i])) > NEAR_TIE * abs(float(t_b[i])):
            continue
        va = {tuple(v) for v in scene.positions[prim_a[i]].tolist()}
        vb = {tuple(v) for v in scene.positions[prim_b[i]].tolist()}
        out[i] = bool(va & vb)
    return out


def _interval(tmin, tmax):
    """The query interval as the GPU receives it (float), widened for the oracle."""
    return float(np.float32(tmin)), float(np.float32(tmax))


def _compare_exact(device, oracle, scene, rays, tmin=0.0, tmax=np.inf):
    """FP64 hit decisions: every id, t and barycentric equals the FP64 oracle's rounded to float."""
    print(f"[intersect parity fp64] scene={scene.name} tris={scene.triangle_count} rays={len(rays)}", flush=True)
    lo, hi = _interval(tmin, tmax)
    bvh = device.upload(scene)
    gpu = bvh.intersect(rays, tmin, tmax, barycentric=True)
    o = oracle.OracleScene(scene)
    r64 = rays.astype(np.float64)
    ref = o.intersect(r64, lo, hi, precision=64, second=True)
    with np.errstate(invalid="ignore"):
        near_tie = np.isfinite(ref["t2"]) & (np.abs(ref["t2"] - ref["t"]) <= NEAR_TIE * np.abs(ref["t"]))
    np.testing.assert_array_equal(gpu["primitive"], ref["prim"])
    hit = ref["prim"] >= 0
    np.testing.assert_array_equal(gpu["t"][hit], ref["t"][hit].astype(np.float32))
    np.testing.assert_array_equal(gpu["barycentric"][hit], ref["bary"][hit].astype(np.float32))
    occ = bvh.occluded(rays, tmin, tmax)
    occ_ref, _ = o.occluded(r64, lo, hi, precision=64)
    np.testing.assert_array_equal(occ, occ_ref)
    print(f"  hits={hit.sum()} near_ties={near_tie.sum()} (all ids equal, near-ties included)", flush=True)
    return gpu


def _compare32(device, oracle, scene, rays, tmin=0.0, tmax=np.inf):
    """FP32 hit decisions: the FP32 twin's results bit for bit; against FP64, ids equal outside
    near-ties and shared-edge crossings."""
    print(f"[intersect parity fp32] scene={scene.name} tris={scene.triangle_count} rays={len(rays)}", flush=True)
    lo, hi = _interval(tmin, tmax)
    bvh = device.upload(scene)
    gpu = bvh.intersect(rays, tmin, tmax, barycentric=True)
    o = oracle.OracleScene(scene)
    r64 = rays.astype(np.float64)
    twin = o.intersect(r64, lo, hi, precision=32)
    ref = o.intersect(r64, lo, hi, precision=64, second=True)
    # FP32 twin: identical ids, t and barycentrics
    np.testing.assert_array_equal(gpu["primitive"], twin["prim"])
    hit = twin["prim"] >= 0
    np.testing.assert_array_equal(gpu["t"][hit], twin["t"][hit].astype(np.float32))
    np.testing.assert_array_equal(gpu["barycentric"][hit], twin["bary"][hit].astype(np.float32))
    # FP64 reference intersector
    t1, t2 = ref["t"], ref["t2"]
    with np.errstate(invalid="ignore"):
        near_tie = np.isfinite(t2) & (np.abs(t2 - t1) <= NEAR_TIE * np.abs(t1))
    mismatch = gpu["primitive"] != ref["prim"]
    outside = mismatch & ~near_tie
    crossing = outside & shared_edge_crossing(scene, gpu["primitive"], gpu["t"], ref["prim"], t1)
    print(f"  hits={hit.sum()} near_ties={near_tie.sum()} id_mismatch={mismatch.sum()} "
          f"outside_near_ties={outside.sum()} (shared-edge crossings {crossing.sum()})")
    assert (outside == crossing).all(), "an id mismatch outside near-ties that is not a shared-edge crossing"
    assert outside.sum() <= max(3, SHARED_EDGE_RATE * len(rays))
    both = (gpu["primitive"] == ref["prim"]) & (ref["prim"] >= 0)
    rel = np.abs(gpu["t"][both] - t1[both]) / np.abs(t1[both])
    assert rel.max(initial=0) <= T_REL
    # any-hit agrees with the twin
    occ = bvh.occluded(rays, tmin, tmax)
    occ_twin, _ = o.occluded(r64, lo, hi, precision=32)
    np.testing.assert_array_equal(occ, occ_twin)
    return gpu


def test_icosphere_random_rays(device, device32, oracle):
    scene = scenegen.icosphere(level=5)
    rays = scenegen.intersection_rays(scene, 1 << 17)
    gpu = _compare_exact(device, oracle, scene, rays)
    assert (gpu["primitive"] >= 0).mean() > 0.3
    _compare32(device32, oracle, scene, rays)


def test_cornell_single_sided_rays(device, device32, oracle):
    scene = scenegen.cornell_box()
    rays = scenegen.intersection_rays(scene, 1 << 15, seed=7, scale=0.9)
    for d, cmp in ((device, _compare_exact), (device32, _compare32)):
        cmp(d, oracle, scene, rays)
        cmp(d, oracle, scene, rays, tmin=1e-3, tmax=1.5)


def test_shared_edges_resolved_in_fp64(device, device32, oracle):
    """Rays aimed at the shared edges of a fan of triangles at three scales (the reference's watertight
    shared-edge cases, intersection_tests.cpp:161-180, 213-237): FP64 decisions give the FP64 ids on
    every ray, including those that pass within float rounding of an edge."""
    from minirender_amd.scene import concat, material
    rng = np.random.default_rng(5)
    for scale, offset in ((1.0, 0.0), (1e-3, 0.0), (1e3, 1e7)):
        k = 16
        ang = np.linspace(0, 2 * np.pi, k + 1)
        ring = np.stack([np.cos(ang), np.sin(ang), np.zeros_like(ang)], -1) * scale + offset
        centre = np.array([0, 0, 0.0]) * scale + offset
        tris = np.stack([np.broadcast_to(centre, (k, 3)), ring[:-1], ring[1:]], 1).astype(np.float32)
        normals = np.tile(np.array([0, 0, 1], np.float32), (k, 3, 1))
        scene = concat([(tris, normals)], [material()])
        # points on the spokes (shared edges) and near them, seen from above at random tilts; the angular
        # offsets are log-uniform over 1e-11..1e-3 so that, at every scale, some rays pass closer to an
        # edge than float rounding resolves (about 2^-19 of the origin distance) and some farther
        a = ang[rng.integers(0, k, 4096)] + rng.choice([-1.0, 1.0], 4096) * 10 ** rng.uniform(-11, -3, 4096)
        r = rng.uniform(0.05, 0.95, 4096) * scale
        target = np.stack([np.cos(a) * r, np.sin(a) * r, np.zeros(4096)], -1) + offset
        d = np.stack([rng.normal(0, 0.3, 4096), rng.normal(0, 0.3, 4096), -np.ones(4096)], -1)
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        o = target - d * 2.0 * max(scale, 1.0)
        rays = np.concatenate([o, d], 1).astype(np.float32)
        rays[:, 3:] /= np.linalg.norm(rays[:, 3:].astype(np.float64), axis=1, keepdims=True).astype(np.float32)
        _compare_exact(device, oracle, scene, rays)


def test_empty_and_single_triangle(device, oracle):
    one = scenegen.floor_scene()
    from minirender_amd.scene import concat, material
    tri = concat([(one.positions[:1], one.normals[:1])], [material()])
    rays = np.array([[0.5, 1, 1.5, 0, -1, 0], [3, 1, -3, 0, -1, 0], [0.5, 1, 1.5, 0, 1, 0]], np.float32)
    gpu = device.upload(tri).intersect(rays)
    o = oracle.OracleScene(tri).intersect(rays.astype(np.float64), precision=64)
    np.testing.assert_array_equal(gpu["primitive"], o["prim"])
    empty = concat([], [])
    e = device.upload(empty).intersect(rays)
    assert (e["primitive"] == -1).all()


def test_coincident_triangles_tie_break(device, device32, oracle):
    """256 coincident triangles: the lowest scene index wins (bvh.cpp:273-276)."""
    from minirender_amd.scene import concat, material
    t = np.array([[[0, 0, 0], [1, 0, 0], [0, 1, 0]]], np.float32)
    n = np.tile(np.array([0, 0, 1], np.float32), (1, 3, 1))
    scene = concat([(t, n)] * 256, [material()] * 256)
    rays = np.array([[0.25, 0.25, 2, 0, 0, -1]], np.float32)
    for d in (device, device32):
        gpu = d.upload(scene).intersect(rays)
        assert gpu["primitive"][0] == 0
        assert gpu["t"][0] == np.float32(2)


def test_full_size_microbench_subset(device, device32, oracle):
    """The 1,310,720-triangle displaced icosphere at full size; every 16th ray of the seeded 16M-ray
    set (1,048,576 rays). FP64 decisions: every id, t and barycentric equals the FP64 reference's.
    FP32 decisions: the twin's; ids match FP64 outside near-ties and shared-edge crossings."""
    scene = scenegen.icosphere(level=8)
    rays = scenegen.intersection_rays(scene, 0, indices=np.arange(0, 1 << 24, 16))
    _compare_exact(device, oracle, scene, rays)
    _compare32(device32, oracle, scene, rays)
