"""GPU closest-hit / any-hit parity against the oracle (FP32 twin: bit-exact; FP64 reference:
hit ids equal outside near-ties, t within 1e-5 relative).

Against FP64 the id rule has measured exceptions outside near-ties: rays that cross the edge shared by
two triangles, where the FP32 watertight test (which the kernels and the FP32 twin compute identically)
puts the crossing on the other side of the edge than FP64 does. They are counted as mismatches, every
one is checked to be such a crossing, and their rate is bounded by SHARED_EDGE_RATE (41 of the 16.7M
microbench rays = 2.4e-6 measured)."""
import numpy as np
import pytest

from minirender_amd import scenegen

pytestmark = pytest.mark.gpu

NEAR_TIE = 1e-6   # |t2 - t1| <= 1e-6 * t1 counts as a near-tie
T_REL = 1e-5
SHARED_EDGE_RATE = 4e-6  # bound on id mismatches outside near-ties (all shared-edge crossings)


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


def _compare(device, oracle, scene, rays, tmin=0.0, tmax=np.inf):
    print(f"[intersect parity] scene={scene.name} tris={scene.triangle_count} rays={len(rays)}", flush=True)
    bvh = device.upload(scene)
    gpu = bvh.intersect(rays, tmin, tmax, barycentric=True)
    o = oracle.OracleScene(scene)
    r64 = rays.astype(np.float64)
    twin = o.intersect(r64, tmin, tmax, precision=32)
    ref = o.intersect(r64, tmin, tmax, precision=64, second=True)
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
    occ_twin, _ = o.occluded(r64, tmin, tmax, precision=32)
    np.testing.assert_array_equal(occ, occ_twin)
    return gpu


def test_icosphere_random_rays(device, oracle):
    scene = scenegen.icosphere(level=5)
    rays = scenegen.intersection_rays(scene, 1 << 17)
    gpu = _compare(device, oracle, scene, rays)
    assert (gpu["primitive"] >= 0).mean() > 0.3


def test_cornell_single_sided_rays(device, oracle):
    scene = scenegen.cornell_box()
    rays = scenegen.intersection_rays(scene, 1 << 15, seed=7, scale=0.9)
    _compare(device, oracle, scene, rays)
    _compare(device, oracle, scene, rays, tmin=1e-3, tmax=1.5)


def test_empty_and_single_triangle(device, oracle):
    one = scenegen.floor_scene()
    from minirender_amd.scene import concat, material
    tri = concat([(one.positions[:1], one.normals[:1])], [material()])
    rays = np.array([[0.5, 1, 1.5, 0, -1, 0], [3, 1, -3, 0, -1, 0], [0.5, 1, 1.5, 0, 1, 0]], np.float32)
    gpu = device.upload(tri).intersect(rays)
    o = oracle.OracleScene(tri).intersect(rays.astype(np.float64), precision=32)
    np.testing.assert_array_equal(gpu["primitive"], o["prim"])
    empty = concat([], [])
    e = device.upload(empty).intersect(rays)
    assert (e["primitive"] == -1).all()


def test_coincident_triangles_tie_break(device, oracle):
    """256 coincident triangles: the lowest scene index wins (bvh.cpp:273-276)."""
    from minirender_amd.scene import concat, material
    t = np.array([[[0, 0, 0], [1, 0, 0], [0, 1, 0]]], np.float32)
    n = np.tile(np.array([0, 0, 1], np.float32), (1, 3, 1))
    scene = concat([(t, n)] * 256, [material()] * 256)
    rays = np.array([[0.25, 0.25, 2, 0, 0, -1]], np.float32)
    gpu = device.upload(scene).intersect(rays)
    assert gpu["primitive"][0] == 0
    assert gpu["t"][0] == np.float32(2)


def test_full_size_microbench_subset(device, oracle):
    """The 1,310,720-triangle displaced icosphere at full size; every 16th ray of the seeded 16M-ray
    set (1,048,576 rays). Ids match the FP64 reference outside near-ties; t within 1e-5 relative."""
    scene = scenegen.icosphere(level=8)
    rays = scenegen.intersection_rays(scene, 0, indices=np.arange(0, 1 << 24, 16))
    _compare(device, oracle, scene, rays)
