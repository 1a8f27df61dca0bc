"""Host BVH builder invariants (no GPU): every primitive referenced once, leaves <= 16 triangles,
depth <= 30 (the traversal stack bound), for ordinary and pathological inputs."""
import ctypes

import numpy as np
import pytest

from minirender_amd import _native as N, scenegen


def build(positions):
    p = np.ascontiguousarray(positions, np.float32).reshape(-1, 9)
    out = np.zeros(7, np.uint64)
    N.check(N.lib().mrg_bvh_build_host(p.ctypes.data_as(ctypes.c_void_p), len(p), out.ctypes.data_as(ctypes.c_void_p)))
    return dict(nodes=int(out[0]), leaves=int(out[1]), depth=int(out[2]), max_leaf=int(out[3]), prims=int(out[4]),
                bad=int(out[5]), stack=int(out[6]))


@pytest.mark.parametrize("n", [1, 2, 3, 5, 17, 1000])
def test_small_random(n):
    rng = np.random.default_rng(n)
    s = build(rng.normal(size=(n, 3, 3)))
    assert s["bad"] == 0 and s["prims"] == n and s["max_leaf"] <= 16 and s["depth"] <= 30


def test_coincident_centroids():
    t = np.tile(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), (5000, 1, 1))
    s = build(t)
    assert s["bad"] == 0 and s["prims"] == 5000 and s["max_leaf"] <= 16 and s["depth"] <= 30


def test_geometric_series_depth_bound():
    # triangles at positions 2^-k: SAH alone would build a very deep chain
    k = np.arange(3000)
    base = np.ldexp(1.0, -(k % 120)).astype(np.float64)[:, None, None]
    tri = np.array([[0, 0, 0], [1e-3, 0, 0], [0, 1e-3, 0]])[None] + np.stack([base, base * 0, base * 0], -1).reshape(-1, 1, 3)
    s = build(tri)
    assert s["bad"] == 0 and s["depth"] <= 30 and s["max_leaf"] <= 16


def test_icosphere_level6():
    sc = scenegen.icosphere(level=6)
    s = build(sc.positions)
    assert s["bad"] == 0 and s["prims"] == sc.triangle_count and s["depth"] <= 30 and s["stack"] <= 64
    assert s["nodes"] < sc.triangle_count / 3  # four-wide nodes: about a third of a binary tree's


@pytest.mark.parametrize("case", ["far_offset", "tiny", "mixed_scale", "flat", "negative_huge"])
def test_quantized_boxes_contain_exact_boxes(case):
    # The device reads child boxes quantized to 8 bits per plane on an exact per-node grid; the build
    # check counts any decoded child box that does not contain its exact box (as "bad").
    rng = np.random.default_rng(7)
    n = 4000
    small = rng.normal(size=(n, 3, 3)) * 1e-3
    centers = rng.uniform(-1, 1, size=(n, 1, 3))
    if case == "far_offset":
        tri = centers * 10 + small + 1e7
    elif case == "tiny":
        tri = (centers + small) * 1e-30
    elif case == "mixed_scale":
        tri = np.concatenate([centers[: n // 2] * 1e6 + small[: n // 2], centers[n // 2:] * 1e-6 + small[n // 2:] * 1e-6])
    elif case == "flat":
        tri = centers + small
        tri[..., 2] = 0.25
    else:
        tri = -(centers + 1.5) * 3e8 + small * 1e5
    s = build(tri.astype(np.float32))
    assert s["bad"] == 0 and s["prims"] == n
