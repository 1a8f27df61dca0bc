"""The oracle reproduces the committed golden fixtures (regression pin of the checker itself)."""
import os

import numpy as np

from minirender_amd import scenegen

G = os.path.join(os.path.dirname(__file__), "golden")


def test_sampler_fixture(oracle):
    f = np.load(os.path.join(G, "sampler.npz"))
    assert list(f["pcg42_54"]) == [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]
    for k, row in zip(f["keys"], f["for_sample"]):
        np.testing.assert_array_equal(oracle.for_sample_u32(int(k[0]), int(k[1]), int(k[2]), 8), row)


def test_intersection_fixture(oracle):
    f = np.load(os.path.join(G, "icosphere5_hits.npz"))
    sc = scenegen.icosphere(level=5)
    np.testing.assert_array_equal(scenegen.intersection_rays(sc, 4096), f["rays"])
    o = oracle.OracleScene(sc)
    r = o.intersect(f["rays"].astype(np.float64), precision=64)
    np.testing.assert_array_equal(r["prim"], f["prim64"])
    np.testing.assert_array_equal(r["t"], f["t64"])


def test_render_fixtures(oracle):
    for name, scene, depth in (("cornell", scenegen.cornell_box(), 8),
                               ("mixed_small", scenegen.mixed_terrain(cells=64, spheres=16, sphere_level=2), 16)):
        f = np.load(os.path.join(G, f"render_{name}_16x16x4.npz"))
        img, _ = oracle.OracleScene(scene).render(scene.camera, 16, 16, 4, seed=0, mode="mis", max_depth=depth,
                                                  precision=32)
        np.testing.assert_array_equal(img, f["image32"])
