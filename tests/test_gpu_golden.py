"""GPU results against the committed golden fixtures (no oracle at run time)."""
import os

import numpy as np
import pytest

from minirender_amd import api, scenegen

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(__file__), "golden")


def test_hits_match_fixture(device):
    f = np.load(os.path.join(G, "icosphere5_hits.npz"))
    print("[golden] icosphere5 4096 rays", flush=True)
    scene = scenegen.icosphere(level=5)
    # FP64 hit decisions (the default): the FP64 fixture's ids and distances, near-ties included
    hits = device.upload(scene).intersect(f["rays"])
    np.testing.assert_array_equal(hits["primitive"], f["prim64"])
    hit = f["prim64"] >= 0
    np.testing.assert_array_equal(hits["t"][hit], f["t64"][hit].astype(np.float32))
    # FP32 decisions: the FP32 twin's fixture; ids equal FP64 outside near-ties
    with api.Device(0, query_precision=32, path_capacity=1 << 16) as d32:
        hits = d32.upload(scene).intersect(f["rays"])
    np.testing.assert_array_equal(hits["primitive"], f["prim32"])
    hit = f["prim32"] >= 0
    np.testing.assert_array_equal(hits["t"][hit], f["t32"][hit].astype(np.float32))
    with np.errstate(invalid="ignore"):
        near = np.abs(f["t2_64"] - f["t64"]) <= 1e-6 * np.abs(f["t64"])
    assert np.all((hits["primitive"] == f["prim64"]) | near)


@pytest.mark.parametrize("name,depth", [("cornell", 8), ("mixed_small", 16)])
def test_render_matches_fixture(device, name, depth):
    f = np.load(os.path.join(G, f"render_{name}_16x16x4.npz"))
    scene = scenegen.cornell_box() if name == "cornell" else scenegen.mixed_terrain(cells=64, spheres=16, sphere_level=2)
    print(f"[golden] render {name} 16x16x4", flush=True)
    job = api.Job(device.upload(scene), scene.camera, api.RenderSettings(16, 16, 4, 0), api.IntegratorSettings("mis", depth))
    job.wait()
    sums, counts = job.film()
    assert (counts == 4).all()
    rel = np.abs(sums - f["sums32"]) / np.maximum(np.abs(f["sums32"]), 1e-6)
    assert rel.max() <= 1e-4
    job.release()
