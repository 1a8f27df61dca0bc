"""Writes the golden fixtures of tests/golden from the CPU oracle (run from the repo root:
`python tests/golden/make_golden.py`). Fixtures are data only: inputs and expected outputs."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from minirender_amd import scenegen  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main():
    # sampler streams: PCG demo KAT and for_sample prefixes
    keys = np.array([[0, 0, 0], [1, 2, 3], [12, 0, 7], [2, 1048575, 0], [0xDEADBEEF, 123456789, 255]], np.uint64)
    streams = np.stack([O.for_sample_u32(int(s), int(p), int(i), 8) for s, p, i in keys])
    np.savez_compressed(os.path.join(HERE, "sampler.npz"), pcg42_54=O.pcg(42, 54, 6), keys=keys, for_sample=streams)
    # closest hits on a 20,480-triangle displaced icosphere (intersection-microbench style rays)
    sc = scenegen.icosphere(level=5)
    rays = scenegen.intersection_rays(sc, 4096)
    o = O.OracleScene(sc)
    r64 = o.intersect(rays.astype(np.float64), precision=64, second=True)
    r32 = o.intersect(rays.astype(np.float64), precision=32)
    np.savez_compressed(os.path.join(HERE, "icosphere5_hits.npz"), rays=rays, prim64=r64["prim"], t64=r64["t"],
                        t2_64=r64["t2"], prim32=r32["prim"], t32=r32["t"])
    # 16x16 x 4 spp renders (FP64 reference and FP32 twin)
    for name, scene, depth in (("cornell", scenegen.cornell_box(), 8),
                               ("mixed_small", scenegen.mixed_terrain(cells=64, spheres=16, sphere_level=2), 16)):
        oo = O.OracleScene(scene)
        out = {}
        for p in (64, 32):
            img, sums, st = oo.render(scene.camera, 16, 16, 4, seed=0, mode="mis", max_depth=depth, precision=p,
                                      sums=True)
            out[f"image{p}"] = img
            out[f"sums{p}"] = sums
            out[f"paths{p}"] = st["paths"]
        np.savez_compressed(os.path.join(HERE, f"render_{name}_16x16x4.npz"), **out)
    print("golden fixtures written")


if __name__ == "__main__":
    main()
