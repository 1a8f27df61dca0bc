"""`mrg_render --render` on the GPU: a JSON v1 scene with OBJ meshes, transforms, an emissive
single-sided quad, a conductor and a directional light, rendered through the C++ session and
written as EXR; the EXR must match the oracle's FP32 twin of the flattened scene (1e-4 relative),
like the other render parity tests, and the summary must count every path (the reference CLI's
"paths: W*H*spp" check, CMakeLists.txt:208-211)."""
import numpy as np
import pytest

import _formats as F
from test_scene_io import CUBE, cli, write_scene

pytestmark = pytest.mark.gpu

QUAD = "v -1 0 -1\nv 1 0 -1\nv 1 0 1\nv -1 0 1\nf 1 4 3 2\n"          # faces +y
LAMP = "v -0.3 0 -0.3\nv 0.3 0 -0.3\nv 0.3 0 0.3\nv -0.3 0 0.3\nf 1 2 3 4\n"  # faces -y


def test_cli_render_matches_oracle(oracle, tmp_path):
    materials = [{"id": "white", "kind": "diffuse", "base_color": [0.73, 0.73, 0.73], "double_sided": False},
                 {"id": "red", "kind": "diffuse", "base_color": [0.65, 0.05, 0.05]},
                 {"id": "metal", "kind": "conductor", "base_color": [0.9, 0.8, 0.6], "roughness": 0.3},
                 {"id": "lamp", "kind": "diffuse", "base_color": [0.0, 0.0, 0.0], "emission": [17, 12, 4],
                  "double_sided": False}]
    objects = [{"id": "floor", "mesh": "quad", "obj": QUAD, "s": [2, 1, 2]},
               {"id": "back", "mesh": "quad", "obj": QUAD, "t": [0, 1, -1], "r": [90, 0, 0], "material": "red"},
               {"id": "box", "mesh": "cube", "obj": CUBE, "t": [-0.3, 0.3, 0.1], "r": [0, 17, 0], "s": [0.3, 0.3, 0.3],
                "material": "metal"},
               {"id": "lamp", "mesh": "lamp", "obj": LAMP, "t": [0, 1.9, 0], "material": "lamp"}]
    path = write_scene(tmp_path, objects, materials=materials, environment=(0.05, 0.06, 0.08),
                       lights=[{"id": "sun", "direction": [-0.3, -1, -0.2], "color": [1, 0.9, 0.8], "intensity": 1.5}],
                       camera={"position": [0, 1, 3.2], "target": [0, 0.8, 0], "up": [0, 1, 0], "vertical_fov": 45,
                               "near": 0.03, "far": 100})
    w, h, spp, depth, seed = 32, 24, 8, 6, 3
    print(f"[cli render] {w}x{h} spp={spp} depth={depth}", flush=True)
    cli("--scene-info", path, tmp_path / "flat.bin")
    out = cli("--render", path, tmp_path / "img", "--width", w, "--height", h, "--spp", spp, "--depth", depth,
              "--seed", seed, "--batch", 2, "--exposure", 0.5, "--integrator", "gpu").stdout
    assert f"paths: {w * h * spp}\n" in out and "triangles: 18\n" in out
    img = F.exr_rgb(tmp_path / "img.exr")
    assert F.read_png(tmp_path / "img.png").shape == (h, w, 3)
    scene = F.read_flat_scene(tmp_path / "flat.bin")
    twin, _ = oracle.OracleScene(scene).render(scene.camera, w, h, spp, seed=seed, max_depth=depth, precision=32)
    diff = np.abs(img.astype(np.float64) - twin.astype(np.float64))
    assert (diff <= 1e-4 * np.maximum(np.abs(twin.astype(np.float64)), 1e-6)).all()
    assert img.max() > 0.05  # the lamp and the sun light the scene


QUAD_UV = "v -1 0 -1\nv 1 0 -1\nv 1 0 1\nv -1 0 1\nvt 0 0\nvt 2 0\nvt 2 2\nvt 0 2\nf 1/1 4/4 3/3 2/2\n"


def test_cli_render_textured_scene_matches_oracle(oracle, tmp_path):
    """A JSON scene whose materials use an sRGB PNG base-color texture (nearest and bilinear) and a
    linear one, lit by an environment map (the sRGB texture, strength 2, rotated 45 degrees), decoded
    by the command line and rendered through the C++ session, against the FP32 twin of the dumped
    flattened scene (textures included)."""
    from test_textures import write_png
    rng = np.random.default_rng(8)
    write_png(tmp_path / "albedo.png", rng.integers(0, 256, (6, 9, 3)), 2, 8)
    write_png(tmp_path / "mask.png", rng.integers(0, 65536, (5, 4, 1)), 0, 16)
    materials = [{"id": "floor", "kind": "diffuse", "base_color": [0.9, 0.9, 0.9], "base_color_texture": "albedo",
                  "texture_filter": "nearest"},
                 {"id": "metal", "kind": "conductor", "base_color": [0.9, 0.8, 0.6], "roughness": 0.4,
                  "base_color_texture": "mask", "texture_filter": "bilinear"},
                 {"id": "back", "kind": "diffuse", "base_color": [0.7, 0.7, 0.7], "base_color_texture": "albedo",
                  "normal_texture": "mask", "normal_scale": 0.7, "normal_flip_y": True}]
    objects = [{"id": "floor", "mesh": "quad", "obj": QUAD_UV, "s": [2, 1, 2], "material": "floor"},
               {"id": "back", "mesh": "quad", "obj": QUAD_UV, "t": [0, 1, -1], "r": [90, 0, 0], "material": "back"},
               {"id": "box", "mesh": "cube", "obj": CUBE, "t": [0.2, 0.3, 0.1], "r": [0, 25, 0], "s": [0.3, 0.3, 0.3],
                "material": "metal"}]
    path = write_scene(tmp_path, objects, materials=materials, environment=(0.4, 0.45, 0.5),
                       lights=[{"id": "sun", "direction": [-0.3, -1, -0.4], "color": [1, 0.9, 0.8], "intensity": 2.0}],
                       camera={"position": [0, 1.2, 3.0], "target": [0, 0.5, 0], "up": [0, 1, 0], "vertical_fov": 45,
                               "near": 0.03, "far": 100},
                       extra={"textures": [{"id": "albedo", "source": "albedo.png", "encoding": "srgb"},
                                           {"id": "mask", "source": "mask.png", "encoding": "linear"}],
                              "environment_map": {"texture": "albedo", "strength": 2.0, "rotation_degrees": 45}})
    w, h, spp, depth, seed = 32, 24, 6, 5, 12
    print(f"[cli textured render] {w}x{h} spp={spp} depth={depth}", flush=True)
    cli("--scene-info", path, tmp_path / "flat.bin")
    out = cli("--render", path, tmp_path / "img", "--width", w, "--height", h, "--spp", spp, "--depth", depth,
              "--seed", seed, "--integrator", "gpu").stdout
    assert f"paths: {w * h * spp}\n" in out
    img = F.exr_rgb(tmp_path / "img.exr")
    scene = F.read_flat_scene(tmp_path / "flat.bin")
    assert len(scene.textures) == 2 and scene.textures[0].shape == (6, 9, 3)
    assert scene.environment_texture == 0 and scene.environment_strength == 2 and scene.environment_rotation_degrees == 45
    twin, _ = oracle.OracleScene(scene).render(scene.camera, w, h, spp, seed=seed, max_depth=depth, precision=32)
    diff = np.abs(img.astype(np.float64) - twin.astype(np.float64))
    assert (diff <= 1e-4 * np.maximum(np.abs(twin.astype(np.float64)), 1e-6)).all()
