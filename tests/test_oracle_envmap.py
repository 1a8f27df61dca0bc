"""Oracle pinning for the environment map (EnvironmentLight, light.cpp:61-191): the reference's own
known answers (light_tests.cpp:71-163) in FP64, the FP32 twin against FP64, and the twin's
project-written atan2 against the C library."""
import math

import numpy as np
import pytest

from minirender_amd.scene import concat


def latlong(u, v):
    """light_tests.cpp:34-37"""
    return np.array([math.sin(v * math.pi) * math.cos(u * 2 * math.pi), math.cos(v * math.pi),
                     math.sin(v * math.pi) * math.sin(u * 2 * math.pi)])


def env_scene(image, environment=(0, 0, 0), strength=1.0, rotation=0.0):
    sc = concat([], [], environment=environment, textures=[np.asarray(image, np.float32)], environment_texture=0,
                 environment_strength=strength, environment_rotation_degrees=rotation)
    return sc


def test_atan2_turns_is_within_three_ulps_of_libm(oracle):
    rng = np.random.default_rng(1)
    y = np.concatenate([rng.standard_normal(200000), [0.0, -0.0, 1.0, -1.0, 0.0, 1e-30, 3.0, -2.0]]).astype(np.float32)
    x = np.concatenate([rng.standard_normal(200000), [1.0, 1.0, 0.0, 0.0, -1.0, 1.0, 3.0, -2.0]]).astype(np.float32)
    got = oracle.atan2_turns(y, x).astype(np.float64)
    want = np.arctan2(y.astype(np.float64), x.astype(np.float64)) / (2 * math.pi)
    ulp = np.spacing(np.abs(want).astype(np.float32)).astype(np.float64)
    assert np.all(np.abs(got - want) <= 3 * np.maximum(ulp, 1e-45))
    assert oracle.atan2_turns([0.0], [-1.0])[0] == 0.5 and oracle.atan2_turns([-0.0], [-1.0])[0] == -0.5
    assert oracle.atan2_turns([1.0], [0.0])[0] == 0.25 and oracle.atan2_turns([-1.0], [0.0])[0] == -0.25


def test_constant_map_samples_uniform_solid_angle(oracle):
    """light_tests.cpp:71-104: a constant map is evaluated and sampled as the uniform sphere."""
    img = np.tile(np.array([2, 1, 0.5], np.float32), (4, 8, 1))
    o = oracle.OracleScene(env_scene(img, environment=(0.1, 0.1, 0.1), strength=2))
    expected = np.array([0.1, 0.1, 0.1], np.float32).astype(np.float64) + [4, 2, 1]
    power = float(expected @ [0.2126, 0.7152, 0.0722]) * 4 * math.pi
    assert abs(o.env_integral() - power) <= 1e-12 * power
    for d in ([0, 1, 0], [0, -1, 0], latlong(0.37, 0.2)):
        rad, pdf = o.env_eval([d])
        assert np.linalg.norm(rad[0] - expected) < 1e-12
        assert abs(pdf[0] - 1 / (4 * math.pi)) < 1e-12
    d, rad, pdf = o.env_sample([[0, 0], [0.05, 0], [np.nextafter(1, 0), np.nextafter(1, 0)]])
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-12) and np.all(np.isfinite(pdf)) and np.all(pdf > 0)
    u = np.random.default_rng(43).random((10000, 2))
    d, rad, pdf = o.env_sample(u)
    _, pdf2 = o.env_eval(d)
    np.testing.assert_allclose(pdf, pdf2, rtol=1e-12)
    assert np.linalg.norm(d.mean(axis=0)) < 0.025


def test_importance_pdf_normalizes_and_keeps_bilinear_support(oracle):
    """light_tests.cpp:106-145."""
    img = np.zeros((4, 8, 3), np.float32)
    img[1, 3] = 100
    o = oracle.OracleScene(env_scene(img))
    integral = 0.0
    for y in range(4):
        angle = (2 * math.pi / 8) * (math.cos(y * math.pi / 4) - math.cos((y + 1) * math.pi / 4))
        dirs = [latlong((x + 0.5) / 8, (y + 0.5) / 4) for x in range(8)]
        integral += o.env_eval(dirs)[1].sum() * angle
    assert abs(integral - 1) < 1e-12
    neighbor = latlong(4.1 / 8, 1.5 / 4)
    rad, pdf = o.env_eval([neighbor])
    assert abs(rad[0, 0] - 40) < 1e-10
    assert abs(pdf[0] - 0.05 / (4 * math.pi)) < 1e-12 * pdf[0]
    u = np.random.default_rng(78).random((20000, 2))
    d, rad, pdf = o.env_sample(u)
    rad2, pdf2 = o.env_eval(d)
    np.testing.assert_allclose(pdf, pdf2, rtol=1e-12)
    assert np.abs(rad - rad2).max() < 1e-12
    uu = np.arctan2(d[:, 2], d[:, 0]) / (2 * math.pi)
    uu -= np.floor(uu)
    vv = np.arctan2(np.hypot(d[:, 0], d[:, 2]), d[:, 1]) / math.pi
    bright = int(np.sum(((uu * 8).astype(int) == 3) & ((vv * 4).astype(int) == 1)))
    solid = (2 * math.pi / 8) * (math.cos(math.pi / 4) - math.cos(math.pi / 2))
    mass = 0.95 + 0.05 * solid / (4 * math.pi)
    assert abs(bright - 20000 * mass) < 5 * math.sqrt(20000 * mass * (1 - mass))
    rotated = oracle.OracleScene(env_scene(img, rotation=90))
    local = latlong(3.5 / 8, 1.5 / 4)
    world = np.array([local[2], local[1], -local[0]])
    r1, p1 = rotated.env_eval([world])
    r0, p0 = o.env_eval([local])
    assert np.linalg.norm(r1 - r0) < 1e-10 and abs(p1[0] - p0[0]) <= 1e-12 * p0[0]


def test_negative_samples_do_not_erase_the_environment(oracle):
    """light_tests.cpp:147-160: mip averaging cancels +1/-1 texels, the selection power and the
    evaluation keep the positive radiance, and sampling falls back to the uniform sphere."""
    img = np.where(np.arange(2048) % 2 == 0, 1.0, -1.0).astype(np.float32)[None, :, None].repeat(3, axis=2)
    o = oracle.OracleScene(env_scene(img))
    assert oracle.texture_level(img, 1)[0, 0].tolist() == [0, 0, 0]
    assert len(o.light_selection()) == 1
    assert abs(o.env_integral() - 2 * math.pi) <= 1e-12 * 2 * math.pi
    assert abs(o.env_eval([latlong(0.5 / 2048, 0.5)])[0][0, 0] - 1) < 1e-10
    assert abs(o.env_eval([[1, 0, 0]])[1][0] - 1 / (4 * math.pi)) < 1e-12


@pytest.mark.parametrize("rotation", [0.0, 37.0, -200.0])
def test_fp32_twin_follows_fp64(oracle, rotation):
    """The twin's evaluate / pdf / sample agree with FP64 to float accuracy on a random map."""
    rng = np.random.default_rng(9)
    img = (rng.random((16, 32, 3)) ** 3 * 4).astype(np.float32)
    o = oracle.OracleScene(env_scene(img, environment=(0.05, 0.05, 0.05), strength=1.5, rotation=rotation))
    dirs = rng.standard_normal((5000, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    r64, p64 = o.env_eval(dirs, precision=64)
    r32, p32 = o.env_eval(dirs.astype(np.float32).astype(np.float64), precision=32)
    # The twin's texel probability is the difference of two float CDF values (each within 2^-24 of
    # the FP64 CDF), divided by the row's solid angle: that bounds its pdf error. A direction on a
    # texel boundary may also round to the neighbouring cell in float: compare the bulk.
    solid_min = (2 * math.pi / 32) * (1 - math.cos(math.pi / 16))
    close = np.abs(p32 - p64) <= 1e-5 * p64 + 0.95 * 2.0 ** -23 / solid_min
    assert close.mean() > 0.995
    np.testing.assert_allclose(r32[close], r64[close], rtol=2e-3, atol=1e-5)
    u = rng.random((5000, 2))
    d64, _, q64 = o.env_sample(u, precision=64)
    d32, _, q32 = o.env_sample(u.astype(np.float32).astype(np.float64), precision=32)
    same = np.abs(q32 - q64) <= 1e-5 * q64 + 0.95 * 2.0 ** -23 / solid_min
    assert same.mean() > 0.995
    assert np.abs(d32[same] - d64[same]).max() < 1e-3
    assert abs(o.env_integral(32) - o.env_integral(64)) <= 1e-6 * o.env_integral(64)
