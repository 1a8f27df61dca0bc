"""Pins the oracle's sampler, warps and discrete distribution against the reference's known-answer
tests (tests/sampling_tests.cpp) and against mathematics (moments, closed forms, library sin/cos)."""
import math

import numpy as np
import pytest

from minirender_amd import rng


def test_pcg_known_answer(oracle):
    """sampling_tests.cpp:59-62: Sampler(42, 54) follows the published pcg32 demo sequence."""
    assert list(oracle.pcg(42, 54, 6)) == [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]


def test_for_sample_streams_independent_and_reproducible(oracle):
    a = oracle.for_sample_u32(12, 0, 7, 64)
    assert (a == oracle.for_sample_u32(12, 0, 7, 64)).all()
    assert a[0] != oracle.for_sample_u32(12, 0, 8, 1)[0]
    assert a[0] != oracle.for_sample_u32(13, 0, 7, 1)[0]
    assert (a != oracle.for_sample_u32(12, 1, 7, 64)).any()


def test_vectorised_streams_match_oracle(oracle):
    """Two independent implementations (numpy host generator vs the C++ oracle) agree bit for bit."""
    r = np.random.default_rng(0)
    for _ in range(20):
        seed, pix, smp = (int(x) for x in r.integers(0, 2**63, 3))
        s = rng.Streams.for_sample(np.uint64(seed), np.array([pix], np.uint64), np.uint64(smp))
        mine = [int(s.next_u32()[0]) for _ in range(16)]
        assert mine == list(oracle.for_sample_u32(seed, pix, smp, 16))


def test_next_1d_is_53_bit_and_twin_rounds(oracle):
    d, f = oracle.next_1d(3, 5, 7, 4096)
    assert ((d >= 0) & (d < 1)).all()
    assert np.all(d * 2.0**53 == np.floor(d * 2.0**53))
    assert ((f >= 0) & (f < 1)).all()
    expect = np.minimum(d.astype(np.float32), np.float32(1 - 2**-24))
    np.testing.assert_array_equal(f, expect)


def _check_frequency(count, p, total):
    expected = p * total
    assert abs(count - expected) < 5 * math.sqrt(expected * (1 - p)) + 8


@pytest.mark.parametrize("precision", [64, 32])
def test_warps_match_analytic_moments(oracle, precision):
    """sampling_tests.cpp:84-139."""
    n = 100000
    st = rng.Streams(np.full(n, 914, np.uint64), np.arange(n, dtype=np.uint64))
    u = np.stack([st.next_1d(), st.next_1d()], axis=1)
    disk = oracle.warp("disk", u, precision)
    r2 = disk[:, 0] ** 2 + disk[:, 1] ** 2
    assert r2.max() <= 1 + 1e-6
    assert abs(r2.mean() - 0.5) < 0.006
    for b in range(4):
        _check_frequency(np.sum(np.minimum((r2 * 4).astype(int), 3) == b), 0.25, n)
    sph = oracle.warp("sphere", u, precision)
    np.testing.assert_allclose(np.linalg.norm(sph, axis=1), 1, atol=1e-6 if precision == 32 else 1e-12)
    assert np.linalg.norm(sph.mean(0)) < 0.01
    cos = oracle.warp("cosine", u, precision)
    assert (cos[:, 2] >= 0).all()
    assert abs(cos[:, 2].mean() - 2 / 3) < 0.006
    for b in range(4):
        lo, hi = b / 4, (b + 1) / 4
        _check_frequency(np.sum(np.minimum((cos[:, 2] * 4).astype(int), 3) == b), hi * hi - lo * lo, n)
    tri = oracle.warp("triangle", u, precision)
    np.testing.assert_allclose(tri.sum(1), 1, atol=1e-6 if precision == 32 else 1e-12)
    assert np.all(np.abs(tri.mean(0) - 1 / 3) < 0.006)
    assert abs((tri[:, 0] * tri[:, 1]).mean() - 1 / 12) < 0.002
    assert (oracle.warp("disk", [[0.5, 0.5]], precision) == 0).all()
    with pytest.raises(RuntimeError):
        oracle.warp("sphere", [[1.0, 0.0]], precision)


def test_project_sincos_against_library():
    """The FP32 twin's sin/cos(2*pi*t) (quarter-turn reduction + Taylor series) stays within 2 float
    ulps of the correctly rounded value computed in FP64 by the C library."""
    from oracle import oracle as O
    t = np.concatenate([np.linspace(-1, 1, 200001, dtype=np.float32), np.float32([0, 0.125, 0.25, 0.5, 0.75, -0.25])])
    s, c = O.sincos_turns(t)
    ref_s = np.sin(2 * np.pi * t.astype(np.float64))
    ref_c = np.cos(2 * np.pi * t.astype(np.float64))
    ulp = np.spacing(np.float32(1.0)) / 2  # absolute error scale near |v| <= 1
    assert np.max(np.abs(s - ref_s)) <= 2 * ulp
    assert np.max(np.abs(c - ref_c)) <= 2 * ulp
    assert s[np.where(t == 0)[0][0]] == 0 and c[np.where(t == 0.25)[0][0]] == 0


def test_discrete_distribution(oracle):
    """sampling_tests.cpp:140-172."""
    p, i, mass, rem = oracle.discrete([0, 1, 3, 0], 0.0)
    np.testing.assert_allclose(p, [0, 0.25, 0.75, 0])
    assert i == 1 and mass == 0.25
    assert oracle.discrete([0, 1, 3, 0], 0.25)[1] == 2
    assert oracle.discrete([0, 1, 3, 0], np.nextafter(1.0, 0.0))[1] == 2
    np.testing.assert_allclose(oracle.discrete([0, 0, 0], 0.5)[0], 1 / 3)
    np.testing.assert_allclose(oracle.discrete([1e308, 1e308], 0.5)[0], 0.5)
    with pytest.raises(RuntimeError):
        oracle.discrete([-1.0], 0.5)
    with pytest.raises(RuntimeError):
        oracle.discrete([0, 1, 3, 0], 1.0)
