"""Vectorised form of the reference sampler streams, for generating seeded inputs on the host.

Follows src/core/sampling.cpp:28-65: SplitMix64 mixing derives a PCG32 (XSH-RR) stream per
(seed, pixel, sample); each dimension consumes two u32 draws and yields 53 random bits. numpy's
uint64 arithmetic wraps modulo 2^64, exactly like the C++ code.
"""
from __future__ import annotations

import numpy as np

_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_MUL = np.uint64(6364136223846793005)
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_SAMPLE_C = np.uint64(0xD1B54A32D192ED03)


def mix64(v):
    v = np.asarray(v, np.uint64)
    with np.errstate(over="ignore"):
        v = (v ^ (v >> np.uint64(30))) * _M1
        v = (v ^ (v >> np.uint64(27))) * _M2
    return v ^ (v >> np.uint64(31))


class Streams:
    """A batch of independent PCG32 streams (one per element)."""

    def __init__(self, seed, sequence):
        seed = np.asarray(seed, np.uint64)
        sequence = np.asarray(sequence, np.uint64)
        seed, sequence = np.broadcast_arrays(seed, sequence)
        self.inc = (sequence << np.uint64(1)) | np.uint64(1)
        self.state = np.zeros_like(self.inc)
        self.next_u32()
        with np.errstate(over="ignore"):
            self.state = self.state + seed
        self.next_u32()

    @classmethod
    def for_sample(cls, seed, pixel, sample):
        seed = np.asarray(seed, np.uint64)
        pixel = np.asarray(pixel, np.uint64)
        sample = np.asarray(sample, np.uint64)
        with np.errstate(over="ignore"):
            return cls(mix64(seed ^ mix64(pixel + _GOLDEN)), mix64(sample + _SAMPLE_C) ^ mix64(pixel))

    def next_u32(self):
        prev = self.state
        with np.errstate(over="ignore"):
            self.state = prev * _MUL + self.inc
        shifted = (((prev >> np.uint64(18)) ^ prev) >> np.uint64(27)).astype(np.uint32)
        rot = (prev >> np.uint64(59)).astype(np.uint32)
        return (shifted >> rot) | (shifted << ((np.uint32(32) - rot) & np.uint32(31)))

    def next_bits53(self):
        high = (self.next_u32() >> np.uint32(5)).astype(np.uint64)
        low = (self.next_u32() >> np.uint32(6)).astype(np.uint64)
        return (high << np.uint64(26)) | low

    def next_1d(self):
        """FP64 dimension in [0, 1), as Sampler::next_1d."""
        return self.next_bits53().astype(np.float64) / 9007199254740992.0

    def next_1d_f32(self):
        """The FP32 rounding of the same 53-bit value, clamped below 1."""
        v = self.next_bits53().astype(np.float32) * np.float32(2.0 ** -53)
        return np.minimum(v, np.float32(1 - 2.0 ** -24))


def sample_sphere(u0, u1):
    """Uniform direction (sampling.cpp:90-95), FP64."""
    z = 1 - 2 * u0
    r = np.sqrt(np.maximum(0.0, 1 - z * z))
    return np.stack([r * np.cos(2 * np.pi * u1), r * np.sin(2 * np.pi * u1), z], axis=-1)
