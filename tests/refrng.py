"""The reference tests' deterministic generator (test/test_mc_moments.jl:55-64): xorshift64 with
shifts 13/7/17, the seed OR-ed with 1, 53-bit mantissas, and Box-Muller normals."""
import math

import numpy as np

_MASK = (1 << 64) - 1


class XorShift:
    def __init__(self, seed=0x2545f4914f6cdd1d):
        self.state = (int(seed) | 1) & _MASK

    def rand(self):
        x = self.state
        x ^= (x << 13) & _MASK
        x ^= x >> 7
        x ^= (x << 17) & _MASK
        self.state = x
        return float(x >> 11) * (1.0 / 9007199254740992.0)

    def randn(self):
        u = self.rand()
        v = self.rand()
        return math.sqrt(-2.0 * math.log(1.0 - u)) * math.cos(math.pi * 2.0 * v)


def seeded(seed):
    return XorShift(seed)


def randn_field(rng, shape):
    """A column-major field of normals, first index fastest (Julia's comprehension order)."""
    n = int(np.prod(shape))
    return np.array([rng.randn() for _ in range(n)])
