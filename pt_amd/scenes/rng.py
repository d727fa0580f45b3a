"""The reference's host random generators (src/rng.cpp:24-60), in Python.

``Splitmix32`` and ``Xshiro128`` (xoshiro128+-style) are used to seed every procedural
scene, as ``calculate_object_intensity`` seeds its sampler (src/scene/lightsample.cpp:43-46):
``Splitmix32(seed)`` gives the four ``Xshiro128`` state words.

``splitmix32_stream`` is the vectorised form of the same Splitmix32 sequence: output i
(0-based) depends only on ``seed + (i + 1) * 0x9e3779b9``, so large arrays are drawn in one
numpy pass and equal the scalar stream exactly.
"""
import numpy as np

M32 = 0xFFFFFFFF
_GOLDEN = 0x9E3779B9
_INV_U32MAX = np.float32(4294967295.0)  # UINT32_MAX as the float the reference divides by


def _mix(z):
    z = ((z ^ (z >> 16)) * 0x85EBCA6B) & M32
    z = ((z ^ (z >> 13)) * 0xC2B2AE35) & M32
    return z ^ (z >> 16)


class Splitmix32:
    def __init__(self, state):
        self.state = state & M32

    def next(self):
        self.state = (self.state + _GOLDEN) & M32
        return _mix(self.state)


def _rotl(x, k):
    return ((x << k) | (x >> (32 - k))) & M32


class Xshiro128:
    def __init__(self, s0, s1, s2, s3):
        self.s = [s0 & M32, s1 & M32, s2 & M32, s3 & M32]

    @classmethod
    def seeded(cls, seed):
        sm = Splitmix32(seed)
        return cls(sm.next(), sm.next(), sm.next(), sm.next())

    def next(self):
        s = self.s
        result = (_rotl((s[0] + s[3]) & M32, 7) + s[0]) & M32
        t = (s[1] << 9) & M32
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = _rotl(s[3], 11)
        return result

    def next_float(self):
        """``static_cast<float>(next()) / UINT32_MAX`` in fp32 (rng.cpp:42-45)."""
        return float(np.float32(self.next()) / _INV_U32MAX)

    def uniform(self, lo, hi):
        return lo + (hi - lo) * self.next_float()

    def log_uniform(self, lo, hi):
        return float(np.exp(np.log(lo) + (np.log(hi) - np.log(lo)) * self.next_float()))


def splitmix32_stream(seed, n):
    """The first n outputs of Splitmix32(seed) as uint32 (vectorised)."""
    i = np.arange(1, n + 1, dtype=np.uint64)
    z = (np.uint64(seed & M32) + i * np.uint64(_GOLDEN)) & np.uint64(M32)
    z = ((z ^ (z >> np.uint64(16))) * np.uint64(0x85EBCA6B)) & np.uint64(M32)
    z = ((z ^ (z >> np.uint64(13))) * np.uint64(0xC2B2AE35)) & np.uint64(M32)
    z = z ^ (z >> np.uint64(16))
    return z.astype(np.uint32)


def splitmix32_floats(seed, n):
    """Stream outputs mapped to [0, 1] the way Xshiro128::next_float maps its words."""
    return splitmix32_stream(seed, n).astype(np.float32) / _INV_U32MAX
