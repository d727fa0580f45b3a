"""The device evaluates Sobol dimensions 1-3 without a loop over the index bits
(pt_amd/csrc/shading.h sobol): dimension 1 (polynomial x + 1, Pascal's triangle mod 2) as a
superset sum over the bit positions followed by a bit reversal, dimensions 2 and 3 by per-byte
XOR tables. Both are the same GF(2) linear map as the definition (XOR of the direction numbers
of the set bits, sample.hlsli:31-67 with the Joe-Kuo polynomials x+1, x^2+x+1, x^3+x+1)."""
import random


def _directions():
    v = [[0] * 32 for _ in range(4)]
    for b in range(32):
        v[0][b] = 1 << (31 - b)
    for d, (s, poly, minit) in enumerate([(1, 0, [1]), (2, 1, [1, 3]), (3, 1, [1, 3, 1])], start=1):
        m = list(minit) + [0] * (32 - s)
        for k in range(s, 32):
            val = m[k - s] ^ ((m[k - s] << s) & 0xFFFFFFFF)
            for j in range(1, s):
                if (poly >> (s - 1 - j)) & 1:
                    val ^= (m[k - j] << j) & 0xFFFFFFFF
            m[k] = val & 0xFFFFFFFF
        v[d] = [(m[k] << (31 - k)) & 0xFFFFFFFF for k in range(32)]
    return v


V = _directions()


def _definition(i, d):
    x = 0
    for b in range(32):
        if (i >> b) & 1:
            x ^= V[d][b]
    return x


def _rev(x):
    return int("{:032b}".format(x)[::-1], 2)


def test_dimension_1_closed_form():
    assert V[1][:4] == [0x80000000, 0xC0000000, 0xA0000000, 0xF0000000]  # Pascal rows 1, 11, 101, 1111
    r = random.Random(5)
    for i in [0, 1, 2, 3, 0xFFFFFFFF, 0x80000000] + [r.getrandbits(32) for _ in range(5000)]:
        z = i
        for k, mask in enumerate([0x55555555, 0x33333333, 0x0F0F0F0F, 0x00FF00FF, 0x0000FFFF]):
            z ^= (z >> (1 << k)) & mask
        assert _rev(z) == _definition(i, 1), hex(i)


def test_dimensions_2_3_byte_tables():
    r = random.Random(6)
    for d in (2, 3):
        t = [[0] * 256 for _ in range(4)]
        for c in range(4):
            for x in range(256):
                for b in range(8):
                    if (x >> b) & 1:
                        t[c][x] ^= V[d][8 * c + b]
        for i in [0, 0xFFFFFFFF] + [r.getrandbits(32) for _ in range(5000)]:
            got = t[0][i & 255] ^ t[1][(i >> 8) & 255] ^ t[2][(i >> 16) & 255] ^ t[3][i >> 24]
            assert got == _definition(i, d)
