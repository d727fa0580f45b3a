"""JPEG decoding of the texture loader against the reference's stb_image 2.30: golden pixels
written by tests/golden/make_jpeg_golden.py, which encodes seeded synthetic images with Pillow
(baseline 4:4:4 / 4:2:2 / 4:2:0, progressive, gray, CMYK, restart intervals, RGB component ids,
quality 100, 1x1 and partial MCUs) and decodes them with oracle/_ref/stbharness (stb_image built
from the reference's include/stb/stb_image.h as src/scene/texturedata.cpp:30-46 uses it).
Vertical-only (4:4:0) and other sampling ratios are not covered: the encoder cannot write them."""
import json
import os

import pytest

import pt_amd
from test_png import decode

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "jpeg_expected.json")))


def _file(name):
    return open(os.path.join(HERE, "golden", "jpeg", name + ".jpg"), "rb").read()


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_jpeg_matches_stb_image(name):
    g = GOLDEN[name]
    w, h, c, px = decode(_file(name))
    assert (w, h, c) == (g["width"], g["height"], g["channels"])
    assert px == bytes.fromhex(g["pixels"])


def test_truncated_jpeg_decodes_what_it_has_or_fails_cleanly():
    data = _file("rgb420_q75")
    for cut in (len(data) // 2, len(data) - 2):  # data cut short; EOI missing
        w, h, c, px = decode(data[:cut])
        assert (w, h, c) == (33, 21, 4) and len(px) == 33 * 21 * 4
    with pytest.raises(pt_amd.PtError):
        decode(data[:100])  # inside the tables: no frame header


def test_corrupt_jpeg_header_is_an_error():
    with pytest.raises(pt_amd.PtError, match="unsupported image"):
        decode(b"\xff\xd8\xff\xe0\x00\x10JFIF\x00" + b"\x00" * 8)


def test_mutated_images_decode_or_fail_cleanly():
    """Seeded byte mutations of the JPEG fixtures and a PNG (bit flips, overwrites, cuts): each
    decodes to a consistent size or raises PtError, never crashes. The same mutator run for 30000
    inputs over an AddressSanitizer/UndefinedBehaviorSanitizer build of the decoders reports
    nothing; it found the signed overflow of corrupt coefficients that the IDCT and the
    coefficient products now compute modulo 2^32 (the bits stb_image's int arithmetic gives)."""
    import numpy as np
    import zlib
    import struct
    rng = np.random.default_rng(2024)
    raw = b"".join(b"\x00" + bytes(rng.integers(0, 256, 12 * 4, dtype=np.uint8)) for _ in range(9))
    chunk = lambda t, d: struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d))
    png = (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", 12, 9, 8, 6, 0, 0, 0)) +
           chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b""))
    seeds = [_file(n) for n in sorted(GOLDEN)] + [png]
    decoded = failed = 0
    for it in range(600):
        d = bytearray(seeds[it % len(seeds)])
        for _ in range(int(rng.integers(1, 9))):
            p = int(rng.integers(0, len(d)))
            k = int(rng.integers(0, 4))
            if k == 0:
                d[p] ^= 1 << int(rng.integers(0, 8))
            elif k == 1:
                d[p] = int(rng.integers(0, 256))
            elif k == 2 and p > 0:
                del d[p:]
            else:
                d[p] = 0xFF
        try:
            w, h, c, px = decode(bytes(d))
        except pt_amd.PtError:
            failed += 1
            continue
        assert len(px) == w * h * c and c in (1, 2, 4)
        decoded += 1
    assert decoded > 0 and failed > 0
