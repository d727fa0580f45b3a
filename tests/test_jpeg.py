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
