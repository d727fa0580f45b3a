"""PNG decoding of the texture loader against the reference's stb_image 2.30 (golden pixels
written by tests/golden/make_png_golden.py with oracle/_ref/stbharness), plus DEFLATE against
Python's zlib."""
import ctypes as C
import json
import os
import zlib

import numpy as np
import pytest

import pt_amd

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "png_expected.json")))


def decode(data):
    lib = pt_amd.host_lib()
    w, h, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
    px = C.POINTER(C.c_uint8)()
    buf = C.create_string_buffer(data, len(data))
    code = lib.pth_decode_image(buf, len(data), C.byref(w), C.byref(h), C.byref(c), C.byref(px))
    if code != 0:
        raise pt_amd.PtError(code, lib.pth_last_error().decode())
    out = bytes(np.ctypeslib.as_array(px, (w.value * h.value * c.value,)))
    lib.pth_free(px)
    return w.value, h.value, c.value, out


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_png_matches_stb_image(name):
    g = GOLDEN[name]
    data = open(os.path.join(HERE, "golden", "png", name + ".png"), "rb").read()
    w, h, c, px = decode(data)
    assert (w, h, c) == (g["width"], g["height"], g["channels"])
    assert px == bytes.fromhex(g["pixels"])


def test_corrupt_png_is_an_error():
    data = open(os.path.join(HERE, "golden", "png", "c6_d8.png"), "rb").read()
    with pytest.raises(pt_amd.PtError):
        decode(data[:40])
    with pytest.raises(pt_amd.PtError, match="unsupported image"):
        decode(b"\xff\xd8\xff\xe0 not a png")
