"""The PNG output path: the oracle's ACES tone mapping (tonemap.cs.hlsl, the settings of
src/pathtrace/ptdispatch.cpp:29, write_png's 8-bit conversion) pinned by values the shader's
arithmetic fixes, and the PNG writer checked by three independent decoders (Python's zlib on
the DEFLATE stream with the chunk CRCs, this package's decoder, and the reference's vendored
stb_image 2.30 when oracle/_ref is built)."""
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

import pt_amd
from oracle_bind import tonemap
from test_png import decode

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STB = os.path.join(ROOT, "oracle", "_ref", "stbharness")


def _fit(v):
    a = v * (v + 0.0245786) - 0.000090537
    b = v * (0.983729 * v + 0.4329510) + 0.238081
    return a / b


def test_tonemap_black_and_white():
    img = np.array([[[0, 0, 0, 1], [1, 1, 1, 1]]], np.float32)
    out = tonemap(img)
    assert out[0, 0].tolist() == [0, 0, 0, 255]
    # to_srgb(1) = 1, both matrices' rows sum to 1 (the odt's third to 0.99999), so white is
    # the fitted curve at 1: 1.0244881 / 1.654761 = 0.61912 -> 157.9 -> 157
    assert abs(_fit(1.0) - 0.619116) < 1e-5
    assert out[0, 1].tolist() == [157, 157, 157, 255]


def test_tonemap_saturates_and_passes_alpha():
    img = np.array([[[1e6, 50.0, 1e-9, 0.5], [-1.0, np.nan, 0.0, 1.0]]], np.float32)
    out = tonemap(img)
    assert out[0, 0, 3] == 127  # (uchar)(0.5 * 255)
    assert out[0, 1].tolist()[:3] == [0, 0, 0] and out[0, 1, 3] == 255  # saturate(NaN) = 0
    big = tonemap(np.full((1, 1, 4), 1e6, np.float32))[0, 0]
    fit_inf = 1.0 / 0.983729  # the fitted curve's limit
    assert big.tolist()[:3] == [255, 255, 255] and fit_inf > 1.0


def test_tonemap_keeps_the_shaders_srgb_threshold():
    # tonemap.cs.hlsl:30-32 switches to the power curve above 0.00031308 (ten times below the
    # sRGB standard's 0.0031308); for x = 0.0005 that gives a negative value where the
    # standard curve gives 12.92 x, which shows in the red channel of (0.0005, 1, 1)
    x = 0.0005
    shader = 1.055 * x ** (1 / 2.4) - 0.055
    standard = 12.92 * x
    assert shader < 0 < standard

    def red(s0):
        r0 = 0.59719 * s0 + 0.35458 + 0.04823
        r1 = 0.07600 * s0 + 0.90834 + 0.01566
        r2 = 0.02840 * s0 + 0.13383 + 0.83777
        o = 1.60475 * _fit(r0) - 0.53108 * _fit(r1) - 0.07367 * _fit(r2)
        return int(min(max(o, 0.0), 1.0) * 255)

    out = tonemap(np.array([[[x, 1, 1, 1]]], np.float32))[0, 0]
    assert red(shader) != red(standard)
    assert abs(int(out[0]) - red(shader)) <= 1


def test_tonemap_is_monotone_and_gray_stays_gray():
    v = np.linspace(0, 4, 2001, dtype=np.float32)
    img = np.stack([v, v, v, np.ones_like(v)], -1)[None]
    out = tonemap(img)[0].astype(int)
    assert np.all(np.diff(out[:, 0]) >= 0)
    assert np.all(np.abs(out[:, 0] - out[:, 2]) <= 1) and np.all(out[:, 0] == out[:, 1])


def _chunks(data):
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, out = 8, []
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        kind, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(kind + body) & 0xFFFFFFFF
        out.append((kind, body))
        pos += 12 + n
    return out


def _unfilter(raw, w, h):
    rb, img, prev = w * 4, [], bytearray(w * 4)
    for y in range(h):
        f, row = raw[y * (rb + 1)], bytearray(raw[y * (rb + 1) + 1:(y + 1) * (rb + 1)])
        for i in range(rb):
            a = row[i - 4] if i >= 4 else 0
            b, c = prev[i], (prev[i - 4] if i >= 4 else 0)
            p = a + b - c
            pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
            pae = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
            row[i] = (row[i] + [0, a, b, (a + b) >> 1, pae][f]) & 255
        img.append(bytes(row))
        prev = row
    return b"".join(img)


IMAGES = {
    "noise_13x7": lambda r: r.integers(0, 256, (7, 13, 4)),
    "gradient_64x48": lambda r: (np.arange(64 * 48 * 4).reshape(48, 64, 4) * 3 % 256),
    "flat_40x30": lambda r: np.full((30, 40, 4), 77),
    "stripes_257x3": lambda r: np.tile(r.integers(0, 256, (1, 257, 4)), (3, 1, 1)),
    "one_pixel": lambda r: r.integers(0, 256, (1, 1, 4)),
}


@pytest.mark.parametrize("name", sorted(IMAGES))
def test_png_writer_round_trip(tmp_path, name):
    img = IMAGES[name](np.random.default_rng(7)).astype(np.uint8)
    h, w = img.shape[:2]
    path = str(tmp_path / (name + ".png"))
    pt_amd.write_png(path, img)
    data = open(path, "rb").read()
    chunks = _chunks(data)
    assert [k for k, _ in chunks] == [b"IHDR", b"IDAT", b"IEND"]
    assert struct.unpack(">IIBBBBB", chunks[0][1]) == (w, h, 8, 6, 0, 0, 0)
    raw = zlib.decompress(chunks[1][1])  # an independent inflater
    assert _unfilter(raw, w, h) == img.tobytes()
    dw, dh, dc, px = decode(data)
    assert (dw, dh, dc) == (w, h, 4) and px == img.tobytes()
    if os.path.exists(STB):
        out = str(tmp_path / (name + ".bin"))
        subprocess.check_call([STB, path, out])
        raw_stb = open(out, "rb").read()
        assert struct.unpack("<III", raw_stb[:12]) == (w, h, 4) and raw_stb[12:] == img.tobytes()


def test_png_writer_compresses(tmp_path):
    img = np.full((256, 256, 4), 200, np.uint8)
    path = str(tmp_path / "flat.png")
    pt_amd.write_png(path, img)
    assert os.path.getsize(path) < 4000  # 256 KiB of pixels in a few long matches


def test_png_writer_errors(tmp_path):
    with pytest.raises(pt_amd.PtError):
        pt_amd.write_png(str(tmp_path / "no_such_dir" / "x.png"), np.zeros((2, 2, 4), np.uint8))
