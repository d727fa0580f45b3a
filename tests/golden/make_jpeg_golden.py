"""Writes JPEG fixtures that cover the texture decoder's cases and the pixels the reference's
vendored stb_image 2.30 decodes from them (oracle/_ref/stbharness, built by `make -C oracle ref`
from include/stb/stb_image.h, the way src/scene/texturedata.cpp:30-46 loads a texture), into
tests/golden/jpeg/ and tests/golden/jpeg_expected.json.

The files are encoded by Pillow from seeded synthetic images: baseline 4:4:4, 4:2:2, 4:2:0,
progressive (with and without chroma subsampling, optimised Huffman tables), gray, CMYK
(Adobe transform 0), restart intervals, RGB component ids, quality 100 (large coefficients),
1x1 and partial-MCU sizes.
Run from the repository root: python tests/golden/make_jpeg_golden.py
"""
import io
import json
import os
import struct
import subprocess
import tempfile

import numpy as np
from PIL import Image

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "jpeg")
STB = os.path.join(HERE, "..", "..", "oracle", "_ref", "stbharness")


def synthetic(w, h, ch, seed):
    r = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    planes = []
    for c in range(ch):
        f = r.uniform(0.05, 0.6, 2)
        p = 128 + 90 * np.sin(x * f[0] + c) * np.cos(y * f[1] - c)
        p += r.normal(0, 18, (h, w))                 # texture noise
        p[(x + 2 * y + 3 * c) % 11 < 2] = 250 - 40 * c  # hard edges
        planes.append(np.clip(p, 0, 255))
    return np.stack(planes, -1).astype(np.uint8)


CASES = {
    # name: (width, height, mode, save options)
    "rgb444_q90": (33, 21, "RGB", dict(quality=90, subsampling=0)),
    "rgb422_q80": (33, 21, "RGB", dict(quality=80, subsampling=1)),
    "rgb420_q75": (33, 21, "RGB", dict(quality=75, subsampling=2)),
    "rgb420_q100": (19, 27, "RGB", dict(quality=100, subsampling=2)),
    "rgb420_prog": (33, 21, "RGB", dict(quality=85, subsampling=2, progressive=True)),
    "rgb444_prog_opt": (24, 17, "RGB", dict(quality=92, subsampling=0, progressive=True, optimize=True)),
    "rgb420_restart": (40, 24, "RGB", dict(quality=80, subsampling=2, restart_marker_blocks=3)),
    "rgb_ids": (21, 13, "RGB", dict(quality=90, keep_rgb=True)),
    "gray": (29, 19, "L", dict(quality=85)),
    "gray_prog": (29, 19, "L", dict(quality=70, progressive=True)),
    "cmyk": (17, 15, "CMYK", dict(quality=90)),
    "rgb_1x1": (1, 1, "RGB", dict(quality=90, subsampling=2)),
    "rgb420_9x3": (9, 3, "RGB", dict(quality=60, subsampling=2)),
}


def stb_decode(path):
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "out.bin")
        subprocess.run([STB, path, out], check=True)
        raw = open(out, "rb").read()
    w, h, c = struct.unpack("<3I", raw[:12])
    return w, h, c, raw[12:]


def main():
    os.makedirs(OUT, exist_ok=True)
    expected = {}
    for i, (name, (w, h, mode, opts)) in enumerate(sorted(CASES.items())):
        ch = {"RGB": 3, "L": 1, "CMYK": 4}[mode]
        px = synthetic(w, h, ch, 100 + i)
        img = Image.fromarray(px[..., 0] if ch == 1 else px, mode)
        buf = io.BytesIO()
        img.save(buf, "JPEG", **opts)
        path = os.path.join(OUT, name + ".jpg")
        open(path, "wb").write(buf.getvalue())
        sw, sh, sc, pixels = stb_decode(path)
        expected[name] = {"width": sw, "height": sh, "channels": sc, "pixels": pixels.hex()}
        print(name, sw, sh, sc, len(buf.getvalue()), "bytes")
    with open(os.path.join(HERE, "jpeg_expected.json"), "w") as f:
        json.dump(expected, f, indent=0, sort_keys=True)


if __name__ == "__main__":
    main()
