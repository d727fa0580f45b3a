"""Writes PNG fixtures that cover the decoder's cases and the pixels the reference's vendored
stb_image 2.30 decodes from them (oracle/_ref/stbharness, built by `make -C oracle ref` from
include/stb/stb_image.h), into tests/golden/png/ and tests/golden/png_expected.json.

Cases: every colour type at every allowed bit depth, tRNS keys and palette alpha, Adam7
interlacing, all five row filters, stored / fixed / dynamic DEFLATE blocks, split IDAT.
Run from the repository root: python tests/golden/make_png_golden.py
"""
import json
import os
import struct
import subprocess
import sys
import tempfile
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "png")
STB = os.path.join(HERE, "..", "..", "oracle", "_ref", "stbharness")

SAMPLES = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}


def chunk(t, data):
    return struct.pack(">I", len(data)) + t + data + struct.pack(">I", zlib.crc32(t + data) & 0xFFFFFFFF)


def paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def pack_row(vals, depth):
    if depth == 16:
        return b"".join(struct.pack(">H", int(v)) for v in vals)
    if depth == 8:
        return bytes(int(v) for v in vals)
    out, acc, nbits = bytearray(), 0, 0
    for v in vals:
        acc = (acc << depth) | int(v)
        nbits += depth
        if nbits == 8:
            out.append(acc)
            acc, nbits = 0, 0
    if nbits:
        out.append(acc << (8 - nbits))
    return bytes(out)


def filter_rows(rows, bpp, filters):
    out, prev = bytearray(), None
    for y, row in enumerate(rows):
        f = filters[y % len(filters)]
        out.append(f)
        for i, x in enumerate(row):
            a = row[i - bpp] if i >= bpp else 0
            b = prev[i] if prev is not None else 0
            c = prev[i - bpp] if (prev is not None and i >= bpp) else 0
            pred = [0, a, b, (a + b) >> 1, paeth(a, b, c)][f]
            out.append((x - pred) & 255)
        prev = row
    return bytes(out)


ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]


def make_png(samples, color, depth, interlace=False, filters=(0,), level=6, plte=None, trns=None, split=1):
    """samples: H x W x n integer array of native sample values."""
    h, w, n = samples.shape
    bpp = max(1, n * depth // 8)
    if interlace:
        raw = b""
        for x0, y0, dx, dy in ADAM7:
            sub = samples[y0::dy, x0::dx]
            if sub.shape[0] == 0 or sub.shape[1] == 0:
                continue
            rows = [pack_row(r.reshape(-1), depth) for r in sub]
            raw += filter_rows(rows, bpp, filters)
    else:
        rows = [pack_row(r.reshape(-1), depth) for r in samples]
        raw = filter_rows(rows, bpp, filters)
    comp = zlib.compress(raw, level)
    png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, color, 0, 0, int(interlace)))
    if plte is not None:
        png += chunk(b"PLTE", bytes(plte))
    if trns is not None:
        png += chunk(b"tRNS", bytes(trns))
    step = (len(comp) + split - 1) // split
    for k in range(split):
        png += chunk(b"IDAT", comp[k * step:(k + 1) * step])
    return png + chunk(b"IEND", b"")


def cases():
    rng = np.random.default_rng(2024)
    out = {}
    for color in (0, 2, 3, 4, 6):
        depths = {0: (1, 2, 4, 8, 16), 3: (1, 2, 4, 8)}.get(color, (8, 16))
        for depth in depths:
            n = SAMPLES[color]
            w, h = 13, 7
            top = (1 << depth) - 1
            if color == 3:
                entries = min(1 << depth, 40)
                s = rng.integers(0, entries, (h, w, 1))
                plte = rng.integers(0, 256, entries * 3).tolist()
            else:
                # smooth ramps (compressible) plus noise
                s = (rng.integers(0, top + 1, (h, w, n)) + np.arange(w)[None, :, None]) % (top + 1)
                plte = None
            for interlace in (False, True):
                name = "c%d_d%d%s" % (color, depth, "_i" if interlace else "")
                out[name] = make_png(s, color, depth, interlace, filters=(0, 1, 2, 3, 4), plte=plte)
            if color in (0, 2):  # tRNS colour key: make sure some pixels match it
                key = s[2, 3]
                t = b"".join(struct.pack(">H", int(v)) for v in key)
                out["c%d_d%d_trns" % (color, depth)] = make_png(s, color, depth, filters=(4, 3), trns=t)
            if color == 3:  # partial palette alpha
                out["c3_d%d_trns" % depth] = make_png(s, color, depth, filters=(1,), plte=plte,
                                                       trns=rng.integers(0, 256, max(1, entries // 2)).tolist())
    big = (np.arange(64 * 48 * 4).reshape(48, 64, 4) * 7 % 251).astype(np.int64)
    out["rgba_stored"] = make_png(big, 6, 8, level=0)
    out["rgba_dynamic_split"] = make_png(big, 6, 8, filters=(2, 4, 1), level=9, split=3)
    tiny = np.full((1, 1, 3), 200)
    out["rgb_1x1_fixed"] = make_png(tiny, 2, 8, level=1)
    return out


def main():
    if not os.path.exists(STB):
        sys.exit("build oracle/_ref/stbharness first: make -C oracle ref")
    os.makedirs(OUT, exist_ok=True)
    expected = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, data in sorted(cases().items()):
            path = os.path.join(OUT, name + ".png")
            with open(path, "wb") as f:
                f.write(data)
            dec = os.path.join(tmp, name + ".bin")
            subprocess.check_call([STB, path, dec])
            raw = open(dec, "rb").read()
            w, h, c = struct.unpack("<III", raw[:12])
            expected[name] = {"width": w, "height": h, "channels": c, "pixels": raw[12:].hex()}
    with open(os.path.join(HERE, "png_expected.json"), "w") as f:
        json.dump(expected, f, indent=0, sort_keys=True)
    print("wrote %d fixtures" % len(expected))


if __name__ == "__main__":
    main()
