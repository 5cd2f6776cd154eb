"""Writes the image fixtures under tests/golden/scenes:

envmap.pfm   8x4 lat-long test image whose rows are distinguishable (row r
             of the image, counted from the top, has red = r + 1, green =
             column + 1, blue = 0.5).  PFM stores rows bottom-up.
checker.png  8x8 RGB, 2x2-texel checker of two colours plus a ramp (texel
             (x, y), y counted from the top row of the file: CHECKER[y, x]);
             rows filtered with None, Sub, Up, Average and Paeth in turn.
pd.png       4x4 palette image (indices PD_INDEX into PD_PALETTE).
normal.png   4x4 RGB tangent-space normal map: NORMAL_TEXELS, (n + 1) / 2
             quantised to 8 bits.
axay.hdr     2x2 Radiance RGBE (flat), AXAY texels, rows top-down.

textured.xml (written by hand) renders them on a floor, a back wall and a
side wall."""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from pngwrite import png_bytes  # noqa: E402

SCENES = os.path.join(HERE, "scenes")


def checker():
    img = np.zeros((8, 8, 3), np.int64)
    for y in range(8):
        for x in range(8):
            a = ((x // 2) + (y // 2)) % 2
            img[y, x] = (200 + 5 * x, 40 + 3 * y, 30) if a else (30, 60 + 4 * x, 180 + 6 * y)
    return img


CHECKER = checker()
PD_PALETTE = np.array([(230, 180, 40), (40, 140, 220), (250, 250, 250)], np.int64)
PD_INDEX = np.array([[0, 1, 2, 0], [1, 2, 0, 1], [2, 0, 1, 2], [0, 1, 2, 2]], np.int64)


def normal_texels():
    n = np.zeros((4, 4, 3), np.float64)
    for y in range(4):
        for x in range(4):
            v = np.array([0.35 * (x - 1.5), 0.3 * (y - 1.5), 1.0])
            n[y, x] = v / np.linalg.norm(v)
    return np.rint((n + 1.0) * 0.5 * 255.0).astype(np.int64)


NORMAL_TEXELS = normal_texels()
AXAY = np.array([[(0.05, 0.6, 0.0), (0.3, 0.3, 0.0)], [(0.6, 0.08, 0.0), (0.15, 0.45, 0.0)]], np.float64)


def rgbe(r, g, b):
    m = max(r, g, b)
    if m < 1e-32:
        return (0, 0, 0, 0)
    e = math.frexp(m)[1]
    f = 256.0 / 2 ** e
    return (int(r * f), int(g * f), int(b * f), e + 128)


def main():
    h, w = 4, 8
    img = np.zeros((h, w, 3), np.float32)
    for r in range(h):
        img[r, :, 0] = r + 1
        img[r, :, 1] = np.arange(w) + 1
        img[r, :, 2] = 0.5
    with open(os.path.join(SCENES, "envmap.pfm"), "wb") as f:
        f.write(b"PF\n%d %d\n-1.0\n" % (w, h))
        f.write(img[::-1].astype("<f4").tobytes())
    with open(os.path.join(SCENES, "checker.png"), "wb") as f:
        f.write(png_bytes(CHECKER, 2))
    with open(os.path.join(SCENES, "pd.png"), "wb") as f:
        f.write(png_bytes(PD_INDEX, 3, palette=PD_PALETTE, filters=(0,)))
    with open(os.path.join(SCENES, "normal.png"), "wb") as f:
        f.write(png_bytes(NORMAL_TEXELS, 2, filters=(4, 1)))
    with open(os.path.join(SCENES, "axay.hdr"), "wb") as f:
        f.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 2 +X 2\n")
        f.write(b"".join(bytes(rgbe(*c)) for row in AXAY for c in row))


if __name__ == "__main__":
    main()
