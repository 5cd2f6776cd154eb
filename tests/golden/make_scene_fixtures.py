"""Writes tests/golden/scenes/envmap.pfm: an 8x4 lat-long test image whose
rows are distinguishable (row r of the image, counted from the top, has
red = r + 1, green = column + 1, blue = 0.5).  PFM stores rows bottom-up."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    h, w = 4, 8
    img = np.zeros((h, w, 3), np.float32)
    for r in range(h):
        img[r, :, 0] = r + 1
        img[r, :, 1] = np.arange(w) + 1
        img[r, :, 2] = 0.5
    with open(os.path.join(HERE, "scenes", "envmap.pfm"), "wb") as f:
        f.write(b"PF\n%d %d\n-1.0\n" % (w, h))
        f.write(img[::-1].astype("<f4").tobytes())


if __name__ == "__main__":
    main()
