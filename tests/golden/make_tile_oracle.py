"""Writes tests/golden/tile_oracle_<name>.npz: the oracle rendered at 65,536
samples per pixel as 16 independent batches of 4,096 samples (seeds 50000 +
batch, disjoint from every seed the GPU tests use), kept as per-batch 16x16
tile means and the overall mean image.  The GPU tile-bias test compares its
own independent batches against these (tests/test_image_parity_gpu.py).

Scenes (drawlab_amd.scenes, small resolutions so the CPU finishes):
  cornell_box       32x32, max depth 8, Russian roulette from 3
  displaced_sphere  32x16, 128x128-grid sphere, level-3 glass, 256x128 sky,
                    max depth 16 (the 2M-triangle workload's materials and
                    lights at a size the CPU renders)

Run once in the build container:  python tests/golden/make_tile_oracle.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from drawlab_amd import scenes  # noqa: E402

BATCHES, SPP, SEED0 = 16, 4096, 50000


def configs():
    return {
        "cornell_box": scenes.cornell_box(32, 32, spp=SPP, max_depth=8, rr_depth=3),
        "displaced_sphere": scenes.displaced_sphere(width=32, height=16, grid=128, glass_level=3,
                                                    env_size=(256, 128), max_depth=16),
    }


def tile_means(img, t=16):
    H, W, _ = img.shape
    return img[:H - H % t, :W - W % t].reshape(H // t, t, W // t, t, 3).mean(axis=(1, 3))


def main(names=None):
    for name, desc in configs().items():
        if names and name not in names:
            continue
        osc = oracle.from_desc(desc)
        tiles, acc = [], None
        t0 = time.time()
        for b in range(BATCHES):
            img, st = osc.render(SPP, seed=SEED0 + b, max_depth=desc.max_depth, rr_depth=desc.rr_depth)
            tiles.append(tile_means(img))
            acc = img.astype(np.float64) if acc is None else acc + img
            print("%s batch %d/%d  %.0f s" % (name, b + 1, BATCHES, time.time() - t0), flush=True)
        np.savez_compressed(os.path.join(HERE, "tile_oracle_%s.npz" % name), tiles=np.stack(tiles).astype(np.float64),
                            image=(acc / BATCHES).astype(np.float32), spp=SPP, batches=BATCHES, seed0=SEED0)


if __name__ == "__main__":
    main(sys.argv[1:])
