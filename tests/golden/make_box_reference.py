"""Writes tests/golden/box256_4096spp.npz: the oracle's 4096-spp render of
the plain PT box workload (256 x 256, depth 8, roulette after depth 3 with
probability 0.5, seed 9000) in both estimators, stored as float16 (relative
rounding 2^-11, far below the 16-spp noise the parity test measures).

Run from the repository root: python tests/golden/make_box_reference.py
(about 3 minutes on 8 cores)."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import Oracle  # noqa: E402
from renderer_amd import scenes  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "box256_4096spp.npz")
SEED = 9000


def main():
    sc = scenes.box_scene(256, 256)
    orc = Oracle(sc.flatten(), accel=1)
    films = {}
    for name, mode in (("compat", 0), ("unbiased", 1)):
        img, rays = orc.render((sc.camera, 64, 64), mode=mode, max_depth=8, rr_start=3, rr_prob=0.5, seed=SEED,
                               threads=os.cpu_count())
        films[name] = img.astype(np.float16)
        print(name, img.mean(axis=(0, 1)), rays, flush=True)
    np.savez_compressed(OUT, seed=np.int64(SEED), **films)
    print("wrote", OUT, os.path.getsize(OUT))


if __name__ == "__main__":
    main()
