"""Write plain_small_truth.npz: the FP64 high-spp image the RMSE check uses.

The truth is the oracle's estimator (src/renderer/cpu.rs:233-342) run in FP64
at 16384 samples per pixel on a stratified 1/16 pixel subset (every 4th row
and column) of the 256x256 small plain scene, with a seed the tests never
use for their estimates.  Run from the repository root:

    python tests/golden/make_truth.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from raydar_amd import scenegen  # noqa: E402

TRUTH_SPP = 16384
TEST_SPP = 16
SEED = 987654321


def main():
    sc = scenegen.plain_small(0)
    ys, xs = np.meshgrid(np.arange(0, 256, 4), np.arange(0, 256, 4), indexing="ij")
    pix = (ys * 256 + xs).reshape(-1).astype(np.uint32)
    acc = oracle.render(sc, TRUTH_SPP, 12, SEED, pixels=pix, fp64=True)
    truth = acc[:, :3] / TRUTH_SPP
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "plain_small_truth.npz")
    np.savez_compressed(out, pixels=pix, truth=truth, truth_spp=TRUTH_SPP, test_spp=TEST_SPP,
                        seed=SEED)
    print(out, truth.shape, truth.mean(axis=0))


if __name__ == "__main__":
    main()
