"""Writes tests/golden/sweep_golden.npz from the reference's FastSweep.

    python tests/golden/make_sweep_golden.py

Input (read-only, present only in the build container):
  * oracle/_ref/libdolref.so -- DistanceOcclusionLib.cpp compiled from its own
    sources by oracle/Makefile (FastSweep 2D and 3D, and the delta
    initialisations that start them)
Output (committed data only -- inputs and expected outputs):
  sweep_golden.npz  for seeded random masks: the starting deltas (unsigned
                    init for cw = 1, border init for cw = 2) and the
                    reference's FastSweep result of each
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle_bind import load_reference, pack_mask  # noqa: E402

DIMS_2D = [(5, 5, 0.2), (33, 17, 0.05), (64, 40, 0.3), (70, 9, 0.01), (130, 67, 0.002)]
DIMS_3D = [(17, 9, 6, 0.05), (33, 5, 4, 0.3), (6, 7, 8, 0.01), (40, 21, 13, 0.002)]


def main():
    R = load_reference()
    if R is None:
        sys.exit("oracle/_ref/libdolref.so is not built (make -C oracle ref)")
    rng = np.random.default_rng(4321)
    out = {}
    for k, (w, h, p) in enumerate(DIMS_2D):
        m = pack_mask(rng.random((h, w)) < p)
        out[f"s2_{k}_dims"] = np.array([w, h])
        for cw, start in ((1, R.init_deltas_2d(w, h, m)), (2, R.border_2d(w, h, m))):
            out[f"s2_{k}_cw{cw}_in"] = start
            out[f"s2_{k}_cw{cw}_out"] = R.fast_sweep_2d(w, h, start, cw)
    for k, (w, h, d, p) in enumerate(DIMS_3D):
        m = pack_mask(rng.random((d, h, w)) < p)
        out[f"s3_{k}_dims"] = np.array([w, h, d])
        for cw, start in ((1, R.init_deltas_3d(w, h, d, m)), (2, R.border_3d(w, h, d, m))):
            out[f"s3_{k}_cw{cw}_in"] = start
            out[f"s3_{k}_cw{cw}_out"] = R.fast_sweep_3d(w, h, d, start, cw)
    path = os.path.join(HERE, "sweep_golden.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
