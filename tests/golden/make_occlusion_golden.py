"""Writes tests/golden/occlusion_golden.npz from the reference's occlusion sweeps.

    python tests/golden/make_occlusion_golden.py

Input (read-only, present only in the build container):
  * oracle/_ref/libdolref.so -- DistanceOcclusionLib.cpp compiled from its own
    sources by oracle/Makefile (InitDirWFromBitMask, OcclusionSweep)
Output (committed data only -- inputs and expected outputs):
  occlusion_golden.npz  for seeded random masks: the mask, the reference's
                        direction field and its swept field (2D: 4 planes;
                        3D: 8 planes, in place and from a second field whose
                        starting values are seeded random floats)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle_bind import load_reference, pack_mask  # noqa: E402

DIMS_2D = [(5, 5, 0.2), (33, 17, 0.05), (64, 40, 0.3), (70, 9, 0.01)]
DIMS_3D = [(17, 9, 6, 0.05), (33, 5, 4, 0.3), (6, 7, 8, 0.01)]


def main():
    R = load_reference()
    if R is None:
        sys.exit("oracle/_ref/libdolref.so is not built (make -C oracle ref)")
    rng = np.random.default_rng(777)
    out = {}
    for k, (w, h, p) in enumerate(DIMS_2D):
        m = pack_mask(rng.random((h, w)) < p)
        dw = R.init_dirw_2d(w, h, m)
        out[f"o2_{k}_dims"] = np.array([w, h])
        out[f"o2_{k}_mask"] = m
        out[f"o2_{k}_dirw"] = dw
        out[f"o2_{k}_swept"] = R.occlusion_sweep_2d(w, h, dw)
    for k, (w, h, d, p) in enumerate(DIMS_3D):
        m = pack_mask(rng.random((d, h, w)) < p)
        dw = R.init_dirw_3d(w, h, d, m)
        o0 = rng.random(dw.shape).astype(np.float32)
        out[f"o3_{k}_dims"] = np.array([w, h, d])
        out[f"o3_{k}_mask"] = m
        out[f"o3_{k}_dirw"] = dw
        out[f"o3_{k}_swept"] = R.occlusion_sweep_3d(w, h, d, dw)
        out[f"o3_{k}_out0"] = o0
        out[f"o3_{k}_from"] = R.occlusion_sweep_3d_from(w, h, d, dw, o0)
    path = os.path.join(HERE, "occlusion_golden.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
