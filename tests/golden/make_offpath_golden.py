"""Writes tests/golden/offpath_golden.npz from the reference's raster-order
methods: Danielsson 2D/3D, Chamfer (int16 and int32), InitOccFromBitMask and
ShadowSweep.

    python tests/golden/make_offpath_golden.py

Input (read-only, present only in the build container):
  * oracle/_ref/libdolref.so -- DistanceOcclusionLib.cpp compiled from its own
    sources by oracle/Makefile
Output (committed data only -- inputs and expected outputs):
  offpath_golden.npz  for seeded random masks: the starting deltas (unsigned
                      init for cw = 1, border init for cw = 2) and the
                      reference's Danielsson result; the initial distances
                      (maxD = INT16_MAX - 16, INT32_MAX - 16) and Chamfer's
                      result; the occlusion image and the shadow of several
                      light directions
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle_bind import load_reference, pack_mask, unpack_mask  # noqa: E402

DIMS_2D = [(5, 5, 0.2), (33, 17, 0.05), (64, 40, 0.3), (70, 9, 0.01), (130, 67, 0.002)]
DIMS_3D = [(17, 9, 6, 0.05), (33, 5, 4, 0.3), (6, 7, 8, 0.01), (40, 21, 13, 0.002)]
LIGHTS = [(0.3, 1.0), (-0.7, 0.5), (1.0, -0.2), (-0.4, -0.9), (0.0, 1.0), (2.0, 0.5)]


def main():
    R = load_reference()
    if R is None:
        sys.exit("oracle/_ref/libdolref.so is not built (make -C oracle ref)")
    rng = np.random.default_rng(8642)
    out = {}
    for k, (w, h, p) in enumerate(DIMS_2D):
        m = pack_mask(rng.random((h, w)) < p)
        out[f"d2_{k}_dims"] = np.array([w, h])
        out[f"d2_{k}_mask"] = m
        for cw, start in ((1, R.init_deltas_2d(w, h, m)), (2, R.border_2d(w, h, m))):
            out[f"d2_{k}_cw{cw}_in"] = start
            out[f"d2_{k}_cw{cw}_out"] = R.danielsson_2d(w, h, start, cw)
        bits = unpack_mask(m, w, h)
        for name, dt, mx in (("i16", np.int16, 32767 - 16), ("i32", np.int32, 2**31 - 1 - 16)):
            start = np.where(bits, 0, mx).astype(dt)
            res, div = R.chamfer(w, h, start)
            out[f"c_{k}_{name}_in"] = start
            out[f"c_{k}_{name}_out"] = res
            out[f"c_{k}_div"] = np.array(div)
        scale = float(rng.uniform(0.05, 1.0))
        occ = R.init_occ(w, h, m, scale)
        out[f"o_{k}_scale"] = np.array(scale, np.float32)
        out[f"o_{k}_occ"] = occ
        for j, (gx, gy) in enumerate(LIGHTS):
            out[f"o_{k}_shadow{j}"] = R.shadow_sweep(gx, gy, w, h, occ)
    out["lights"] = np.array(LIGHTS, np.float32)
    for k, (w, h, d, p) in enumerate(DIMS_3D):
        m = pack_mask(rng.random((d, h, w)) < p)
        out[f"d3_{k}_dims"] = np.array([w, h, d])
        for cw, start in ((1, R.init_deltas_3d(w, h, d, m)), (2, R.border_3d(w, h, d, m))):
            out[f"d3_{k}_cw{cw}_in"] = start
            out[f"d3_{k}_cw{cw}_out"] = R.danielsson_3d(w, h, d, start, cw)
    path = os.path.join(HERE, "offpath_golden.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
