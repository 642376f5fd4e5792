"""Writes tests/golden/voxel_golden.npz from the reference's voxeliser.

    python tests/golden/make_voxel_golden.py

Inputs (read-only, present only in the build container):
  * oracle/_ref/libmeshref.so -- MeshSupport.cpp compiled from its own source
    by oracle/Makefile (ReadObjFile, FindBounds, CreateBitMaskFromTriangles)
  * the reference's sample mesh data/plant2.obj
Outputs (committed data only -- inputs and expected outputs):
  voxel_golden.npz  the sample mesh's positions and indices, its bounds, and
                    the reference's masks at several grid sizes; a seeded
                    random triangle soup (thin, degenerate, partly outside the
                    box) with its masks
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_DIR = os.environ.get("REF_DIR", "/root/reference")
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle_bind import load_mesh_reference  # noqa: E402

PLANT_DIMS = [(64, 64, 64), (100, 80, 60), (33, 17, 40)]
SOUP_DIMS = [(40, 40, 40), (31, 50, 9)]


def triangle_soup(rng, n=400):
    """Triangles of every shape around the unit box: small, long and thin,
    degenerate (repeated or collinear vertices), some sticking out of it."""
    p = rng.uniform(-0.2, 1.2, (n, 3, 3)).astype(np.float32)
    small = rng.uniform(0, 1, (n // 2, 1, 3)).astype(np.float32) + rng.normal(0, 0.01, (n // 2, 3, 3)).astype(np.float32)
    p[: n // 2] = small
    p[n - 20:n - 10, 2] = p[n - 20:n - 10, 1]                      # two equal vertices
    t = rng.uniform(0, 1, (10, 1)).astype(np.float32)
    p[n - 10:, 2] = p[n - 10:, 0] + t * (p[n - 10:, 1] - p[n - 10:, 0])   # collinear
    return p.reshape(-1, 3), np.arange(3 * n, dtype=np.int32)


def main():
    ref = load_mesh_reference()
    if ref is None:
        sys.exit("oracle/_ref/libmeshref.so is not built (make -C oracle)")
    out = {}
    pos, idx = ref.read_obj(os.path.join(REF_DIR, "data", "plant2.obj"))
    bbox = ref.find_bounds(pos)
    out["plant_positions"], out["plant_indices"], out["plant_bbox"] = pos, idx, bbox
    for w, h, d in PLANT_DIMS:
        out[f"plant_mask_{w}_{h}_{d}"] = ref.mask_from_triangles(w, h, d, pos, idx, bbox)
    rng = np.random.default_rng(20261204)
    sp, si = triangle_soup(rng)
    sbox = np.array([0, 0, 0, 1, 1, 1], np.float32)
    out["soup_positions"], out["soup_indices"], out["soup_bbox"] = sp, si, sbox
    for w, h, d in SOUP_DIMS:
        out[f"soup_mask_{w}_{h}_{d}"] = ref.mask_from_triangles(w, h, d, sp, si, sbox)
    np.savez_compressed(os.path.join(HERE, "voxel_golden.npz"), **out)
    print("wrote", os.path.join(HERE, "voxel_golden.npz"), sorted(out))


if __name__ == "__main__":
    main()
