"""Writes the golden fixtures of tests/golden/ from the reference.

    python tests/golden/make_golden.py

Inputs (read-only, present only in the build container):
  * oracle/_ref/libdolref.so -- DistanceOcclusionLib.cpp compiled from its own
    sources by oracle/Makefile; produces the reference vectors
  * the reference's sample images data/{tiger,kiwi}.png and its published
    signed outputs images/{tiger,kiwi}-sdf*.png
Outputs (committed data only -- inputs and expected outputs):
  reference_vectors.npz  small full arrays + SHA-256 digests of large ones
  samples.npz            grey sample images and the published outputs
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_DIR = os.environ.get("REF_DIR", "/root/reference")
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle_bind import empty_mask, load_reference, pack_mask  # noqa: E402


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def grey(path):
    """stbi_load(..., 1): RGB(A) -> (77 r + 150 g + 29 b) >> 8, alpha dropped
    (stb_image_mini.h:665-668, 702)."""
    from PIL import Image

    im = np.asarray(Image.open(path).convert("RGBA")).astype(np.int32)
    return ((im[..., 0] * 77 + im[..., 1] * 150 + im[..., 2] * 29) >> 8).astype(np.uint8)


def rgba(path):
    from PIL import Image

    return np.asarray(Image.open(path).convert("RGBA"))


def gray_png(path):
    from PIL import Image

    return np.asarray(Image.open(path).convert("L"))


def main():
    R = load_reference()
    if R is None:
        sys.exit("oracle/_ref/libdolref.so is not built (make -C oracle ref)")
    rng = np.random.default_rng(1234)
    out = {}

    # config 1 of the tool: -p 100 on 512 x 512, seed 12345 (DistanceOcclusionTool.cpp:1762)
    w = h = 512
    m = empty_mask(w, h)
    R.mask_add_points_2d(w, h, 100, 12345, m)
    out["cfg1_mask"] = m.copy()
    start = R.init_deltas_2d(w, h, m)
    out["cfg1_jfa_cw1_sha256"] = np.array(digest(R.jfa_2d(w, h, start, 1)))
    f = R.felzenszwalb_2d(w, h, R.init_distances_f32(w, h, m))
    out["cfg1_felz_sha256"] = np.array(digest(f))
    # the tool's -b 4 block plus points (DistanceOcclusionTool.cpp:1921-1924)
    R.mask_add_block_2d(w, h, 4, m)
    out["cfg1b_mask_sha256"] = np.array(digest(m))

    # small full vectors, 2D
    for k, (w, h, p) in enumerate([(5, 5, 0.2), (4, 3, 0.0), (33, 17, 0.1), (64, 40, 0.5), (70, 9, 0.02)]):
        bits = rng.random((h, w)) < p
        m = pack_mask(bits)
        out[f"d2_{k}_dims"] = np.array([w, h])
        out[f"d2_{k}_mask"] = m
        b = R.border_2d(w, h, m)
        out[f"d2_{k}_border"] = b
        out[f"d2_{k}_jfa_cw2"] = R.jfa_2d(w, h, b, 2)
        i0 = R.init_deltas_2d(w, h, m)
        out[f"d2_{k}_jfa_cw1"] = R.jfa_2d(w, h, i0, 1)
        out[f"d2_{k}_brute_cw1"] = R.brute_force_2d(w, h, i0, 1)
        out[f"d2_{k}_felz"] = R.felzenszwalb_2d(w, h, R.init_distances_f32(w, h, m))

    # small full vectors, 3D
    for k, (w, h, d, p) in enumerate([(17, 9, 6, 0.1), (33, 5, 4, 0.5), (6, 7, 8, 0.02), (3, 3, 3, 0.0)]):
        bits = rng.random((d, h, w)) < p
        m = pack_mask(bits)
        out[f"d3_{k}_dims"] = np.array([w, h, d])
        out[f"d3_{k}_mask"] = m
        b = R.border_3d(w, h, d, m)
        out[f"d3_{k}_border"] = b
        out[f"d3_{k}_jfa_cw2"] = R.jfa_3d(w, h, d, b, 2)
        i0 = R.init_deltas_3d(w, h, d, m)
        out[f"d3_{k}_jfa_cw1"] = R.jfa_3d(w, h, d, i0, 1)
        out[f"d3_{k}_brute_cw1"] = R.brute_force_3d(w, h, d, i0, 1)

    # float Felzenszwalb beyond 4096, where float is no longer exact
    # (odd squares above 2^24 round, so (float)(q*q) differs from q*q)
    w, h = 4300, 3
    bits = np.zeros((h, w), bool)
    bits[0, 0] = True
    m = pack_mask(bits)
    out["felz_wide_dims"] = np.array([w, h])
    out["felz_wide_mask"] = m
    out["felz_wide"] = R.felzenszwalb_2d(w, h, R.init_distances_f32(w, h, m))

    # the reference's samples: signed border pipeline deltas (border init +
    # JumpFlood in half cells), tool thresholds t = 128, lessThan (dark = inside)
    samples = {}
    for name in ("tiger", "kiwi"):
        g = grey(os.path.join(REF_DIR, "data", f"{name}.png"))
        h, w = g.shape
        bits = g < 128
        m = pack_mask(bits)
        b = R.border_2d(w, h, m)
        samples[f"{name}_grey"] = g
        out[f"{name}_border_jfa_cw2"] = R.jfa_2d(w, h, b, 2)
        samples[f"{name}_sdf"] = gray_png(os.path.join(REF_DIR, "images", f"{name}-sdf.png"))
        samples[f"{name}_sdf_delta"] = rgba(os.path.join(REF_DIR, "images", f"{name}-sdf-delta.png"))
        samples[f"{name}_sdf_rg"] = rgba(os.path.join(REF_DIR, "images", f"{name}-sdf-rg.png"))

    np.savez_compressed(os.path.join(HERE, "reference_vectors.npz"), **out)
    np.savez_compressed(os.path.join(HERE, "samples.npz"), **samples)
    for f in ("reference_vectors.npz", "samples.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
