"""Writes tests/golden/image_golden.npz: the image files the reference's own
tool writes, decoded to pixel arrays, for the oracle's image writers
(oc_border_image, oc_distance_u8, oc_scalar_u8_*) to be checked against.

The tool is oracle/_ref/distocc_ref, built by `make -C oracle tool` from the
reference's DistanceOcclusionLib.cpp and DistanceOcclusionTool.cpp where they
lie; its masks come from its generators (-b block, then -p points), which the
oracle restates.  Each case stores its arguments and the decoded images (one
per slice for volumes).  Run from the repository root:

    python tests/golden/make_image_golden.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
TOOL = os.path.join(ROOT, "oracle", "_ref", "distocc_ref")

# (key, tool arguments, dims, sides, points, seed, method, image type, maxLen override)
CASES = [
    ("s2_t0", ["-s", "-m", "2", "-w", "70", "52", "-b", "3", "-p", "20", "5"], (70, 52), 3, 20, 5, 2, 0, None),
    ("s2_t1", ["-s", "-m", "2", "-i", "1", "-w", "70", "52", "-b", "3", "-p", "20", "5"], (70, 52), 3, 20, 5, 2, 1, 127),
    ("s2_t2", ["-s", "-m", "2", "-i", "2", "-w", "70", "52", "-b", "3", "-p", "20", "5"], (70, 52), 3, 20, 5, 2, 2, None),
    ("s2_t3", ["-s", "-m", "2", "-i", "3", "-w", "70", "52", "-b", "3", "-p", "20", "5"], (70, 52), 3, 20, 5, 2, 3, None),
    ("s3_t0", ["-s", "-m", "2", "-W", "22", "18", "9", "-b", "3", "-p", "12", "7"], (22, 18, 9), 3, 12, 7, 2, 0, None),
    ("s3_t1", ["-s", "-m", "2", "-i", "1", "-W", "22", "18", "9", "-b", "3", "-p", "12", "7"], (22, 18, 9), 3, 12, 7, 2, 1, 127),
    ("s3_t2", ["-s", "-m", "2", "-i", "2", "-W", "22", "18", "9", "-b", "3", "-p", "12", "7"], (22, 18, 9), 3, 12, 7, 2, 2, None),
    ("s3_t3", ["-s", "-m", "2", "-i", "3", "-W", "22", "18", "9", "-b", "3", "-p", "12", "7"], (22, 18, 9), 3, 12, 7, 2, 3, None),
    ("d2_jfa", ["-d", "-m", "2", "-w", "66", "40", "-p", "25", "9"], (66, 40), 0, 25, 9, 2, 0, None),
    ("d3_jfa", ["-d", "-m", "2", "-W", "20", "17", "11", "-p", "15", "3"], (20, 17, 11), 0, 15, 3, 2, 0, None),
    ("d2_chamfer", ["-d", "-m", "4", "-w", "66", "40", "-p", "25", "9"], (66, 40), 0, 25, 9, 4, 0, None),
    ("d2_felz", ["-d", "-m", "5", "-w", "66", "40", "-p", "25", "9"], (66, 40), 0, 25, 9, 5, 0, None),
]


def read_png(path):
    from PIL import Image

    im = Image.open(path)
    return np.asarray(im.convert("L") if im.mode in ("L", "P") else im.convert("RGBA"))


def main():
    if not os.path.exists(TOOL):
        sys.exit("oracle/_ref/distocc_ref is not built (make -C oracle tool)")
    out = {}
    for key, args, dims, sides, points, seed, method, itype, ml in CASES:
        with tempfile.TemporaryDirectory() as d:
            subprocess.run([TOOL] + args + ["-o", "g"], cwd=d, check=True, capture_output=True)
            files = sorted(f for f in os.listdir(d) if f.endswith(".png"))
            imgs = np.stack([read_png(os.path.join(d, f)) for f in files])
        out[f"{key}_dims"] = np.array(dims, np.int32)
        out[f"{key}_gen"] = np.array([sides, points, seed, method, itype, -1 if ml is None else ml], np.int32)
        out[f"{key}_img"] = imgs
        print(key, files[:2], imgs.shape)
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "image_golden.npz"), **out)


if __name__ == "__main__":
    main()
