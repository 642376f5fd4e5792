"""distocc_hip end to end: the tool's distance flows on the GPU library (the
algorithms through DOL::, masks from bytes and the images through the C ABI).
Every image equals the oracle's for the same mask; masks come from the
tool's generators (-b block, then -p points, default seed 12345) or from raw
bytes (-f, with -t / -r)."""
import os
import subprocess

import numpy as np
import pytest

from oracle_bind import ROOT, empty_mask, unpack_mask

pytestmark = pytest.mark.gpu

CLI = os.path.join(ROOT, "distance-occlusion_amd", "cli", "distocc_hip")


def run(tmp_path, *args):
    out = tmp_path / "out.bin"
    if out.exists():
        out.unlink()
    r = subprocess.run([CLI, *args, "-o", str(out)], capture_output=True, text=True, timeout=120)
    return r, out


def gen_mask(oracle, dims, points, sides=0, seed=12345):
    m = empty_mask(*dims)
    add = (oracle.mask_add_block_2d, oracle.mask_add_points_2d) if len(dims) == 2 else \
        (oracle.mask_add_block_3d, oracle.mask_add_points_3d)
    if sides:
        add[0](*dims, sides, m)
    if points:
        add[1](*dims, points, seed, m)
    return m


@pytest.mark.parametrize("method", [0, 1, 2, 3])
def test_cli_distance_delta_methods(gpu, oracle, tmp_path, method):
    """-d with Danielsson, FastSweep, JumpFlood, BruteForce: EncodeU8(Dist / maxLen)"""
    w, h = 160, 96
    r, out = run(tmp_path, "-d", "-m", str(method), "-w", str(w), str(h), "-p", "40", "-c")
    assert r.returncode == 0, r.stderr
    assert "Max error" in r.stderr
    m = gen_mask(oracle, (w, h), 40)
    i0 = oracle.init_deltas_2d(w, h, m)
    delta = {0: oracle.danielsson_2d, 1: oracle.fast_sweep_2d, 2: oracle.jfa_2d, 3: oracle.brute_force_2d}[method](
        w, h, i0, 1)
    np.testing.assert_array_equal(np.fromfile(out, np.uint8), oracle.distance_u8(2, delta, 0.25 * w))


def test_cli_distance_chamfer_felzenszwalb(gpu, oracle, tmp_path):
    w = 300
    m = gen_mask(oracle, (w, w), 40, 5)
    r, out = run(tmp_path, "-d", "-m", "5", "-w", str(w), "-b", "5", "-p", "40", "-c")
    assert r.returncode == 0, r.stderr
    f = oracle.sqrt_f32(oracle.felzenszwalb_2d(w, w, oracle.init_distances_f32(w, w, m)))
    np.testing.assert_array_equal(np.fromfile(out, np.uint8), oracle.scalar_u8(f, 0.25 * w))
    r, out = run(tmp_path, "-d", "-m", "4", "-w", str(w), "-b", "5", "-p", "40")
    assert r.returncode == 0, r.stderr
    start = np.where(unpack_mask(m, w, w), 0, 32767 - 16).astype(np.int16)
    cd, cm = oracle.chamfer(w, w, start)
    np.testing.assert_array_equal(np.fromfile(out, np.uint8), oracle.scalar_u8(cd, 0.25 * w * cm))


@pytest.mark.parametrize("image_type", [0, 1, 2, 3])
def test_cli_signed_border_images(gpu, oracle, tmp_path, image_type):
    """-s -i 0..3, 2D (not square) and 3D, with the border check (-c)"""
    w, h = 200, 130
    r, out = run(tmp_path, "-s", "-i", str(image_type), "-w", str(w), str(h), "-b", "4", "-p", "30", "-c")
    assert r.returncode == 0, r.stderr
    m = gen_mask(oracle, (w, h), 30, 4)
    ml = 127.0 if image_type == 1 else 0.25 * w
    d = oracle.jfa_2d(w, h, oracle.border_2d(w, h, m), 2)
    exp = oracle.border_image(w, h, 0, m, d, 2 * ml, image_type)
    np.testing.assert_array_equal(np.fromfile(out, np.uint8).reshape(exp.shape), exp)
    w, h, dd = 36, 28, 20
    r, out = run(tmp_path, "-s", "-i", str(image_type), "-W", str(w), str(h), str(dd), "-b", "7", "-p", "10", "-c")
    assert r.returncode == 0, r.stderr
    m = gen_mask(oracle, (w, h, dd), 10, 7)
    ml = 127.0 if image_type == 1 else 0.25 * w
    d = oracle.jfa_3d(w, h, dd, oracle.border_3d(w, h, dd, m), 2)
    exp = oracle.border_image(w, h, dd, m, d, 2 * ml, image_type)
    np.testing.assert_array_equal(np.fromfile(out, np.uint8).reshape(exp.shape), exp)


@pytest.mark.parametrize("method", [0, 1, 3])
def test_cli_signed_other_methods(gpu, oracle, tmp_path, method):
    w = 120
    r, out = run(tmp_path, "-s", "-m", str(method), "-w", str(w), "-b", "4", "-p", "20", "-c")
    assert r.returncode == 0, r.stderr
    m = gen_mask(oracle, (w, w), 20, 4)
    b = oracle.border_2d(w, w, m)
    d = {0: oracle.danielsson_2d, 1: oracle.fast_sweep_2d, 3: oracle.brute_force_2d}[method](w, w, b, 2)
    np.testing.assert_array_equal(np.fromfile(out, np.uint8).reshape(w, w),
                                  oracle.border_image(w, w, 0, m, d, 2 * 0.25 * w, 0))


def test_cli_two_field(gpu, oracle, tmp_path):
    """-S: exterior and interior fields; images 0 and 2"""
    w = 200
    m = gen_mask(oracle, (w, w), 30, 4)
    d0 = oracle.jfa_2d(w, w, oracle.init_deltas_2d(w, w, m, False), 1)
    d1 = oracle.jfa_2d(w, w, oracle.init_deltas_2d(w, w, m, True), 1)
    for it in (0, 2):
        r, out = run(tmp_path, "-S", "-i", str(it), "-w", str(w), "-b", "4", "-p", "30", "-c")
        assert r.returncode == 0, r.stderr
        exp = oracle.two_field_image_2d(w, w, d0, d1, 0.25 * w, it)
        np.testing.assert_array_equal(np.fromfile(out, np.uint8).reshape(exp.shape), exp)
    w = 24
    r, out = run(tmp_path, "-S", "-W", str(w), "-b", "5", "-p", "10")
    assert r.returncode == 0, r.stderr
    m = gen_mask(oracle, (w, w, w), 10, 5)
    d0 = oracle.jfa_3d(w, w, w, oracle.init_deltas_3d(w, w, w, m, False), 1)
    d1 = oracle.jfa_3d(w, w, w, oracle.init_deltas_3d(w, w, w, m, True), 1)
    np.testing.assert_array_equal(np.fromfile(out, np.uint8).reshape(w, w, w),
                                  oracle.two_field_image_3d(w, w, w, d0, d1, 0.25 * w, 0))


@pytest.mark.parametrize("threshold,reverse", [(128, False), (60, False), (200, True)])
def test_cli_raw_bytes(gpu, oracle, tmp_path, rng, threshold, reverse):
    """-f raw bytes with -t and -r: BitMaskFromImageFile's test, dark occupied
    by default"""
    w, h = 150, 90
    img = rng.integers(0, 256, (h, w)).astype(np.uint8)
    img[20:60, 30:90] = 10      # a dark block
    src = tmp_path / "in.raw"
    img.tofile(src)
    args = ["-s", "-w", str(w), str(h), "-f", str(src), "-t", str(threshold)] + (["-r"] if reverse else [])
    r, out = run(tmp_path, *args)
    assert r.returncode == 0, r.stderr
    m = oracle.mask_from_bytes(w, h, img.reshape(-1), threshold, not reverse)
    d = oracle.jfa_2d(w, h, oracle.border_2d(w, h, m), 2)
    np.testing.assert_array_equal(np.fromfile(out, np.uint8).reshape(h, w),
                                  oracle.border_image(w, h, 0, m, d, 2 * 0.25 * w, 0))


def test_cli_verbose_and_errors(gpu, tmp_path):
    r, _ = run(tmp_path, "-s", "-v", "-w", "64", "-p", "5")
    assert r.returncode == 0, r.stderr
    assert "border init" in r.stdout and "propagate" in r.stdout
    r, _ = run(tmp_path, "-s", "-m", "5")
    assert r.returncode != 0
    r, _ = run(tmp_path, "-d", "-x")
    assert r.returncode != 0
    r, _ = run(tmp_path, "-s", "-w", "64", "-f", str(tmp_path / "missing.raw"))
    assert r.returncode != 0
