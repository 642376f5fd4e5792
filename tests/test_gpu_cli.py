"""The DOL:: drop-in interface end to end through distocc_hip, the tool-like
command built on it: its outputs equal the oracle's for the tool's own
cases (-p points with the default seed 12345, -b blocks)."""
import os
import subprocess

import numpy as np
import pytest

from oracle_bind import ROOT, empty_mask

pytestmark = pytest.mark.gpu

CLI = os.path.join(ROOT, "distance-occlusion_amd", "cli", "distocc_hip")


def run(tmp_path, *args):
    out = tmp_path / "out.bin"
    r = subprocess.run([CLI, *args, "-o", str(out)], capture_output=True, text=True, timeout=120)
    return r, out


def mask2(oracle, w, points, sides=0, seed=12345):
    m = empty_mask(w, w)
    if sides:
        oracle.mask_add_block_2d(w, w, sides, m)
    if points:
        oracle.mask_add_points_2d(w, w, points, seed, m)
    return m


def mask3(oracle, w, points, sides=0, seed=12345):
    m = empty_mask(w, w, w)
    if sides:
        oracle.mask_add_block_3d(w, w, w, sides, m)
    if points:
        oracle.mask_add_points_3d(w, w, w, points, seed, m)
    return m


def test_cli_distance_jfa(gpu, oracle, tmp_path):
    r, out = run(tmp_path, "-d", "-m", "2", "-w", "512", "-p", "100", "-c")
    assert r.returncode == 0, r.stderr
    assert "Max error" in r.stderr
    m = mask2(oracle, 512, 100)
    exp = oracle.dist_2d(oracle.jfa_2d(512, 512, oracle.init_deltas_2d(512, 512, m), 1))
    np.testing.assert_array_equal(np.fromfile(out, np.float32).reshape(512, 512), exp)


def test_cli_distance_felzenszwalb(gpu, oracle, tmp_path):
    r, out = run(tmp_path, "-d", "-m", "5", "-w", "300", "-b", "5", "-p", "40", "-c")
    assert r.returncode == 0, r.stderr
    m = mask2(oracle, 300, 40, 5)
    exp = np.sqrt(oracle.felzenszwalb_2d(300, 300, oracle.init_distances_f32(300, 300, m)))
    np.testing.assert_array_equal(np.fromfile(out, np.float32).reshape(300, 300), exp)


def test_cli_signed_2d(gpu, oracle, tmp_path):
    r, out = run(tmp_path, "-s", "-w", "256", "-b", "4", "-p", "30")
    assert r.returncode == 0, r.stderr
    m = mask2(oracle, 256, 30, 4)
    d = oracle.jfa_2d(256, 256, oracle.border_2d(256, 256, m), 2)
    exp = oracle.sdf_u8_2d(256, 256, m, d, 2 * 0.25 * 256)
    np.testing.assert_array_equal(np.fromfile(out, np.uint8).reshape(256, 256), exp)


def test_cli_volume(gpu, oracle, tmp_path):
    w = 40
    r, out = run(tmp_path, "-d", "-W", str(w), "-p", "20", "-c")
    assert r.returncode == 0, r.stderr
    m = mask3(oracle, w, 20)
    dl = oracle.jfa_3d(w, w, w, oracle.init_deltas_3d(w, w, w, m), 1).astype(np.int32)
    exp = np.sqrt((dl ** 2).sum(-1).astype(np.float32))
    np.testing.assert_array_equal(np.fromfile(out, np.float32).reshape(w, w, w), exp)
    r, out = run(tmp_path, "-s", "-W", str(w), "-b", "7", "-p", "10")
    assert r.returncode == 0, r.stderr
    m = mask3(oracle, w, 10, 7)
    d = oracle.jfa_3d(w, w, w, oracle.border_3d(w, w, w, m), 2)
    np.testing.assert_array_equal(np.fromfile(out, np.uint8).reshape(w, w, w),
                                  oracle.sdf_u8_3d(w, w, w, m, d, 2 * 0.25 * w))


def test_cli_two_field(gpu, oracle, tmp_path):
    """-S: exterior and interior JumpFlood fields, (Dist(d0) - Dist(d1)) / maxLen"""
    r, out = run(tmp_path, "-S", "-w", "200", "-b", "4", "-p", "30")
    assert r.returncode == 0, r.stderr
    m = mask2(oracle, 200, 30, 4)
    d0 = oracle.jfa_2d(200, 200, oracle.init_deltas_2d(200, 200, m, False), 1)
    d1 = oracle.jfa_2d(200, 200, oracle.init_deltas_2d(200, 200, m, True), 1)
    np.testing.assert_array_equal(np.fromfile(out, np.uint8).reshape(200, 200),
                                  oracle.two_field_image_2d(200, 200, d0, d1, 0.25 * 200, 0))
    w = 24
    r, out = run(tmp_path, "-S", "-W", str(w), "-b", "5", "-p", "10")
    assert r.returncode == 0, r.stderr
    m = mask3(oracle, w, 10, 5)
    d0 = oracle.jfa_3d(w, w, w, oracle.init_deltas_3d(w, w, w, m, False), 1)
    d1 = oracle.jfa_3d(w, w, w, oracle.init_deltas_3d(w, w, w, m, True), 1)
    np.testing.assert_array_equal(np.fromfile(out, np.uint8).reshape(w, w, w),
                                  oracle.two_field_image_3d(w, w, w, d0, d1, 0.25 * w, 0))


def test_cli_rejects_unknown(gpu, tmp_path):
    r, _ = run(tmp_path, "-s", "-m", "5")
    assert r.returncode != 0


def test_cli_fast_sweep(gpu, oracle, tmp_path):
    """-m 1: the tool's FindDeltas with FastSweep, in -d (checked against
    the exact EDT, as -c does) and -s flows, 2D and 3D."""
    r, out = run(tmp_path, "-d", "-m", "1", "-w", "300", "-p", "40", "-c")
    assert r.returncode == 0, r.stderr
    m = mask2(oracle, 300, 40)
    exp = oracle.dist_2d(oracle.fast_sweep_2d(300, 300, oracle.init_deltas_2d(300, 300, m), 1))
    np.testing.assert_array_equal(np.fromfile(out, np.float32).reshape(300, 300), exp)
    r, out = run(tmp_path, "-s", "-m", "1", "-w", "256", "-b", "4", "-p", "30")
    assert r.returncode == 0, r.stderr
    m = mask2(oracle, 256, 30, 4)
    d = oracle.fast_sweep_2d(256, 256, oracle.border_2d(256, 256, m), 2)
    np.testing.assert_array_equal(np.fromfile(out, np.uint8).reshape(256, 256),
                                  oracle.sdf_u8_2d(256, 256, m, d, 2 * 0.25 * 256))
    w = 40
    r, out = run(tmp_path, "-s", "-m", "1", "-W", str(w), "-b", "7", "-p", "10")
    assert r.returncode == 0, r.stderr
    m = mask3(oracle, w, 10, 7)
    d = oracle.fast_sweep_3d(w, w, w, oracle.border_3d(w, w, w, m), 2)
    np.testing.assert_array_equal(np.fromfile(out, np.uint8).reshape(w, w, w),
                                  oracle.sdf_u8_3d(w, w, w, m, d, 2 * 0.25 * w))
