"""GPU parity for the reference's methods off the data-parallel path:
BruteForce, Danielsson, Chamfer, the occlusion image and the shadow sweep,
bit-exact against the oracle on seeded inputs and against the reference's
own outputs (tests/golden/reference_vectors.npz, offpath_golden.npz).  Also
the cell width 3 (any cw the int16 deltas hold is accepted) for JumpFlood
and FastSweep.
"""
import os

import numpy as np
import pytest

from oracle_bind import ROOT, empty_mask, pack_mask, unpack_mask

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(ROOT, "tests", "golden")


def rand_mask(rng, dims, p):
    return pack_mask(rng.random(tuple(reversed(dims))) < p)


# ------------------------------------------------------------- BruteForce


@pytest.mark.parametrize("dims", [(1, 1), (1, 9), (9, 1), (5, 5), (33, 20), (64, 64), (100, 37)])
@pytest.mark.parametrize("cw", [1, 2, 3])
def test_brute_force_2d(gpu, oracle, rng, dims, cw):
    w, h = dims
    for p in (0.0, 0.01, 0.1, 0.5):
        m = rand_mask(rng, dims, p)
        start = oracle.init_deltas_2d(w, h, m) if cw != 2 else oracle.border_2d(w, h, m)
        np.testing.assert_array_equal(gpu.brute_force_2d(w, h, start, cw), oracle.brute_force_2d(w, h, start, cw))


@pytest.mark.parametrize("dims", [(1, 1, 1), (5, 4, 3), (17, 9, 6), (33, 8, 5)])
@pytest.mark.parametrize("cw", [1, 2])
def test_brute_force_3d(gpu, oracle, rng, dims, cw):
    w, h, d = dims
    for p in (0.0, 0.01, 0.1, 0.5):
        m = rand_mask(rng, dims, p)
        start = oracle.init_deltas_3d(w, h, d, m) if cw == 1 else oracle.border_3d(w, h, d, m)
        np.testing.assert_array_equal(gpu.brute_force_3d(w, h, d, start, cw),
                                      oracle.brute_force_3d(w, h, d, start, cw))


def test_brute_force_golden(gpu, oracle):
    """The reference's BruteForce outputs (cw = 1 from the unsigned init)."""
    vec = np.load(os.path.join(GOLDEN, "reference_vectors.npz"))
    for k in range(5):
        w, h = (int(v) for v in vec[f"d2_{k}_dims"])
        i0 = oracle.init_deltas_2d(w, h, vec[f"d2_{k}_mask"])
        np.testing.assert_array_equal(gpu.brute_force_2d(w, h, i0, 1), vec[f"d2_{k}_brute_cw1"])
    for k in range(4):
        w, h, d = (int(v) for v in vec[f"d3_{k}_dims"])
        i0 = oracle.init_deltas_3d(w, h, d, vec[f"d3_{k}_mask"])
        np.testing.assert_array_equal(gpu.brute_force_3d(w, h, d, i0, 1), vec[f"d3_{k}_brute_cw1"])


def test_brute_force_config1(gpu, oracle):
    """512 x 512 with 100 points (seed 12345): BruteForce and JumpFlood agree
    with the oracle's BruteForce, and JumpFlood is never shorter."""
    w = h = 512
    m = empty_mask(w, h)
    oracle.mask_add_points_2d(w, h, 100, 12345, m)
    i0 = oracle.init_deltas_2d(w, h, m)
    exact = oracle.brute_force_2d(w, h, i0, 1)
    np.testing.assert_array_equal(gpu.brute_force_2d(w, h, i0, 1), exact)
    jfa = gpu.jump_flood_2d(w, h, i0, 1).astype(np.int64)
    e = exact.astype(np.int64)
    assert ((jfa ** 2).sum(-1) >= (e ** 2).sum(-1)).all()


# ------------------------------------------------------------- Danielsson


@pytest.mark.parametrize("dims", [(1, 1), (1, 40), (40, 1), (5, 3), (33, 20), (64, 64), (130, 67), (1500, 5)])
@pytest.mark.parametrize("cw", [1, 2, 3])
def test_danielsson_2d(gpu, oracle, rng, dims, cw):
    w, h = dims
    for p in (0.0, 0.002, 0.05, 0.5):
        m = rand_mask(rng, dims, p)
        start = oracle.init_deltas_2d(w, h, m) if cw != 2 else oracle.border_2d(w, h, m)
        np.testing.assert_array_equal(gpu.danielsson_2d(w, h, start, cw), oracle.danielsson_2d(w, h, start, cw))


@pytest.mark.parametrize("dims", [(1, 1, 1), (5, 4, 3), (17, 9, 6), (1, 1, 30), (64, 20, 5), (1100, 3, 2)])
@pytest.mark.parametrize("cw", [1, 2])
def test_danielsson_3d(gpu, oracle, rng, dims, cw):
    w, h, d = dims
    for p in (0.0, 0.002, 0.05, 0.5):
        m = rand_mask(rng, dims, p)
        start = oracle.init_deltas_3d(w, h, d, m) if cw == 1 else oracle.border_3d(w, h, d, m)
        np.testing.assert_array_equal(gpu.danielsson_3d(w, h, d, start, cw),
                                      oracle.danielsson_3d(w, h, d, start, cw))


# ------------------------------------------------------------- Chamfer


@pytest.mark.parametrize("dims", [(1, 1), (1, 50), (50, 1), (2, 2), (33, 20), (64, 64), (4097, 3), (9000, 2),
                                  (130, 257)])
def test_chamfer(gpu, oracle, rng, dims):
    w, h = dims
    for p in (0.0, 0.003, 0.1):
        m = rand_mask(rng, dims, p)
        bits = unpack_mask(m, w, h)
        for dt in (np.int16, np.int32):
            # InitDistancesFromBitMask: 0 on set bits, the type's maximum elsewhere
            start = np.where(bits, 0, np.iinfo(dt).max).astype(dt)
            got, mult = gpu.chamfer(w, h, start)
            exp, emult = oracle.chamfer(w, h, start)
            assert mult == emult == 3
            np.testing.assert_array_equal(got, exp)
    # arbitrary values, negative ones included
    for dt, lo, hi in ((np.int16, -32768, 32767), (np.int32, -(1 << 30), 1 << 30)):
        start = rng.integers(lo, hi, (h, w)).astype(dt)
        np.testing.assert_array_equal(gpu.chamfer(w, h, start)[0], oracle.chamfer(w, h, start)[0])


# ------------------------------------------------------------- occlusion image and shadows


LIGHTS = [(0.3, 1.0), (-0.7, 0.2), (1.0, -1.0), (-0.25, -0.9), (0.0, 1.0), (2.0, 0.0), (0.6, 0.6)]


@pytest.mark.parametrize("dims", [(1, 1), (1, 30), (30, 1), (5, 3), (33, 20), (64, 64), (1500, 7), (6, 1300)])
def test_occlusion_image_and_shadow(gpu, oracle, rng, dims):
    w, h = dims
    for p in (0.0, 0.05, 0.5):
        m = rand_mask(rng, dims, p)
        for scale in (1.0, 0.37):
            occ = gpu.init_occ(w, h, m, scale)
            np.testing.assert_array_equal(occ.view(np.uint32), oracle.init_occ(w, h, m, scale).view(np.uint32))
            for gx, gy in LIGHTS:
                np.testing.assert_array_equal(gpu.shadow_sweep(gx, gy, w, h, occ).view(np.uint32),
                                              oracle.shadow_sweep(gx, gy, w, h, occ).view(np.uint32))


def test_offpath_golden(gpu):
    """The reference's own Danielsson, Chamfer, occlusion-image and shadow
    outputs (tests/golden/make_offpath_golden.py)."""
    g = np.load(os.path.join(GOLDEN, "offpath_golden.npz"))
    lights = [tuple(float(v) for v in l) for l in g["lights"]]
    for k in range(5):
        w, h = (int(v) for v in g[f"d2_{k}_dims"])
        for cw in (1, 2):
            np.testing.assert_array_equal(gpu.danielsson_2d(w, h, g[f"d2_{k}_cw{cw}_in"], cw), g[f"d2_{k}_cw{cw}_out"])
        for name in ("i16", "i32"):
            res, mult = gpu.chamfer(w, h, g[f"c_{k}_{name}_in"])
            np.testing.assert_array_equal(res, g[f"c_{k}_{name}_out"])
            assert mult == int(g[f"c_{k}_div"])
        occ = gpu.init_occ(w, h, g[f"d2_{k}_mask"], float(g[f"o_{k}_scale"]))
        np.testing.assert_array_equal(occ, g[f"o_{k}_occ"])
        for j, (gx, gy) in enumerate(lights):
            np.testing.assert_array_equal(gpu.shadow_sweep(gx, gy, w, h, occ).view(np.uint32),
                                          g[f"o_{k}_shadow{j}"].view(np.uint32))
    for k in range(4):
        w, h, d = (int(v) for v in g[f"d3_{k}_dims"])
        for cw in (1, 2):
            np.testing.assert_array_equal(gpu.danielsson_3d(w, h, d, g[f"d3_{k}_cw{cw}_in"], cw),
                                          g[f"d3_{k}_cw{cw}_out"])


# ------------------------------------------------------------- other cell widths


@pytest.mark.parametrize("dims", [(5, 5), (33, 20), (64, 64), (257, 70)])
@pytest.mark.parametrize("cw", [3, 5])
def test_jfa_2d_other_cw(gpu, oracle, rng, dims, cw):
    w, h = dims
    for p in (0.0, 0.01, 0.3):
        start = oracle.init_deltas_2d(w, h, rand_mask(rng, dims, p))
        np.testing.assert_array_equal(gpu.jump_flood_2d(w, h, start, cw), oracle.jfa_2d(w, h, start, cw))
        np.testing.assert_array_equal(gpu.fast_sweep_2d(w, h, start, cw), oracle.fast_sweep_2d(w, h, start, cw))


@pytest.mark.parametrize("dims", [(5, 4, 3), (33, 8, 5), (64, 20, 9)])
@pytest.mark.parametrize("cw", [3, 5])
def test_jfa_3d_other_cw(gpu, oracle, rng, dims, cw):
    w, h, d = dims
    for p in (0.0, 0.01, 0.3):
        start = oracle.init_deltas_3d(w, h, d, rand_mask(rng, dims, p))
        np.testing.assert_array_equal(gpu.jump_flood_3d(w, h, d, start, cw), oracle.jfa_3d(w, h, d, start, cw))
        np.testing.assert_array_equal(gpu.fast_sweep_3d(w, h, d, start, cw),
                                      oracle.fast_sweep_3d(w, h, d, start, cw))


def test_offpath_invalid_arguments(gpu):
    import dolhip

    z = np.zeros((4, 4, 2), np.int16)
    for call in (lambda: gpu.danielsson_2d(0, 4, z), lambda: gpu.danielsson_2d(4, 4, z, 0),
                 lambda: gpu.danielsson_2d(50000, 1, np.zeros((1, 50000, 2), np.int16)),
                 lambda: gpu.brute_force_2d(4, 4, z, -1), lambda: gpu.chamfer(0, 1, np.zeros(1, np.int32)),
                 lambda: gpu.shadow_sweep(1.0, 1.0, 0, 3, np.zeros(3, np.float32))):
        with pytest.raises(dolhip.DolHipError):
            call()
