"""GPU parity: every entry point of libdolhip against the oracle (and the
compiled reference where it was built), bit-exact, on seeded random masks.

Sizes straddle the 32-bit word boundaries of the bitmask rows (31, 32, 33,
63, 64, 65 ...) and include single rows/columns/slices, empty and full masks.
"""
import os

import numpy as np
import pytest

from oracle_bind import ROOT, empty_mask, pack_mask, unpack_mask

pytestmark = pytest.mark.gpu

SIZES_2D = [(1, 1), (1, 7), (7, 1), (2, 2), (5, 5), (31, 9), (32, 17), (33, 33), (63, 20),
            (64, 64), (65, 3), (100, 71), (129, 130), (257, 64)]
SIZES_3D = [(1, 1, 1), (2, 3, 4), (5, 5, 5), (31, 7, 6), (32, 9, 3), (33, 4, 9), (40, 17, 11),
            (65, 6, 5), (70, 33, 2)]
DENSITIES = [0.0, 0.01, 0.2, 0.5, 0.95, 1.0]


def rand_mask(rng, dims, p):
    bits = rng.random(tuple(reversed(dims))) < p
    return pack_mask(bits)


def blob_mask(rng, dims, n=6):
    """Union of random balls: large uniform regions with curved borders."""
    shape = tuple(reversed(dims))
    grids = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    bits = np.zeros(shape, dtype=bool)
    for _ in range(n):
        c = [rng.uniform(0, s) for s in shape]
        r = rng.uniform(1, max(2, max(shape) / 3))
        d2 = sum((g - cc) ** 2 for g, cc in zip(grids, c))
        bits |= d2 <= r * r
    return pack_mask(bits)


def masks_for(rng, dims):
    out = [rand_mask(rng, dims, p) for p in DENSITIES]
    out.append(blob_mask(rng, dims))
    return out


# ------------------------------------------------------------- init


@pytest.mark.parametrize("dims", SIZES_2D)
def test_init_deltas_2d(gpu, oracle, rng, dims):
    w, h = dims
    for m in masks_for(rng, dims):
        for inv in (False, True):
            np.testing.assert_array_equal(gpu.init_deltas_2d(w, h, m, inv), oracle.init_deltas_2d(w, h, m, inv))


@pytest.mark.parametrize("dims", SIZES_3D)
def test_init_deltas_3d(gpu, oracle, rng, dims):
    w, h, d = dims
    for m in masks_for(rng, dims):
        for inv in (False, True):
            np.testing.assert_array_equal(gpu.init_deltas_3d(w, h, d, m, inv), oracle.init_deltas_3d(w, h, d, m, inv))


@pytest.mark.parametrize("dims", SIZES_2D)
def test_border_2d(gpu, oracle, reference, rng, dims):
    w, h = dims
    for m in masks_for(rng, dims):
        got = gpu.init_border_deltas_2d(w, h, m)
        np.testing.assert_array_equal(got, oracle.border_2d(w, h, m))
        if reference is not None:
            np.testing.assert_array_equal(got, reference.border_2d(w, h, m))


@pytest.mark.parametrize("dims", [(32, 1), (96, 5), (160, 13), (320, 37), (2048, 3), (4096, 2)])
def test_border_2d_word_rows(gpu, oracle, rng, dims):
    """Widths that are a multiple of 32 (whole mask words per row): waves
    spanning several rows, partial last waves."""
    w, h = dims
    for m in masks_for(rng, dims):
        np.testing.assert_array_equal(gpu.init_border_deltas_2d(w, h, m), oracle.border_2d(w, h, m))


@pytest.mark.parametrize("dims", SIZES_3D)
def test_border_3d(gpu, oracle, reference, rng, dims):
    w, h, d = dims
    for m in masks_for(rng, dims):
        got = gpu.init_border_deltas_3d(w, h, d, m)
        np.testing.assert_array_equal(got, oracle.border_3d(w, h, d, m))
        if reference is not None:
            np.testing.assert_array_equal(got, reference.border_3d(w, h, d, m))


@pytest.mark.parametrize("dims", [(64, 5, 4), (96, 12, 10), (128, 3, 17), (32, 1, 1), (64, 1, 9), (64, 9, 1)])
def test_border_3d_word_rows(gpu, oracle, reference, rng, dims):
    """Whole-word rows (word ends inside the row and at its end) and
    volumes one row or one slice thick."""
    w, h, d = dims
    for m in masks_for(rng, dims):
        got = gpu.init_border_deltas_3d(w, h, d, m)
        np.testing.assert_array_equal(got, oracle.border_3d(w, h, d, m))
        if reference is not None:
            np.testing.assert_array_equal(got, reference.border_3d(w, h, d, m))


def test_init_distances(gpu, oracle, rng):
    # widths that are multiples of 4 (and of 8) take the 16-byte-a-thread
    # kernel for 4-byte (and 2-byte) cells; (2056, 5) leaves threads past
    # the row's end in its last workgroup
    for w, h in [(1, 1), (33, 17), (100, 64), (4, 3), (8, 2), (1024, 9), (2056, 5)]:
        m = rand_mask(rng, (w, h), 0.3)
        np.testing.assert_array_equal(gpu.init_distances_f32(w, h, m), oracle.init_distances_f32(w, h, m))
        exp = unpack_mask(m, w, h)
        np.testing.assert_array_equal(gpu.init_distances_i32(w, h, m, 1234), np.where(exp, 0, 1234))
        np.testing.assert_array_equal(gpu.init_distances_i16(w, h, m, 77), np.where(exp, 0, 77))


@pytest.mark.parametrize("w,rows", [(1, 1), (31, 3), (32, 5), (33, 7), (100, 9), (256, 4), (257, 3)])
def test_mask_from_bytes(gpu, oracle, rng, w, rows):
    data = rng.integers(0, 256, size=w * rows, dtype=np.uint8)
    for thr, lt in [(128, True), (128, False), (0, False), (255, True), (1, True)]:
        np.testing.assert_array_equal(gpu.mask_from_bytes(w, rows, data, thr, lt),
                                      oracle.mask_from_bytes(w, rows, data, thr, lt))


# ------------------------------------------------------------- jump flood


@pytest.mark.parametrize("dims", SIZES_2D)
@pytest.mark.parametrize("cw", [1, 2])
def test_jfa_2d(gpu, oracle, rng, dims, cw):
    w, h = dims
    for m in masks_for(rng, dims):
        start = oracle.border_2d(w, h, m) if cw == 2 else oracle.init_deltas_2d(w, h, m)
        np.testing.assert_array_equal(gpu.jump_flood_2d(w, h, start, cw), oracle.jfa_2d(w, h, start, cw))


@pytest.mark.parametrize("dims", SIZES_3D)
@pytest.mark.parametrize("cw", [1, 2])
def test_jfa_3d(gpu, oracle, rng, dims, cw):
    w, h, d = dims
    for m in masks_for(rng, dims):
        start = oracle.border_3d(w, h, d, m) if cw == 2 else oracle.init_deltas_3d(w, h, d, m)
        np.testing.assert_array_equal(gpu.jump_flood_3d(w, h, d, start, cw), oracle.jfa_3d(w, h, d, start, cw))


@pytest.mark.parametrize("dims", [(8, 8, 8), (64, 20, 12), (128, 9, 40), (512, 16, 16), (1024, 4, 3),
                                  (1032, 2, 3)])
@pytest.mark.parametrize("cw", [1, 2])
def test_jfa_3d_staged_rows(gpu, oracle, rng, dims, cw):
    """Widths that are multiples of 8 (rows staged whole in LDS up to 1024
    cells; 1032 takes the per-cell kernel)."""
    w, h, d = dims
    for m in masks_for(rng, dims):
        start = oracle.border_3d(w, h, d, m) if cw == 2 else oracle.init_deltas_3d(w, h, d, m)
        np.testing.assert_array_equal(gpu.jump_flood_3d(w, h, d, start, cw), oracle.jfa_3d(w, h, d, start, cw))


@pytest.mark.parametrize("bound", [3, 2047, 2049, 4000, 32767])
@pytest.mark.parametrize("cw", [1, 2])
def test_jfa_3d_rows_any_field(gpu, oracle, rng, bound, cw):
    """Caller fields that no initialiser makes: components uniform in
    [-bound, bound] with a fifth of the cells kMaxDelta3 (their y and z
    random).  Fields within 2048 take the keyed path of the row kernel, the
    others the general one, pass by pass; both must give the reference's
    wrapped int16 results."""
    for dims in [(64, 9, 7), (96, 33, 5), (512, 3, 4)]:
        w, h, d = dims
        f = rng.integers(-bound, bound + 1, size=(d, h, w, 3)).astype(np.int16)
        mx = rng.random((d, h, w)) < 0.2
        f[mx, 0] = 25000
        np.testing.assert_array_equal(gpu.jump_flood_3d(w, h, d, f, cw), oracle.jfa_3d(w, h, d, f, cw),
                                      err_msg=f"{dims}")


def test_jfa_2d_config1(gpu, oracle, reference):
    """The tool's -p 100 case: 512x512, 100 points from seed 12345."""
    w = h = 512
    m = empty_mask(w, h)
    oracle.mask_add_points_2d(w, h, 100, 12345, m)
    start = oracle.init_deltas_2d(w, h, m)
    got = gpu.jump_flood_2d(w, h, start, 1)
    np.testing.assert_array_equal(got, oracle.jfa_2d(w, h, start, 1))
    if reference is not None:
        np.testing.assert_array_equal(got, reference.jfa_2d(w, h, start, 1))
    # error against the exact transform stays below one cell (DistanceOcclusionTool.cpp:280-284)
    exact = np.sqrt(oracle.edt_2d(w, h, m).astype(np.float64))
    err = np.abs(oracle.dist_2d(got).astype(np.float64) - exact)
    assert err.max() < 1.0
    assert oracle.check_deltas_2d(w, h, got, 1) == 0


# ------------------------------------------------------------- Felzenszwalb


@pytest.mark.parametrize("dims", SIZES_2D + [(8192, 16), (16, 4100)])
def test_felzenszwalb_float(gpu, oracle, reference, rng, dims):
    w, h = dims
    ms = masks_for(rng, dims) if w * h < 20000 else [rand_mask(rng, dims, 0.0005), rand_mask(rng, dims, 0.01)]
    for m in ms:
        f0 = oracle.init_distances_f32(w, h, m)
        got = gpu.felzenszwalb_2d(w, h, f0)
        exp = oracle.felzenszwalb_2d(w, h, f0)
        np.testing.assert_array_equal(got.view(np.uint32), exp.view(np.uint32))
        if reference is not None:
            np.testing.assert_array_equal(got.view(np.uint32), reference.felzenszwalb_2d(w, h, f0).view(np.uint32))


def _felz_fields(rng, w, h):
    """Float fields for the Felzenszwalb envelope: random values, small
    integers (many ties), a ramp down each column whose last sample pops a
    stack deeper than the envelope kernel's LDS ring, a sawtooth of such
    ramps (each column its own period and phase: the stack grows past the
    ring, a zero pops far below it, and the next ramp pushes again over the
    depths the ring no longer held), and FLT_MAX with zeros."""
    fmax = np.finfo(np.float32).max
    yield rng.uniform(0, 1e4, (h, w)).astype(np.float32)
    yield rng.integers(0, 8, (h, w)).astype(np.float32)
    ramp = (1e5 - 100.0 * np.arange(h, dtype=np.float64))[:, None] * np.ones((1, w))
    ramp[-1, :] = 0.0
    yield ramp.astype(np.float32)
    period = rng.integers(40, 200, w)
    r = (np.arange(h)[:, None] + rng.integers(0, 200, w)[None, :]) % period[None, :]
    saw = 1e5 - 100.0 * r
    saw[r == period[None, :] - 1] = 0.0
    yield saw.astype(np.float32)
    yield np.where(rng.random((h, w)) < 0.02, 0.0, fmax).astype(np.float32)


@pytest.mark.parametrize("dims", [(1, 1), (1, 300), (300, 1), (97, 130), (513, 67), (64, 700), (4100, 8),
                                  (32767, 3)])
def test_felzenszwalb_float_any_field(gpu, oracle, reference, rng, dims):
    """DOL::Felzenszwalb on arbitrary float fields, bit for bit (the deep-pop
    ramp exercises the stack reads from device memory)."""
    w, h = dims
    for f0 in _felz_fields(rng, w, h):
        got = gpu.felzenszwalb_2d(w, h, f0)
        exp = oracle.felzenszwalb_2d(w, h, f0)
        np.testing.assert_array_equal(got.view(np.uint32), exp.view(np.uint32))
        if reference is not None:
            np.testing.assert_array_equal(got.view(np.uint32), reference.felzenszwalb_2d(w, h, f0).view(np.uint32))


def test_felzenszwalb_float_nan(gpu, oracle, rng):
    """NaN samples: the envelope's z loses its order, and the lines that hold
    a NaN take the reference's running scan; compared as floats (NaN equal to
    NaN; the payload bits are not the reference's concern)."""
    w, h = 150, 90
    f0 = rng.uniform(0, 1e3, (h, w)).astype(np.float32)
    f0[rng.random((h, w)) < 0.01] = np.nan
    np.testing.assert_array_equal(gpu.felzenszwalb_2d(w, h, f0), oracle.felzenszwalb_2d(w, h, f0))


# ------------------------------------------------------------- exact EDT


@pytest.mark.parametrize("dims", SIZES_2D + [(600, 3), (3, 600)])
def test_edt_2d(gpu, oracle, rng, dims):
    w, h = dims
    for m in masks_for(rng, dims):
        for inv in (False, True):
            d2, fl = gpu.edt_2d(w, h, m, inv)
            exp = oracle.edt_2d(w, h, m, inv)
            np.testing.assert_array_equal(d2, exp)
            expf = np.where(exp == 0x7FFFFFFF, np.sqrt(np.float32(np.finfo(np.float32).max)),
                            np.sqrt(exp.astype(np.float32)))
            np.testing.assert_array_equal(fl, expf.astype(np.float32))


@pytest.mark.parametrize("dims", SIZES_3D + [(64, 64, 64)])
def test_edt_3d(gpu, oracle, rng, dims):
    w, h, d = dims
    for m in masks_for(rng, dims):
        for inv in (False, True):
            d2, fl = gpu.edt_3d(w, h, d, m, inv)
            exp = oracle.edt_3d(w, h, d, m, inv)
            np.testing.assert_array_equal(d2, exp)
            finite = exp != 0x7FFFFFFF
            np.testing.assert_array_equal(fl[finite], np.sqrt(exp[finite].astype(np.float32)))


def test_signed_edt(gpu, oracle, rng):
    for dims in [(33, 17), (64, 64, 8)]:
        m = blob_mask(rng, dims)
        if len(dims) == 2:
            got = gpu.signed_edt_2d(*dims, m)
            a = oracle.edt_2d(*dims, m, False)
            b = oracle.edt_2d(*dims, m, True)
        else:
            got = gpu.signed_edt_3d(*dims, m)
            a = oracle.edt_3d(*dims, m, False)
            b = oracle.edt_3d(*dims, m, True)
        exp = np.sqrt(a.astype(np.float32)) - np.sqrt(b.astype(np.float32))
        np.testing.assert_array_equal(got, exp.astype(np.float32))


# ------------------------------------------------------------- encoders


def test_encoders_2d(gpu, oracle, rng):
    for w, h in [(33, 17), (128, 96)]:
        m = blob_mask(rng, (w, h))
        delta = oracle.jfa_2d(w, h, oracle.border_2d(w, h, m), 2)
        max_len = 0.5 * max(w, h)
        np.testing.assert_array_equal(gpu.signed_distance_2d(w, h, m, delta), oracle.signed_dist_2d(w, h, m, delta))
        np.testing.assert_array_equal(gpu.encode_sdf_u8_2d(w, h, m, delta, max_len),
                                      oracle.sdf_u8_2d(w, h, m, delta, max_len))
        np.testing.assert_array_equal(gpu.encode_sdf_delta_rgba_2d(w, h, delta, max_len),
                                      oracle.sdf_delta_rgba_2d(w, h, delta, max_len))
        np.testing.assert_array_equal(gpu.encode_sdf_rg_rgba_2d(w, h, m, delta, max_len),
                                      oracle.sdf_rg_rgba_2d(w, h, m, delta, max_len))


@pytest.mark.parametrize("dims", [(33, 20, 9), (32, 12, 7), (64, 5, 3)])
def test_encoders_3d(gpu, oracle, rng, dims):
    w, h, d = dims
    m = blob_mask(rng, (w, h, d))
    delta = oracle.jfa_3d(w, h, d, oracle.border_3d(w, h, d, m), 2)
    np.testing.assert_array_equal(gpu.signed_distance_3d(w, h, d, m, delta), oracle.signed_dist_3d(w, h, d, m, delta))
    np.testing.assert_array_equal(gpu.encode_sdf_u8_3d(w, h, d, m, delta, 16.5),
                                  oracle.sdf_u8_3d(w, h, d, m, delta, 16.5))


# ------------------------------------------------------------- errors


# ------------------------------------------------------------- voxeliser


def test_mask_from_triangles_golden(gpu, oracle):
    """The reference's masks of its sample mesh and of a triangle soup."""
    import os

    from oracle_bind import ROOT

    g = np.load(os.path.join(ROOT, "tests", "golden", "voxel_golden.npz"))
    for key in sorted(g.files):
        kind = key.split("_")[0]
        if "_mask_" not in key:
            continue
        w, h, d = (int(v) for v in key.split("_")[2:])
        got = gpu.mask_from_triangles(w, h, d, g[kind + "_positions"], g[kind + "_indices"], g[kind + "_bbox"])
        np.testing.assert_array_equal(got, g[key])


@pytest.mark.parametrize("dims", [(1, 1, 1), (7, 40, 3), (64, 64, 64), (130, 33, 70)])
def test_mask_from_triangles(gpu, oracle, rng, dims):
    w, h, d = dims
    for n in (1, 50, 2000):
        pos = rng.uniform(-0.2, 1.2, (3 * n, 3)).astype(np.float32)
        pos[: 3 * (n // 2)] = (rng.uniform(0, 1, (n // 2, 1, 3)) + rng.normal(0, 0.03, (n // 2, 3, 3))
                               ).astype(np.float32).reshape(-1, 3)
        bbox = np.array([0, 0, 0, 1, 1, 1], np.float32)
        idx = rng.permutation(3 * n).astype(np.int32)
        np.testing.assert_array_equal(gpu.mask_from_triangles(w, h, d, pos, idx, bbox),
                                      oracle.mask_from_triangles(w, h, d, pos, idx, bbox))
        # no index array: consecutive vertices; ORed into an existing mask
        base = rand_mask(rng, dims, 0.05)
        np.testing.assert_array_equal(gpu.mask_from_triangles(w, h, d, pos, None, bbox, base.copy()),
                                      oracle.mask_from_triangles(w, h, d, pos, None, bbox, base.copy()))


def test_mask_from_triangles_large_triangle(gpu, oracle):
    """One triangle across a 256^3 box (about 10^7 cells in its range)."""
    pos = np.array([[0.01, 0.02, 0.03], [0.97, 0.1, 0.9], [0.3, 0.99, 0.5]], np.float32)
    bbox = np.array([0, 0, 0, 1, 1, 1], np.float32)
    np.testing.assert_array_equal(gpu.mask_from_triangles(256, 256, 256, pos, None, bbox),
                                  oracle.mask_from_triangles(256, 256, 256, pos, None, bbox))


def test_mask_from_triangles_mid_ranges(gpu, oracle, rng):
    """Triangles whose cell ranges hold from a few to a few hundred cells (both
    sides of the one-pass / rows-first switch at 64 cells), some flat in one
    axis and some on cell-centre planes, on a 64 x 48 x 40 grid."""
    w, h, d = 64, 48, 40
    n = 3000
    centre = rng.uniform(0, 1, (n, 1, 3))
    size = rng.choice([0.02, 0.05, 0.1, 0.2], (n, 1, 1))
    pos = centre + size * rng.normal(0, 1, (n, 3, 3))
    flat = rng.integers(0, 4, n)                       # 3: not flat
    for a in range(3):
        sel = flat == a
        pos[sel, :, a] = pos[sel, :1, a]              # one coordinate shared by the three vertices
    dims = np.array([w, h, d])
    on_centre = rng.random(n) < 0.25
    a = flat[on_centre]
    keep = a < 3
    rows = np.nonzero(on_centre)[0][keep]
    a = a[keep]
    pos[rows, :, a] = ((np.floor(pos[rows, 0, a] * dims[a]) + 0.5) / dims[a])[:, None]
    pos = pos.astype(np.float32).reshape(-1, 3)
    bbox = np.array([0, 0, 0, 1, 1, 1], np.float32)
    np.testing.assert_array_equal(gpu.mask_from_triangles(w, h, d, pos, None, bbox),
                                  oracle.mask_from_triangles(w, h, d, pos, None, bbox))


# ------------------------------------------------------------- two-field (-S) images


@pytest.mark.parametrize("dims", [(1, 1), (5, 1), (1, 7), (33, 20), (64, 64), (100, 71)])
@pytest.mark.parametrize("cw", [1, 2])
def test_two_field_2d(gpu, oracle, rng, dims, cw):
    w, h = dims
    for m in [rand_mask(rng, dims, p) for p in (0.0, 0.05, 0.5, 1.0)] + [blob_mask(rng, dims)]:
        d0, d1 = gpu.two_field_deltas_2d(w, h, m, cw)
        np.testing.assert_array_equal(d0, oracle.jfa_2d(w, h, oracle.init_deltas_2d(w, h, m, False), cw))
        np.testing.assert_array_equal(d1, oracle.jfa_2d(w, h, oracle.init_deltas_2d(w, h, m, True), cw))
        for max_len in (0.25 * max(w, h), 3.0):
            for t in range(4):
                np.testing.assert_array_equal(gpu.encode_two_field_2d(w, h, d0, d1, max_len, t),
                                              oracle.two_field_image_2d(w, h, d0, d1, max_len, t))


@pytest.mark.parametrize("dims", [(1, 1, 1), (5, 5, 5), (33, 7, 6), (64, 12, 9)])
@pytest.mark.parametrize("cw", [1, 2])
def test_two_field_3d(gpu, oracle, rng, dims, cw):
    w, h, d = dims
    for m in [rand_mask(rng, dims, p) for p in (0.0, 0.05, 0.5)] + [blob_mask(rng, dims)]:
        d0, d1 = gpu.two_field_deltas_3d(w, h, d, m, cw)
        np.testing.assert_array_equal(d0, oracle.jfa_3d(w, h, d, oracle.init_deltas_3d(w, h, d, m, False), cw))
        np.testing.assert_array_equal(d1, oracle.jfa_3d(w, h, d, oracle.init_deltas_3d(w, h, d, m, True), cw))
        for t in range(4):
            np.testing.assert_array_equal(gpu.encode_two_field_3d(w, h, d, d0, d1, 0.25 * max(w, h, d), t),
                                          oracle.two_field_image_3d(w, h, d, d0, d1, 0.25 * max(w, h, d), t))


def test_two_field_rejects_bad_image_type(gpu):
    d0 = np.zeros((2, 2, 2), np.int16)
    with pytest.raises(Exception):
        gpu.encode_two_field_2d(2, 2, d0, d0, 1.0, 4)
    with pytest.raises(Exception):
        gpu.encode_two_field_2d(2, 2, d0, d0, 0.0, 0)


def test_invalid_arguments(gpu):
    import dolhip

    with pytest.raises(dolhip.DolHipError) as e:
        gpu.jump_flood_2d(0, 5, np.zeros((5, 0, 2), np.int16))
    assert e.value.code == dolhip.ERR_INVALID_ARG
    with pytest.raises(dolhip.DolHipError) as e:
        gpu.jump_flood_2d(4, 4, np.zeros((4, 4, 2), np.int16), cw=16385)
    assert e.value.code == dolhip.ERR_INVALID_ARG
    for cw in (0, -2):
        with pytest.raises(dolhip.DolHipError):
            gpu.jump_flood_3d(2, 2, 2, np.zeros((2, 2, 2, 3), np.int16), cw=cw)
    with pytest.raises(dolhip.DolHipError):
        gpu._call("dolhip_init_deltas_2d", 4, 4, None, None, 0)


def _pattern_masks(rng, dims):
    """Adversarial patterns for the envelope: walls (every site survives),
    checkerboards, single voxels, diagonal lines, plus random and blobs."""
    shape = tuple(reversed(dims))
    out = [np.zeros(shape, bool), np.ones(shape, bool)]
    one = np.zeros(shape, bool)
    one[tuple(int(rng.integers(0, s)) for s in shape)] = True
    out.append(one)
    wall = np.zeros(shape, bool)
    wall[..., shape[-1] // 3] = True          # plane x = const
    out.append(wall)
    wall2 = np.zeros(shape, bool)
    wall2[..., shape[-2] // 2, :] = True      # plane y = const
    out.append(wall2)
    idx = np.indices(shape).sum(axis=0)
    out.append(idx % 2 == 0)
    out.append(idx % 7 == 0)
    out.append(rng.random(shape) < 0.001)
    out.append(rng.random(shape) < 0.3)
    return [pack_mask(b) for b in out] + [blob_mask(rng, dims, 12)]


@pytest.mark.parametrize("dims", [(512, 64, 24), (130, 257, 33), (33, 512, 5), (5, 7, 512),
                                  (512, 512, 3), (31, 1, 512), (1, 40, 1), (64, 3, 512),
                                  (96, 2, 512)])
def test_edt_3d_tiled_patterns(gpu, oracle, rng, dims):
    w, h, d = dims
    for m in _pattern_masks(rng, dims):
        for inv in (False, True):
            d2, fl = gpu.edt_3d(w, h, d, m, inv)
            exp = oracle.edt_3d(w, h, d, m, inv)
            np.testing.assert_array_equal(d2, exp)
            finite = exp != 0x7FFFFFFF
            np.testing.assert_array_equal(fl[finite], np.sqrt(exp[finite].astype(np.float32)))


@pytest.mark.parametrize("dims", [(400, 8, 3), (402, 3, 8), (300, 300, 2), (12, 400, 6)])
def test_edt_3d_far_sites(gpu, oracle, rng, dims):
    """Sites farther than 181 cells along x or y from short lines: the 16-bit
    scan of wide tiles cannot give these distances exactly and must hand the
    tile to the exact path."""
    w, h, d = dims
    shape = (d, h, w)
    cases = []
    for corner in [(0, 0, 0), (d - 1, h - 1, w - 1), (d // 2, 0, w - 1)]:
        b = np.zeros(shape, bool)
        b[corner] = True
        cases.append(b)
    b = np.zeros(shape, bool)
    b[:, :, 0] = True                              # plane x = 0
    cases.append(b)
    b = np.zeros(shape, bool)
    b[0, :, :] = rng.random((h, w)) < 0.0005      # a few sites in slice 0 only
    cases.append(b)
    for bits in cases:
        m = pack_mask(bits)
        d2, fl = gpu.edt_3d(w, h, d, m)
        exp = oracle.edt_3d(w, h, d, m)
        np.testing.assert_array_equal(d2, exp)
        finite = exp != 0x7FFFFFFF
        np.testing.assert_array_equal(fl[finite], np.sqrt(exp[finite].astype(np.float32)))


@pytest.mark.parametrize("dims", [(512, 512, 512), (511, 384, 300), (512, 512), (301, 512)])
def test_edt_corner_site_every_root(gpu, dims):
    """One site at the origin: d2 = x^2 + y^2 (+ z^2) at every cell, a closed
    form, so the outputs hold every squared distance a grid of this size can
    produce, and each float must be IEEE sqrt of it bit for bit.  (The tiled
    path's roots rest on the hardware root of an integer below 2^20 being
    correctly rounded or one ulp low; this covers every value it meets.)"""
    shape = tuple(reversed(dims))
    bits = np.zeros(shape, bool)
    bits[(0,) * len(dims)] = True
    m = pack_mask(bits)
    sq = [np.arange(n, dtype=np.int32) ** 2 for n in shape]
    if len(dims) == 3:
        d2, fl = gpu.edt_3d(*dims, m)
        exp = sq[0][:, None, None] + sq[1][None, :, None] + sq[2][None, None, :]
    else:
        d2, fl = gpu.edt_2d(*dims, m)
        exp = sq[0][:, None] + sq[1][None, :]
    np.testing.assert_array_equal(d2, exp)
    np.testing.assert_array_equal(fl.view(np.uint32), np.sqrt(exp.astype(np.float32)).view(np.uint32))


@pytest.mark.parametrize("dims", [(64, 40, 24), (33, 20, 9), (64, 64)])
def test_edt_one_output(gpu, oracle, rng, dims):
    """Either output may be omitted (a null pointer)."""
    m = blob_mask(rng, dims)
    if len(dims) == 2:
        exp = oracle.edt_2d(*dims, m, False)
        d2, fl = gpu.edt_2d(*dims, m, want_dist=False)
        assert fl is None
        np.testing.assert_array_equal(d2, exp)
        d2, fl = gpu.edt_2d(*dims, m, want_d2=False)
    else:
        exp = oracle.edt_3d(*dims, m, False)
        d2, fl = gpu.edt_3d(*dims, m, want_dist=False)
        assert fl is None
        np.testing.assert_array_equal(d2, exp)
        d2, fl = gpu.edt_3d(*dims, m, want_d2=False)
    assert d2 is None
    finite = exp != 0x7FFFFFFF
    np.testing.assert_array_equal(fl[finite], np.sqrt(exp[finite].astype(np.float32)))
    assert np.all(np.isinf(fl[~finite]) | (fl[~finite] > 1e19))


@pytest.mark.parametrize("dims", [(512, 512), (300, 511), (512, 7), (7, 512), (513, 40), (40, 513)])
def test_edt_2d_patterns(gpu, oracle, rng, dims):
    w, h = dims
    for m in _pattern_masks(rng, dims):
        for inv in (False, True):
            d2, _ = gpu.edt_2d(w, h, m, inv)
            np.testing.assert_array_equal(d2, oracle.edt_2d(w, h, m, inv))


def test_edt_3d_scene_slab(gpu, oracle):
    """A 512 x 512 x 48 slab of the seeded ellipsoid/box scene."""
    import dolhip

    w = h = 512
    d = 48
    m = dolhip.load().scene_shapes_3d(w, h, d, 40, 14, seed=11)
    d2, _ = gpu.edt_3d(w, h, d, m)
    np.testing.assert_array_equal(d2, oracle.edt_3d(w, h, d, m))


# ------------------------------------------------------------- FastSweep


@pytest.mark.parametrize("dims", [(1, 1), (5, 3), (1, 90), (90, 1), (64, 64), (65, 64), (64, 65),
                                  (130, 67), (200, 129), (257, 70)])
def test_fast_sweep_2d(gpu, oracle, rng, dims):
    w, h = dims
    for p in (0.0, 0.002, 0.05, 0.5):
        m = rand_mask(rng, dims, p)
        for cw, start in ((1, oracle.init_deltas_2d(w, h, m)), (2, oracle.border_2d(w, h, m))):
            np.testing.assert_array_equal(gpu.fast_sweep_2d(w, h, start, cw), oracle.fast_sweep_2d(w, h, start, cw))


@pytest.mark.parametrize("dims", [(1, 1, 1), (5, 4, 3), (17, 9, 6), (64, 64, 3), (70, 65, 5), (33, 130, 4),
                                  (1, 1, 40), (129, 3, 7)])
def test_fast_sweep_3d(gpu, oracle, rng, dims):
    w, h, d = dims
    for p in (0.0, 0.002, 0.05, 0.5):
        m = rand_mask(rng, dims, p)
        for cw, start in ((1, oracle.init_deltas_3d(w, h, d, m)), (2, oracle.border_3d(w, h, d, m))):
            np.testing.assert_array_equal(gpu.fast_sweep_3d(w, h, d, start, cw),
                                          oracle.fast_sweep_3d(w, h, d, start, cw))


def test_fast_sweep_golden(gpu):
    """The reference's own outputs (tests/golden/sweep_golden.npz)."""
    g = np.load(os.path.join(ROOT, "tests", "golden", "sweep_golden.npz"))
    n = 0
    for key in g.files:
        if not key.endswith("_dims"):
            continue
        pre = key[: -len("_dims")]
        dims = tuple(int(v) for v in g[key])
        for cw in (1, 2):
            start, exp = g[f"{pre}_cw{cw}_in"], g[f"{pre}_cw{cw}_out"]
            got = gpu.fast_sweep_2d(*dims, start, cw) if len(dims) == 2 else gpu.fast_sweep_3d(*dims, start, cw)
            np.testing.assert_array_equal(got, exp)
            n += 1
    assert n == 18


def test_fast_sweep_large_single_seed(gpu):
    """Full-size property: one seed, every cell ends with the exact offset to
    it (2D anywhere; 3D for a seed on the last row of the last slice)."""
    w, h = 1500, 1100
    start = np.full((h, w, 2), 32767, np.int16)
    start[617, 1234] = 0
    out = gpu.fast_sweep_2d(w, h, start, 1)
    ys, xs = np.mgrid[0:h, 0:w]
    np.testing.assert_array_equal(out[..., 0], 1234 - xs)
    np.testing.assert_array_equal(out[..., 1], 617 - ys)
    w, h, d = 150, 131, 97
    start = np.full((d, h, w, 3), 25000, np.int16)
    start[d - 1, h - 1, 77] = 0
    out = gpu.fast_sweep_3d(w, h, d, start, 1)
    zs, ys, xs = np.mgrid[0:d, 0:h, 0:w]
    np.testing.assert_array_equal(out[..., 0], 77 - xs)
    np.testing.assert_array_equal(out[..., 1], h - 1 - ys)
    np.testing.assert_array_equal(out[..., 2], d - 1 - zs)


# ------------------------------------------------------------- occlusion sweeps


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("dims", [(1, 1), (5, 3), (1, 40), (40, 1), (8, 15), (9, 16), (64, 61), (130, 67), (257, 33)])
def test_occlusion_2d(gpu, oracle, rng, dims):
    w, h = dims
    for p in (0.0, 0.01, 0.2, 1.0):
        m = rand_mask(rng, dims, p)
        dw = gpu.init_dirw_2d(w, h, m)
        np.testing.assert_array_equal(dw, oracle.init_dirw_2d(w, h, m))
        np.testing.assert_array_equal(_bits(gpu.occlusion_sweep_2d(w, h, dw)), _bits(oracle.occlusion_sweep_2d(w, h, dw)))


@pytest.mark.parametrize("dims", [(1, 1, 1), (5, 4, 3), (17, 9, 6), (8, 15, 1), (33, 31, 5), (64, 20, 9), (3, 2, 40)])
def test_occlusion_3d(gpu, oracle, rng, dims):
    w, h, d = dims
    for p in (0.0, 0.01, 0.2):
        m = rand_mask(rng, dims, p)
        dw = gpu.init_dirw_3d(w, h, d, m)
        np.testing.assert_array_equal(dw, oracle.init_dirw_3d(w, h, d, m))
        np.testing.assert_array_equal(_bits(gpu.occlusion_sweep_3d(w, h, d, dw)),
                                      _bits(oracle.occlusion_sweep_3d(w, h, d, dw)))
        o0 = rng.random(dw.shape).astype(np.float32)
        np.testing.assert_array_equal(_bits(gpu.occlusion_sweep_3d_from(w, h, d, dw, o0)),
                                      _bits(oracle.occlusion_sweep_3d_from(w, h, d, dw, o0)))


def test_occlusion_golden(gpu):
    """The reference's own outputs (tests/golden/occlusion_golden.npz)."""
    g = np.load(os.path.join(ROOT, "tests", "golden", "occlusion_golden.npz"))
    for k in range(4):
        w, h = (int(v) for v in g[f"o2_{k}_dims"])
        dw = gpu.init_dirw_2d(w, h, g[f"o2_{k}_mask"])
        np.testing.assert_array_equal(dw, g[f"o2_{k}_dirw"])
        np.testing.assert_array_equal(_bits(gpu.occlusion_sweep_2d(w, h, dw)), _bits(g[f"o2_{k}_swept"]))
    for k in range(3):
        w, h, d = (int(v) for v in g[f"o3_{k}_dims"])
        dw = gpu.init_dirw_3d(w, h, d, g[f"o3_{k}_mask"])
        np.testing.assert_array_equal(_bits(gpu.occlusion_sweep_3d(w, h, d, dw)), _bits(g[f"o3_{k}_swept"]))
        np.testing.assert_array_equal(_bits(gpu.occlusion_sweep_3d_from(w, h, d, dw, g[f"o3_{k}_out0"])),
                                      _bits(g[f"o3_{k}_from"]))


def test_occlusion_large(gpu, oracle, rng):
    """2048 x 1500 against the oracle, and a full-size 1D closed form: a set
    pixel at the start of each row decays by 0.75 a cell along plane 0."""
    w, h = 2048, 1500
    m = rand_mask(rng, (w, h), 0.002)
    dw = oracle.init_dirw_2d(w, h, m)
    np.testing.assert_array_equal(_bits(gpu.occlusion_sweep_2d(w, h, dw)), _bits(oracle.occlusion_sweep_2d(w, h, dw)))
    w = 3000
    dw = np.zeros((4, 1, w), np.float32)
    dw[:, 0, 0] = 1.0
    exp = np.empty(w, np.float32)
    exp[0] = 1.0
    for x in range(1, w):
        exp[x] = np.float32(0.75) * exp[x - 1]
    np.testing.assert_array_equal(gpu.occlusion_sweep_2d(w, 1, dw)[0, 0], exp)


@pytest.mark.parametrize("shape", [(4, 1, 1), (4, 37, 53), (8, 3, 5, 7), (8, 9, 17, 33)])
def test_ao_encoders(gpu, oracle, rng, shape):
    """AO images of swept fields (random fields, all-zero and all-one fields)
    against the oracle, byte for byte."""
    for fill in (None, 0.0, 1.0):
        d = rng.random(shape).astype(np.float32) if fill is None else np.full(shape, fill, np.float32)
        for rgba in (False, True):
            np.testing.assert_array_equal(gpu.encode_ao(d, rgba), oracle.encode_ao(d, rgba))
    # a swept field from a mask
    dims = tuple(reversed(shape[1:]))
    m = rand_mask(rng, dims, 0.05)
    dw = oracle.init_dirw_2d(*dims, m) if len(dims) == 2 else oracle.init_dirw_3d(*dims, m)
    dw = oracle.occlusion_sweep_2d(*dims, dw) if len(dims) == 2 else oracle.occlusion_sweep_3d(*dims, dw)
    for rgba in (False, True):
        np.testing.assert_array_equal(gpu.encode_ao(dw, rgba), oracle.encode_ao(dw, rgba))


@pytest.mark.parametrize("seed", range(4))
def test_sweeps_random_sizes(gpu, oracle, seed):
    """FastSweep and the occlusion sweeps on seeded random sizes (ragged tile
    edges in every direction) and densities, against the oracle."""
    r = np.random.default_rng(1000 + seed)
    for _ in range(3):
        w, h = (int(v) for v in r.integers(1, 300, 2))
        m = rand_mask(r, (w, h), float(r.choice([0.0, 0.001, 0.02, 0.3])))
        for cw, start in ((1, oracle.init_deltas_2d(w, h, m)), (2, oracle.border_2d(w, h, m))):
            np.testing.assert_array_equal(gpu.fast_sweep_2d(w, h, start, cw), oracle.fast_sweep_2d(w, h, start, cw))
        dw = oracle.init_dirw_2d(w, h, m)
        np.testing.assert_array_equal(_bits(gpu.occlusion_sweep_2d(w, h, dw)), _bits(oracle.occlusion_sweep_2d(w, h, dw)))
        w, h, d = (int(v) for v in r.integers(1, 40, 3))
        m = rand_mask(r, (w, h, d), float(r.choice([0.0, 0.002, 0.05])))
        start = oracle.border_3d(w, h, d, m)
        np.testing.assert_array_equal(gpu.fast_sweep_3d(w, h, d, start, 2), oracle.fast_sweep_3d(w, h, d, start, 2))
        dw = oracle.init_dirw_3d(w, h, d, m)
        np.testing.assert_array_equal(_bits(gpu.occlusion_sweep_3d(w, h, d, dw)),
                                      _bits(oracle.occlusion_sweep_3d(w, h, d, dw)))


# ------------------------------------------------------------- long lines (scan + stack fallback)


@pytest.mark.parametrize("dims", [(1100, 700), (3000, 40), (40, 3000), (600, 3), (2049, 1)])
def test_edt_2d_long_lines(gpu, oracle, rng, dims):
    """Lines above 512 samples: the outward scan, and the stack kernel for the
    lines whose scan reaches its cap (long empty stretches: one site, empty
    and full masks), both polarities."""
    w, h = dims
    masks = [rand_mask(rng, dims, p) for p in (0.0, 0.001, 0.02, 0.3, 1.0)]
    one = np.zeros((h, w), bool)
    one[h // 3, (2 * w) // 3] = True
    masks.append(pack_mask(one))
    for m in masks:
        for inv in (False, True):
            d2, fl = gpu.edt_2d(w, h, m, inv)
            exp = oracle.edt_2d(w, h, m, inv)
            np.testing.assert_array_equal(d2, exp)
            expf = np.where(exp == 0x7FFFFFFF, np.sqrt(np.float32(np.finfo(np.float32).max)),
                            np.sqrt(exp.astype(np.float32)))
            np.testing.assert_array_equal(fl, expf.astype(np.float32))


@pytest.mark.parametrize("dims", [(600, 5, 7), (5, 600, 4), (4, 5, 1030), (530, 20, 9)])
def test_edt_3d_long_lines(gpu, oracle, rng, dims):
    w, h, d = dims
    for p in (0.0, 0.002, 0.05, 1.0):
        m = rand_mask(rng, dims, p)
        for inv in (False, True):
            d2, _ = gpu.edt_3d(w, h, d, m, inv)
            np.testing.assert_array_equal(d2, oracle.edt_3d(w, h, d, m, inv))
