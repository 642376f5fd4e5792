"""GPU parity of the tool's image writers (the border method's four image
types in 2D and 3D, the unsigned distance images of the delta and scalar
methods, the float roots) against the oracle, bit for bit, and of whole
device pipelines (_dev entry points on torch tensors) against the images the
reference's own tool wrote (tests/golden/image_golden.npz).
"""
import os

import numpy as np
import pytest

from oracle_bind import ROOT, empty_mask, pack_mask, unpack_mask

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def torch():
    import torch as t

    return t


def rand_mask(rng, dims, p):
    return pack_mask(rng.random(tuple(reversed(dims))) < p)


def odd_deltas(rng, shape):
    """deltas as the propagation leaves them, plus the odd cases: zeros (Dist
    0: the angle image divides by it), the int16 extremes whose 32-bit
    squared length wraps, MAX cells"""
    d = rng.integers(-600, 600, shape).astype(np.int16)
    flat = d.reshape(-1, shape[-1])
    n = flat.shape[0]
    k = max(1, n // 20)
    flat[rng.integers(0, n, k)] = 0
    flat[rng.integers(0, n, k)] = -32768
    flat[rng.integers(0, n, k)] = 32767
    flat[rng.integers(0, n, k), 0] = 25000
    return d


@pytest.mark.parametrize("dims", [(1, 1), (5, 3), (33, 20), (64, 64), (257, 70)])
@pytest.mark.parametrize("image_type", [0, 1, 2, 3])
def test_border_image_2d(gpu, oracle, rng, dims, image_type):
    w, h = dims
    m = rand_mask(rng, dims, 0.4)
    for delta in (odd_deltas(rng, (h, w, 2)), oracle.jfa_2d(w, h, oracle.border_2d(w, h, m), 2)):
        for max_len in (0.5 * max(w, h), 254.0, 3.7):
            np.testing.assert_array_equal(gpu.border_image(w, h, 0, m, delta, max_len, image_type),
                                          oracle.border_image(w, h, 0, m, delta, max_len, image_type))


@pytest.mark.parametrize("dims", [(1, 1, 1), (5, 4, 3), (33, 17, 6), (64, 20, 9)])
@pytest.mark.parametrize("image_type", [0, 1, 2, 3])
def test_border_image_3d(gpu, oracle, rng, dims, image_type):
    w, h, d = dims
    m = rand_mask(rng, dims, 0.4)
    for delta in (odd_deltas(rng, (d, h, w, 3)), oracle.jfa_3d(w, h, d, oracle.border_3d(w, h, d, m), 2)):
        for max_len in (0.5 * max(w, h), 254.0):
            np.testing.assert_array_equal(gpu.border_image(w, h, d, m, delta, max_len, image_type),
                                          oracle.border_image(w, h, d, m, delta, max_len, image_type))


@pytest.mark.parametrize("dims", [(1, 1, 0), (33, 20, 0), (257, 70, 0), (5, 4, 3), (40, 17, 9)])
def test_distance_and_scalar_images(gpu, oracle, rng, dims):
    w, h, d = dims
    dd = 3 if d else 2
    delta = odd_deltas(rng, (max(d, 1), h, w, dd))
    for max_len in (0.25 * max(w, h), 7.0):
        exp = oracle.distance_u8(dd, delta, max_len).reshape((d, h, w) if d else (h, w))
        np.testing.assert_array_equal(gpu.distance_u8(w, h, d, delta, max_len), exp)
    if d:
        return
    fvals = (rng.random((h, w)) * 1e5).astype(np.float32)
    fvals.reshape(-1)[:: 7] = np.float32(np.finfo(np.float32).max)
    fvals.reshape(-1)[:: 11] = 0
    roots = gpu.sqrt_f32(w, h, fvals)
    np.testing.assert_array_equal(roots.view(np.uint32), oracle.sqrt_f32(fvals).reshape(h, w).view(np.uint32))
    for vals in (roots, rng.integers(-5, 40000, (h, w)).astype(np.int32), rng.integers(-300, 32767, (h, w)).astype(np.int16)):
        for max_len in (0.25 * max(w, h), 96.0):
            np.testing.assert_array_equal(gpu.scalar_u8(w, h, vals, max_len),
                                          oracle.scalar_u8(vals, max_len).reshape(h, w))


def _tool_mask(oracle, dims, sides, points, seed):
    m = empty_mask(*dims)
    gen = (oracle.mask_add_block_2d, oracle.mask_add_points_2d) if len(dims) == 2 else \
        (oracle.mask_add_block_3d, oracle.mask_add_points_3d)
    if sides:
        gen[0](*dims, sides, m)
    if points:
        gen[1](*dims, points, seed, m)
    return m


def test_tool_images_on_device(gpu, oracle, torch):
    """The tool's flows run as device pipelines (mask on the device, the
    _dev entry points, torch tensors), each image equal to the file the
    reference's tool wrote: border init + JumpFlood(cw = 2) + each border
    image, unsigned init + JumpFlood(cw = 1) + the distance image, Chamfer and
    Felzenszwalb + their images."""
    g = np.load(os.path.join(GOLDEN, "image_golden.npz"))
    keys = sorted(k[:-5] for k in g.files if k.endswith("_dims"))
    dev = torch.device("cuda")
    for key in keys:
        dims = tuple(int(v) for v in g[f"{key}_dims"])
        sides, points, seed, method, itype, ml = (int(v) for v in g[f"{key}_gen"])
        img = g[f"{key}_img"]
        w, h = dims[0], dims[1]
        d = dims[2] if len(dims) == 3 else 0
        dd = max(d, 1)
        cells = w * h * dd
        m = _tool_mask(oracle, dims, sides, points, seed)
        m_t = torch.from_numpy(m.view(np.int32)).to(dev)
        max_len = float(ml) if ml >= 0 else 0.25 * max(w, h)
        if key.startswith("s"):
            nc = 3 if d else 2
            delta = torch.empty(cells * nc, dtype=torch.int16, device=dev)
            out = torch.empty(cells * (1 if itype == 0 else 4), dtype=torch.uint8, device=dev)
            if d:
                wk = torch.empty(gpu.workspace("jump_flood_3d", w, h, d), dtype=torch.uint8, device=dev)
                gpu.call_dev("signed_border_3d", w, h, d, m_t, delta, wk)
                gpu.call_dev("encode_border_image_3d", w, h, d, m_t, delta, 2 * max_len, itype, out)
            else:
                wk = torch.empty(gpu.workspace("jump_flood_2d", w, h), dtype=torch.uint8, device=dev)
                gpu.call_dev("signed_border_2d", w, h, m_t, delta, wk)
                gpu.call_dev("encode_border_image_2d", w, h, m_t, delta, 2 * max_len, itype, out)
        elif method == 2:
            nc = 3 if d else 2
            delta = torch.empty(cells * nc, dtype=torch.int16, device=dev)
            out = torch.empty(cells, dtype=torch.uint8, device=dev)
            if d:
                wk = torch.empty(gpu.workspace("jump_flood_3d", w, h, d), dtype=torch.uint8, device=dev)
                gpu.call_dev("init_deltas_3d", w, h, d, m_t, delta, 0)
                gpu.call_dev("jump_flood_3d", w, h, d, delta, 1, wk)
                gpu.call_dev("encode_distance_u8_3d", w, h, d, delta, max_len, out)
            else:
                wk = torch.empty(gpu.workspace("jump_flood_2d", w, h), dtype=torch.uint8, device=dev)
                gpu.call_dev("init_deltas_2d", w, h, m_t, delta, 0)
                gpu.call_dev("jump_flood_2d", w, h, delta, 1, wk)
                gpu.call_dev("encode_distance_u8_2d", w, h, delta, max_len, out)
        elif method == 4:
            bits = unpack_mask(m, w, h)
            dist = torch.from_numpy(np.where(bits, 0, 32767 - 16).astype(np.int16)).to(dev)
            mult = np.zeros(1, np.int32)
            gpu.call_dev("chamfer_i16", w, h, dist, mult)
            out = torch.empty(cells, dtype=torch.uint8, device=dev)
            gpu.call_dev("encode_scalar_u8_i16", w, h, dist, max_len * int(mult[0]), out)
        else:
            f = torch.empty(cells, dtype=torch.float32, device=dev)
            fw = torch.empty(max(256, gpu.workspace("felzenszwalb_2d", w, h)), dtype=torch.uint8, device=dev)
            gpu.call_dev("init_distances_f32", w, h, m_t, f, float(np.finfo(np.float32).max))
            gpu.call_dev("felzenszwalb_2d", w, h, f, fw)
            gpu.call_dev("sqrt_f32", w, h, f, f)
            out = torch.empty(cells, dtype=torch.uint8, device=dev)
            gpu.call_dev("encode_scalar_u8_f32", w, h, f, max_len, out)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.cpu().numpy().reshape(img.shape), img, err_msg=key)


def test_image_writers_reject_bad_arguments(gpu):
    import dolhip

    m = np.zeros(4, np.uint32)
    dl = np.zeros((4, 4, 2), np.int16)
    for call in (lambda: gpu.border_image(4, 4, 0, m, dl, 1.0, 4), lambda: gpu.border_image(0, 4, 0, m, dl, 1.0, 0),
                 lambda: gpu.distance_u8(4, 0, 0, dl, 1.0)):
        with pytest.raises(dolhip.DolHipError):
            call()
