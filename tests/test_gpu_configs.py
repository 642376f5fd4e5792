"""GPU parity at the benchmark configurations' full sizes.

Each test builds the configuration's seeded input with the library's host
generators, runs the GPU path through the C ABI on device buffers, and
compares every output with the oracle run on the same input (the oracle's
parallel loops use the host's cores, at most 16):

  8192^2 disc/box bitmask  signed border init + JumpFlood (cw = 2), int16x2
                           deltas, float signed distances, sdf byte image
  512^3 ellipsoid/box      signed 3D border init + JumpFlood (cw = 2), int16x3
                           deltas and float distances; exact squared EDT
                           (int32 and float)
  20,000 ragged glyphs     batched signed byte SDFs, per-image launches and
                           the oracle, byte for byte
  2048^2, 4096 seeds       nearest-seed index exact under float64, float
                           distances, seed gradient against the float64
                           oracle gradient and central differences (h = 1e-3)
"""
import sys
import time

import numpy as np
import pytest

from oracle_bind import host_threads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t

    return t


def _say(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


def _mismatch_report(got, exp, what):
    bad = np.argwhere(got != exp)
    if bad.size:
        first = tuple(bad[0])
        return f"{what}: {len(bad)} differ, first at {first}: {got[first]} vs {exp[first]}"
    return ""


# ------------------------------------------------------------ 8192^2 2D


def test_signed_border_8192(gpu, torch, oracle):
    w = h = 8192
    mask = gpu.scene_shapes_2d(w, h, 2000, seed=2)
    dev = torch.device("cuda")
    m_t = torch.from_numpy(mask.view(np.int32)).to(dev)
    delta = torch.empty(w * h * 2, dtype=torch.int16, device=dev)
    dist = torch.empty(w * h, dtype=torch.float32, device=dev)
    sdf = torch.empty(w * h, dtype=torch.uint8, device=dev)
    work = torch.empty(gpu.workspace("jump_flood_2d", w, h), dtype=torch.uint8, device=dev)
    _say("8192^2 signed border: GPU")
    gpu.call_dev("signed_border_2d", w, h, m_t, delta, work)
    gpu.call_dev("signed_distance_2d", w, h, m_t, delta, dist)
    gpu.call_dev("encode_sdf_u8_2d", w, h, m_t, delta, 0.5 * w, sdf)
    torch.cuda.synchronize()
    got = delta.cpu().numpy().reshape(h, w, 2)

    t = host_threads()
    _say(f"8192^2 signed border: oracle on {t} threads")
    t0 = time.perf_counter()
    exp = oracle.jfa_2d(w, h, oracle.border_2d(w, h, mask), 2, threads=t)
    _say(f"oracle {time.perf_counter() - t0:.1f} s")
    msg = _mismatch_report(got, exp, "deltas")
    assert not msg, msg
    np.testing.assert_array_equal(dist.cpu().numpy().reshape(h, w).view(np.uint32),
                                  oracle.signed_dist_2d(w, h, mask, exp).view(np.uint32))
    np.testing.assert_array_equal(sdf.cpu().numpy().reshape(h, w), oracle.sdf_u8_2d(w, h, mask, exp, 0.5 * w))


def test_two_field_8192(gpu, torch, oracle):
    """the -S mode's exterior and interior fields (cw = 1) on the same scene"""
    w = h = 8192
    mask = gpu.scene_shapes_2d(w, h, 2000, seed=2)
    dev = torch.device("cuda")
    m_t = torch.from_numpy(mask.view(np.int32)).to(dev)
    d0 = torch.empty(w * h * 2, dtype=torch.int16, device=dev)
    d1 = torch.empty(w * h * 2, dtype=torch.int16, device=dev)
    img = torch.empty(w * h, dtype=torch.uint8, device=dev)
    work = torch.empty(gpu.workspace("two_field_2d", w, h), dtype=torch.uint8, device=dev)
    gpu.call_dev("two_field_deltas_2d", w, h, m_t, d0, d1, 1, work)
    gpu.call_dev("encode_two_field_2d", w, h, d0, d1, 0.25 * w, 0, img)
    torch.cuda.synchronize()
    t = host_threads()
    e0 = oracle.jfa_2d(w, h, oracle.init_deltas_2d(w, h, mask, False), 1, threads=t)
    msg = _mismatch_report(d0.cpu().numpy().reshape(h, w, 2), e0, "exterior")
    assert not msg, msg
    e1 = oracle.jfa_2d(w, h, oracle.init_deltas_2d(w, h, mask, True), 1, threads=t)
    msg = _mismatch_report(d1.cpu().numpy().reshape(h, w, 2), e1, "interior")
    assert not msg, msg
    np.testing.assert_array_equal(img.cpu().numpy().reshape(h, w), oracle.two_field_image_2d(w, h, e0, e1, 0.25 * w, 0))


# ------------------------------------------------------------ 512^3 3D


@pytest.fixture(scope="module")
def scene512(gpu):
    n = 512
    return n, gpu.scene_shapes_3d(n, n, n, 300, 100, seed=3)


def test_signed_jfa_512cube(gpu, torch, oracle, scene512):
    n, mask = scene512
    dev = torch.device("cuda")
    m_t = torch.from_numpy(mask.view(np.int32)).to(dev)
    delta = torch.empty(n ** 3 * 3, dtype=torch.int16, device=dev)
    dist = torch.empty(n ** 3, dtype=torch.float32, device=dev)
    work = torch.empty(gpu.workspace("jump_flood_3d", n, n, n), dtype=torch.uint8, device=dev)
    _say("512^3 signed JumpFlood: GPU")
    gpu.call_dev("signed_border_3d", n, n, n, m_t, delta, work)
    gpu.call_dev("signed_distance_3d", n, n, n, m_t, delta, dist)
    torch.cuda.synchronize()
    del work
    got = delta.cpu().numpy().reshape(n, n, n, 3)
    got_dist = dist.cpu().numpy().reshape(n, n, n)
    del delta, dist
    torch.cuda.empty_cache()

    t = host_threads()
    _say(f"512^3 signed JumpFlood: oracle border init, then 10 passes on {t} threads")
    t0 = time.perf_counter()
    exp = oracle.border_3d(n, n, n, mask)
    exp = oracle.jfa_3d(n, n, n, exp, 2, threads=t)
    _say(f"oracle {time.perf_counter() - t0:.1f} s")
    msg = _mismatch_report(got, exp, "deltas")
    assert not msg, msg
    np.testing.assert_array_equal(got_dist.view(np.uint32), oracle.signed_dist_3d(n, n, n, mask, exp).view(np.uint32))


def test_edt_512cube(gpu, torch, oracle, scene512):
    n, mask = scene512
    dev = torch.device("cuda")
    m_t = torch.from_numpy(mask.view(np.int32)).to(dev)
    d2 = torch.empty(n ** 3, dtype=torch.int32, device=dev)
    fl = torch.empty(n ** 3, dtype=torch.float32, device=dev)
    work = torch.empty(gpu.workspace("edt_3d", n, n, n), dtype=torch.uint8, device=dev)
    gpu.call_dev("edt_3d", n, n, n, m_t, 0, d2, fl, work)
    torch.cuda.synchronize()
    got = d2.cpu().numpy().reshape(n, n, n)
    got_f = fl.cpu().numpy().reshape(n, n, n)
    exp = oracle.edt_3d(n, n, n, mask)
    msg = _mismatch_report(got, exp, "squared distances")
    assert not msg, msg
    # sqrtf of the exact value, correctly rounded (numpy's float32 sqrt is)
    exp_f = np.sqrt(np.where(exp == 0x7FFFFFFF, np.float32(np.finfo(np.float32).max), exp.astype(np.float32)))
    np.testing.assert_array_equal(got_f.view(np.uint32), exp_f.view(np.uint32))


# ------------------------------------------------------------ 20,000 glyphs


def test_glyphs_20000(gpu, torch, oracle):
    import dolhip

    gw, gh, goff, gdata = gpu.scene_glyphs(n=20000, min_size=16, max_size=256, seed=4)
    dev = torch.device("cuda")
    b = dolhip.Batch(gpu, gw, gh, goff, threshold=128, less_than=False)
    d_in = torch.from_numpy(gdata).to(dev)
    out_b = torch.empty_like(d_in)
    out_p = torch.empty_like(d_in)
    b.tune(d_in, out_b)
    out_b.fill_(0)
    b.run(d_in, out_b, graph=True)
    b.run_per_image(d_in, out_p)
    torch.cuda.synchronize()
    batched = out_b.cpu().numpy()
    per_image = out_p.cpu().numpy()
    b.close()
    t = host_threads()
    _say(f"20,000 glyphs: oracle on {t} threads")
    exp = oracle.glyph_sdf_batch(gw, gh, goff, gdata, 128, False, threads=t)
    assert np.array_equal(per_image, exp), _mismatch_report(per_image, exp, "per-image")
    assert np.array_equal(batched, exp), _mismatch_report(batched, exp, "batched")


# ------------------------------------------------------------ 2048^2, 4096 seeds


def _fd_local(oracle_dist, seeds, wt, k, c, eps):
    """Central difference of sum(w * d) in seed k's coordinate c, in float64,
    over the only cells whose nearest distance can change: those no further
    from seed k than their current nearest distance plus 2 eps.  Also returns
    whether any of them is within 2 eps of a tie between seed k and another
    seed: there the loss has a kink inside [p - eps, p + eps] and the central
    difference averages two slopes."""
    h, w = oracle_dist.shape
    s64 = seeds.astype(np.float64)
    px, py = s64[k]
    y, x = np.indices((h, w))
    near = np.hypot(x - px, y - py) <= oracle_dist + 2 * eps
    cy, cx = np.nonzero(near)
    others = np.delete(s64, k, axis=0)
    d_others = np.full(cx.size, np.inf)
    for s0 in range(0, others.shape[0], 512):
        blk = others[s0:s0 + 512]
        d = np.hypot(cx[:, None] - blk[None, :, 0], cy[:, None] - blk[None, :, 1]).min(axis=1)
        d_others = np.minimum(d_others, d)
    wts = wt[cy, cx].astype(np.float64)
    pp, pm = s64[k].copy(), s64[k].copy()
    pp[c] += eps
    pm[c] -= eps
    lp = (wts * np.minimum(d_others, np.hypot(cx - pp[0], cy - pp[1]))).sum()
    lm = (wts * np.minimum(d_others, np.hypot(cx - pm[0], cy - pm[1]))).sum()
    kink = bool((np.abs(d_others - np.hypot(cx - px, cy - py)) <= 2 * eps).any())
    return (lp - lm) / (2 * eps), kink


def test_seeds_2048(gpu, torch, oracle):
    w = h = 2048
    n = 4096
    seeds, wt = gpu.scene_seeds(w, h, n, seed=5)
    dev = torch.device("cuda")
    s_t = torch.from_numpy(seeds).to(dev)
    idx = torch.empty((h, w), dtype=torch.int32, device=dev)
    dist = torch.empty((h, w), dtype=torch.float32, device=dev)
    grad = torch.empty((n, 2), dtype=torch.float32, device=dev)
    work = torch.empty(gpu.workspace("seed", w, h, n), dtype=torch.uint8, device=dev)
    gpu.call_dev("seed_nearest_2d", w, h, n, s_t, idx, dist, work)
    gpu.call_dev("seed_grad_2d", w, h, n, s_t, idx, dist, torch.from_numpy(wt).to(dev), grad)
    torch.cuda.synchronize()
    t = host_threads()
    _say(f"2048^2 seeds: float64 oracle on {t} threads")
    eidx, edist = oracle.seed_nearest_2d(w, h, seeds, threads=t)
    gi = idx.cpu().numpy()
    msg = _mismatch_report(gi, eidx, "nearest-seed index")
    assert not msg, msg
    gd = dist.cpu().numpy().astype(np.float64)
    # float32 distance of a float64 distance below 2900: within one float32 ulp
    assert np.all(np.abs(gd - edist) <= np.spacing(edist.astype(np.float32)).astype(np.float64))
    eg = oracle.seed_grad_2d(w, h, seeds, eidx, wt)
    g = grad.cpu().numpy().astype(np.float64)
    rel = np.abs(g - eg).max() / np.abs(eg).max()
    _say(f"seed gradient: max |g - g64| / max |g64| = {rel:.3e}")
    assert rel <= 1e-4
    # central differences at seeds whose loss is smooth within +-eps
    eps = 1e-3
    checked = 0
    for k in range(0, n, 97):
        for c in range(2):
            fd, kink = _fd_local(edist, seeds, wt, k, c, eps)
            if kink:
                continue
            checked += 1
            assert abs(fd - g[k, c]) <= 1e-3 * max(1.0, abs(fd)), (k, c, fd, g[k, c])
    assert checked >= 40, checked
