"""GPU parity of the batched glyph SDFs and the continuous-seed distance.

Batched glyphs: every kernel variant of the persistent launch, the hipGraph
replay (several replays with eager work between them) and the per-image
launches must give the bytes the oracle gives for each mask: threshold,
border init, JumpFlood in half cells, EncodeU8Signed with
maxLen = 0.5 max(w, h).

Seeds: the nearest-seed index must equal the float64 brute force exactly
(ties to the lowest index); distances match float64 to float32 rounding;
gradients match the float64 oracle gradient (relative 1e-4 of the largest
component, float32 accumulation in varying order) and central finite
differences of the float64 loss.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t

    return t


def glyph_expected(oracle, w, h, data, off):
    out = np.empty(data.size, np.uint8)
    for i in range(len(w)):
        ww, hh, o = int(w[i]), int(h[i]), int(off[i])
        g = data[o:o + ww * hh]
        m = oracle.mask_from_bytes(ww, hh, g, 128, False)
        d = oracle.jfa_2d(ww, hh, oracle.border_2d(ww, hh, m), 2)
        out[o:o + ww * hh] = oracle.sdf_u8_2d(ww, hh, m, d, 0.5 * max(ww, hh)).reshape(-1)
    return out


@pytest.fixture(scope="module")
def glyphs(gpu, oracle):
    w, h, off, data = gpu.scene_glyphs(n=160, min_size=16, max_size=256, seed=21)
    # a few edge shapes: thin, single row/column, the largest
    return w, h, off, data, glyph_expected(oracle, w, h, data, off)


def test_batch_all_variants(gpu, torch, glyphs):
    import dolhip

    w, h, off, data, exp = glyphs
    b = dolhip.Batch(gpu, w, h, off, threshold=128, less_than=False)
    d_in = torch.from_numpy(data).cuda()
    d_out = torch.empty_like(d_in)
    for v in range(b.variants()):
        b.set_variant(v)
        d_out.fill_(7)
        b.run(d_in, d_out)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(d_out.cpu().numpy(), exp, err_msg=f"variant {v}")
    b.close()


def test_batch_per_image_and_graph(gpu, torch, glyphs):
    import dolhip

    w, h, off, data, exp = glyphs
    b = dolhip.Batch(gpu, w, h, off, threshold=128, less_than=False)
    d_in = torch.from_numpy(data).cuda()
    d_out = torch.empty_like(d_in)
    b.run_per_image(d_in, d_out)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d_out.cpu().numpy(), exp)
    b.tune(d_in, d_out)
    cells, waves = b.config()
    assert cells > 0 and waves > 0
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        for rep in range(4):
            d_out.fill_(0)                       # eager work between replays
            b.run(d_in, d_out, graph=True, stream=s)
            s.synchronize()
            np.testing.assert_array_equal(d_out.cpu().numpy(), exp, err_msg=f"replay {rep}")
    # graph mode from the default stream too
    d_out.fill_(1)
    b.run(d_in, d_out, graph=True)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d_out.cpu().numpy(), exp)
    b.close()


def test_batch_edge_sizes(gpu, torch, oracle):
    import dolhip

    rng = np.random.default_rng(3)
    sizes = [(1, 1), (1, 256), (256, 1), (16, 16), (33, 31), (256, 256), (255, 200), (2, 3)]
    w = np.array([s[0] for s in sizes], np.int32)
    h = np.array([s[1] for s in sizes], np.int32)
    off = np.concatenate([[0], np.cumsum(w.astype(np.int64) * h)[:-1]])
    data = (rng.random(int((w.astype(np.int64) * h).sum())) < 0.3).astype(np.uint8) * 255
    exp = glyph_expected(oracle, w, h, data, off)
    b = dolhip.Batch(gpu, w, h, off, threshold=128, less_than=False)
    d_in = torch.from_numpy(data).cuda()
    d_out = torch.empty_like(d_in)
    for v in range(b.variants()):
        b.set_variant(v)
        b.run(d_in, d_out)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(d_out.cpu().numpy(), exp, err_msg=f"variant {v}")
    b.close()


def test_batch_many_small_and_uniform(gpu, torch, oracle):
    """Hundreds of small masks share each bin (several to a wavefront's
    groups), with empty and full masks among them (no border cell: every cell
    keeps the reference's wrapped MAX values and encodes to 1 or 255), and
    the dark-is-inside threshold."""
    import dolhip

    rng = np.random.default_rng(11)
    n = 700
    w = rng.integers(1, 48, n).astype(np.int32)
    h = rng.integers(1, 48, n).astype(np.int32)
    off = np.concatenate([[0], np.cumsum(w.astype(np.int64) * h)[:-1]])
    data = rng.integers(0, 256, int((w.astype(np.int64) * h).sum())).astype(np.uint8)
    for k in range(0, n, 7):   # every 7th mask uniform
        data[off[k]:off[k] + w[k] * h[k]] = 255 if k % 2 else 0
    for less_than in (False, True):
        exp = np.empty(data.size, np.uint8)
        for i in range(n):
            ww, hh, o = int(w[i]), int(h[i]), int(off[i])
            m = oracle.mask_from_bytes(ww, hh, data[o:o + ww * hh], 128, less_than)
            d = oracle.jfa_2d(ww, hh, oracle.border_2d(ww, hh, m), 2)
            exp[o:o + ww * hh] = oracle.sdf_u8_2d(ww, hh, m, d, 0.5 * max(ww, hh)).reshape(-1)
        b = dolhip.Batch(gpu, w, h, off, threshold=128, less_than=less_than)
        d_in = torch.from_numpy(data).cuda()
        d_out = torch.empty_like(d_in)
        for v in range(b.variants()):
            b.set_variant(v)
            d_out.fill_(7)
            b.run(d_in, d_out)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(d_out.cpu().numpy(), exp, err_msg=f"variant {v} less_than {less_than}")
        b.close()


def test_batch_rejects_bad_sizes(gpu):
    import dolhip

    with pytest.raises(dolhip.DolHipError):
        dolhip.Batch(gpu, [300], [10], [0])


# ------------------------------------------------------------ seeds


def _seed_case(gpu, w, h, n, seed):
    s, wt = gpu.scene_seeds(w, h, n, seed=seed)
    return s, wt


@pytest.mark.parametrize("w,h,n", [(64, 48, 1), (96, 80, 37), (256, 192, 300), (300, 20, 50)])
def test_seed_nearest_exact(gpu, torch, oracle, w, h, n):
    s, _ = _seed_case(gpu, w, h, n, seed=w + n)
    st = torch.from_numpy(s).cuda()
    idx = torch.empty((h, w), dtype=torch.int32, device="cuda")
    dist = torch.empty((h, w), dtype=torch.float32, device="cuda")
    work = torch.empty(gpu.workspace("seed", w, h, n), dtype=torch.uint8, device="cuda")
    gpu.call_dev("seed_nearest_2d", w, h, n, st, idx, dist, work)
    torch.cuda.synchronize()
    eidx, edist = oracle.seed_nearest_2d(w, h, s)
    np.testing.assert_array_equal(idx.cpu().numpy(), eidx)
    np.testing.assert_allclose(dist.cpu().numpy(), edist, rtol=1e-6, atol=1e-6)


def test_seed_ties_go_to_lowest_index(gpu, torch, oracle):
    # duplicated seeds and seeds on exact midpoints
    w, h = 40, 30
    s = np.array([[10, 10], [10, 10], [20, 10], [15, 10], [15, 10], [0, 0], [39, 29]], np.float32)
    st = torch.from_numpy(s).cuda()
    idx = torch.empty((h, w), dtype=torch.int32, device="cuda")
    dist = torch.empty((h, w), dtype=torch.float32, device="cuda")
    work = torch.empty(gpu.workspace("seed", w, h, len(s)), dtype=torch.uint8, device="cuda")
    gpu.call_dev("seed_nearest_2d", w, h, len(s), st, idx, dist, work)
    torch.cuda.synchronize()
    eidx, _ = oracle.seed_nearest_2d(w, h, s)
    np.testing.assert_array_equal(idx.cpu().numpy(), eidx)


def test_seed_grad_vs_oracle(gpu, torch, oracle):
    w, h, n = 256, 192, 300
    s, wt = _seed_case(gpu, w, h, n, seed=9)
    st = torch.from_numpy(s).cuda()
    idx = torch.empty((h, w), dtype=torch.int32, device="cuda")
    dist = torch.empty((h, w), dtype=torch.float32, device="cuda")
    work = torch.empty(gpu.workspace("seed", w, h, n), dtype=torch.uint8, device="cuda")
    gpu.call_dev("seed_nearest_2d", w, h, n, st, idx, dist, work)
    grad = torch.empty((n, 2), dtype=torch.float32, device="cuda")
    gout = torch.from_numpy(wt).cuda()
    gpu.call_dev("seed_grad_2d", w, h, n, st, idx, dist, gout, grad)
    torch.cuda.synchronize()
    eidx, _ = oracle.seed_nearest_2d(w, h, s)
    eg = oracle.seed_grad_2d(w, h, s, eidx, wt)
    g = grad.cpu().numpy().astype(np.float64)
    scale = np.abs(eg).max()
    assert np.abs(g - eg).max() <= 1e-4 * scale


def test_seed_autograd_function(gpu, torch, oracle):
    """The libtorch autograd Function: backward vs the float64 oracle gradient
    and vs central finite differences (h = 1e-3) of the float64 loss."""
    import build

    ext = build.build_torch_ext()
    w, h, n = 128, 96, 64
    s, wt = _seed_case(gpu, w, h, n, seed=17)
    st = torch.from_numpy(s).cuda().requires_grad_(True)
    wt_t = torch.from_numpy(wt).cuda()
    dist, idx = ext.seed_distance(st, w, h)
    loss = (wt_t * dist).sum()
    loss.backward()
    g = st.grad.cpu().numpy().astype(np.float64)
    eidx, edist = oracle.seed_nearest_2d(w, h, s)
    np.testing.assert_array_equal(idx.cpu().numpy(), eidx)
    eg = oracle.seed_grad_2d(w, h, s, eidx, wt)
    assert np.abs(g - eg).max() <= 1e-4 * np.abs(eg).max()

    y, x = np.indices((h, w))

    def loss64(seeds):
        d2 = (x[..., None] - seeds[:, 0]) ** 2 + (y[..., None] - seeds[:, 1]) ** 2
        return float((wt.astype(np.float64) * np.sqrt(d2.min(-1))).sum())

    s64 = s.astype(np.float64)
    eps = 1e-3
    for k in range(0, n, 7):
        for c in range(2):
            sp, sm = s64.copy(), s64.copy()
            sp[k, c] += eps
            sm[k, c] -= eps
            fd = (loss64(sp) - loss64(sm)) / (2 * eps)
            assert abs(fd - g[k, c]) <= 2e-2 * max(1.0, abs(fd)), (k, c, fd, g[k, c])
