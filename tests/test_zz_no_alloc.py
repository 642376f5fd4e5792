"""The allocation-free JumpFlood 2D entry point (dolhip_jump_flood_2d_noalloc)
against dolhip_jump_flood_2d_dev, called through the C ABI: two calls in a
row on one device, the second on a smaller field (the scratch array keeps the
first call's contents), bit-exact; and the size limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _field(gpu, oracle, w, h, n, seed, cw):
    from oracle_bind import empty_mask

    m = empty_mask(w, h)
    oracle.mask_add_points_2d(w, h, n, seed, m)
    return oracle.border_2d(w, h, m) if cw == 2 else oracle.init_deltas_2d(w, h, m)


def test_noalloc_matches_dev_twice(gpu, oracle):
    import torch

    L = gpu.L
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for (w, h, n, seed, cw) in [(512, 512, 100, 12345, 1), (300, 200, 40, 7, 2)]:
        start = _field(gpu, oracle, w, h, n, seed, cw)
        a = torch.from_numpy(start.reshape(-1).copy()).cuda()
        b = a.clone()
        work = torch.empty(gpu.workspace("jump_flood_2d", w, h), dtype=torch.uint8, device="cuda")
        assert L.dolhip_jump_flood_2d_dev(w, h, ctypes.c_void_p(a.data_ptr()), cw,
                                          ctypes.c_void_p(work.data_ptr()), stream) == 0, gpu.last_error()
        assert L.dolhip_jump_flood_2d_noalloc(w, h, ctypes.c_void_p(b.data_ptr()), cw, stream) == 0, \
            gpu.last_error()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(b.cpu().numpy(), a.cpu().numpy(), err_msg=f"{w}x{h} cw={cw}")
        np.testing.assert_array_equal(b.cpu().numpy().reshape(h, w, 2), oracle.jfa_2d(w, h, start, cw))


def test_noalloc_rejects_large_fields(gpu):
    import torch

    d = torch.empty(4, dtype=torch.int16, device="cuda")   # never touched: the size is checked first
    rc = gpu.L.dolhip_jump_flood_2d_noalloc(2049, 2048, ctypes.c_void_p(d.data_ptr()), 1, None)
    assert rc == 1 and "scratch" in gpu.last_error()
