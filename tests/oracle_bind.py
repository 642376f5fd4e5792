"""ctypes bindings for the test oracle (oracle/liboracle.so) and, when it was
built, the compiled reference (oracle/_ref/libdolref.so).

Test infrastructure only: tests, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module as the checker; the product library never
does.  Arrays are numpy; delta fields are int16 arrays of shape (..., 2) or
(..., 3), which is the cCellDelta2/cCellDelta3 layout
(DistanceOcclusionLib.h:41-53).
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libdolref.so")
REF_OMP_SO = os.path.join(ROOT, "oracle", "_ref", "libdolref_omp.so")
MESH_REF_SO = os.path.join(ROOT, "oracle", "_ref", "libmeshref.so")

_P = ctypes.c_void_p
_I = ctypes.c_int
_U = ctypes.c_uint32
_F = ctypes.c_float


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def mask_stride(w):
    return (w + 31) // 32


def empty_mask(*dims):
    """dims = (w, h) or (w, h, d); returns a zeroed uint32 mask."""
    w = dims[0]
    rows = int(np.prod(dims[1:]))
    return np.zeros(mask_stride(w) * rows, dtype=np.uint32)


def unpack_mask(mask, *dims):
    """bitmask -> bool array of shape (d, h, w) or (h, w)."""
    w = dims[0]
    rows = int(np.prod(dims[1:]))
    ws = mask_stride(w)
    m = mask.reshape(rows, ws)
    bits = ((m[:, :, None] >> np.arange(32, dtype=np.uint32)) & 1).astype(bool)
    bits = bits.reshape(rows, ws * 32)[:, :w]
    return bits.reshape(tuple(reversed(dims)))


def pack_mask(bits):
    """bool array (..., w) -> LSB-first padded uint32 words."""
    bits = np.asarray(bits, dtype=bool)
    w = bits.shape[-1]
    rows = bits.reshape(-1, w)
    ws = mask_stride(w)
    pad = np.zeros((rows.shape[0], ws * 32), dtype=np.uint64)
    pad[:, :w] = rows
    pad = pad.reshape(rows.shape[0], ws, 32)
    words = (pad << np.arange(32, dtype=np.uint64)).sum(axis=2).astype(np.uint32)
    return words.reshape(-1)


class Oracle:
    """The CPU restatement (oracle/dol_oracle.c)."""

    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (run make -C oracle)")
        L = ctypes.CDLL(path)
        self.L = L
        L.oc_edt_float.restype = ctypes.c_float
        L.oc_edt_float.argtypes = [ctypes.c_int32]
        L.oc_encode_u8.restype = ctypes.c_uint8
        L.oc_encode_u8.argtypes = [_F]
        L.oc_encode_u8_signed.restype = ctypes.c_uint8
        L.oc_encode_u8_signed.argtypes = [_F]
        L.oc_check_deltas_2d.restype = ctypes.c_long
        L.oc_check_deltas_3d.restype = ctypes.c_long
        L.oc_floor_pow2.restype = _I
        L.oc_rng.restype = _U

    # generators
    def mask_add_points_2d(self, w, h, n, seed, mask):
        self.L.oc_mask_add_points_2d(_I(w), _I(h), _I(n), _U(seed), _ptr(mask))

    def mask_add_points_3d(self, w, h, d, n, seed, mask):
        self.L.oc_mask_add_points_3d(_I(w), _I(h), _I(d), _I(n), _U(seed), _ptr(mask))

    def mask_add_block_2d(self, w, h, sides, mask):
        self.L.oc_mask_add_block_2d(_I(w), _I(h), _I(sides), _ptr(mask))

    def mask_add_block_3d(self, w, h, d, sides, mask):
        self.L.oc_mask_add_block_3d(_I(w), _I(h), _I(d), _I(sides), _ptr(mask))

    def mask_from_bytes(self, w, rows, data, threshold=128, less_than=True):
        data = np.ascontiguousarray(data, dtype=np.uint8)
        mask = np.zeros(mask_stride(w) * rows, dtype=np.uint32)
        self.L.oc_mask_from_bytes(_I(w), _I(rows), _ptr(data), _I(threshold), _I(int(less_than)), _ptr(mask))
        return mask

    # init
    def init_deltas_2d(self, w, h, mask, invert=False):
        out = np.empty((h, w, 2), dtype=np.int16)
        self.L.oc_init_deltas_2d(_I(w), _I(h), _ptr(mask), _ptr(out), _I(int(invert)))
        return out

    def init_deltas_3d(self, w, h, d, mask, invert=False):
        out = np.empty((d, h, w, 3), dtype=np.int16)
        self.L.oc_init_deltas_3d(_I(w), _I(h), _I(d), _ptr(mask), _ptr(out), _I(int(invert)))
        return out

    def init_distances_f32(self, w, h, mask, max_d=np.finfo(np.float32).max):
        out = np.empty((h, w), dtype=np.float32)
        self.L.oc_init_distances_f32(_I(w), _I(h), _ptr(mask), _ptr(out), _F(max_d))
        return out

    def border_2d(self, w, h, mask):
        out = np.empty((h, w, 2), dtype=np.int16)
        self.L.oc_border_2d(_I(w), _I(h), _ptr(mask), _ptr(out))
        return out

    def border_3d(self, w, h, d, mask):
        out = np.empty((d, h, w, 3), dtype=np.int16)
        self.L.oc_border_3d(_I(w), _I(h), _I(d), _ptr(mask), _ptr(out))
        return out

    # propagation (in place on a copy)
    def jfa_2d(self, w, h, delta, cw=1, threads=1):
        out = np.ascontiguousarray(delta, dtype=np.int16).copy()
        self.L.oc_jfa_2d_mt(_I(w), _I(h), _ptr(out), _I(cw), _I(threads))
        return out

    def jfa_3d(self, w, h, d, delta, cw=1, threads=1):
        out = np.ascontiguousarray(delta, dtype=np.int16).copy()
        self.L.oc_jfa_3d_mt(_I(w), _I(h), _I(d), _ptr(out), _I(cw), _I(threads))
        return out

    def fast_sweep_2d(self, w, h, delta, cw=1):
        out = np.ascontiguousarray(delta, dtype=np.int16).copy()
        self.L.oc_fast_sweep_2d(_I(w), _I(h), _ptr(out), _I(cw))
        return out

    def fast_sweep_3d(self, w, h, d, delta, cw=1):
        out = np.ascontiguousarray(delta, dtype=np.int16).copy()
        self.L.oc_fast_sweep_3d(_I(w), _I(h), _I(d), _ptr(out), _I(cw))
        return out

    def init_dirw_2d(self, w, h, mask):
        out = np.empty((4, h, w), dtype=np.float32)
        self.L.oc_init_dirw_2d(_I(w), _I(h), _ptr(mask), _ptr(out))
        return out

    def init_dirw_3d(self, w, h, d, mask):
        out = np.empty((8, d, h, w), dtype=np.float32)
        self.L.oc_init_dirw_3d(_I(w), _I(h), _I(d), _ptr(mask), _ptr(out))
        return out

    def occlusion_sweep_2d(self, w, h, dirw):
        out = np.ascontiguousarray(dirw, dtype=np.float32).copy()
        self.L.oc_occlusion_sweep_2d(_I(w), _I(h), _ptr(out))
        return out

    def occlusion_sweep_3d(self, w, h, d, dirw):
        out = np.ascontiguousarray(dirw, dtype=np.float32).copy()
        self.L.oc_occlusion_sweep_3d(_I(w), _I(h), _I(d), _ptr(out))
        return out

    def occlusion_sweep_3d_from(self, w, h, d, dirw_in, dirw_out):
        src = np.ascontiguousarray(dirw_in, dtype=np.float32)
        out = np.ascontiguousarray(dirw_out, dtype=np.float32).copy()
        self.L.oc_occlusion_sweep_3d_from(_I(w), _I(h), _I(d), _ptr(src), _ptr(out))
        return out

    def encode_ao(self, dirw, dirw_mode=False):
        d = np.ascontiguousarray(dirw, dtype=np.float32)
        planes = d.shape[0]
        n = d.size // planes
        out = np.empty((n, 4) if dirw_mode else n, dtype=np.uint8)
        fn = self.L.oc_encode_dirw if dirw_mode else self.L.oc_encode_ao
        fn(ctypes.c_size_t(n), _I(planes), _ptr(d), _ptr(out))
        return out.reshape(d.shape[1:] + ((4,) if dirw_mode else ()))

    def jfa_pass_2d(self, step, w, h, delta, cw=1):
        src = np.ascontiguousarray(delta, dtype=np.int16)
        out = np.empty_like(src)
        self.L.oc_jfa_pass_2d(_I(step), _I(w), _I(h), _ptr(src), _ptr(out), _I(cw))
        return out

    def jfa_pass_3d(self, step, w, h, d, delta, cw=1):
        src = np.ascontiguousarray(delta, dtype=np.int16)
        out = np.empty_like(src)
        self.L.oc_jfa_pass_3d(_I(step), _I(w), _I(h), _I(d), _ptr(src), _ptr(out), _I(cw))
        return out

    def felzenszwalb_2d(self, w, h, dist):
        out = np.ascontiguousarray(dist, dtype=np.float32).copy()
        self.L.oc_felzenszwalb_2d(_I(w), _I(h), _ptr(out))
        return out

    def brute_force_2d(self, w, h, delta, cw=1):
        out = np.ascontiguousarray(delta, dtype=np.int16).copy()
        self.L.oc_brute_force_2d(_I(w), _I(h), _ptr(out), _I(cw))
        return out

    def brute_force_3d(self, w, h, d, delta, cw=1):
        out = np.ascontiguousarray(delta, dtype=np.int16).copy()
        self.L.oc_brute_force_3d(_I(w), _I(h), _I(d), _ptr(out), _I(cw))
        return out

    # exact EDT
    def edt_2d(self, w, h, mask, invert=False):
        out = np.empty((h, w), dtype=np.int32)
        self.L.oc_edt_2d(_I(w), _I(h), _ptr(mask), _I(int(invert)), _ptr(out))
        return out

    def edt_3d(self, w, h, d, mask, invert=False):
        out = np.empty((d, h, w), dtype=np.int32)
        self.L.oc_edt_3d(_I(w), _I(h), _I(d), _ptr(mask), _I(int(invert)), _ptr(out))
        return out

    def edt_brute_2d(self, w, h, mask, invert=False):
        out = np.empty((h, w), dtype=np.int32)
        self.L.oc_edt_brute_2d(_I(w), _I(h), _ptr(mask), _I(int(invert)), _ptr(out))
        return out

    def edt_brute_3d(self, w, h, d, mask, invert=False):
        out = np.empty((d, h, w), dtype=np.int32)
        self.L.oc_edt_brute_3d(_I(w), _I(h), _I(d), _ptr(mask), _I(int(invert)), _ptr(out))
        return out

    # encoders
    def signed_dist_2d(self, w, h, mask, delta):
        out = np.empty((h, w), dtype=np.float32)
        self.L.oc_signed_dist_2d(_I(w), _I(h), _ptr(mask), _ptr(np.ascontiguousarray(delta)), _ptr(out))
        return out

    def signed_dist_3d(self, w, h, d, mask, delta):
        out = np.empty((d, h, w), dtype=np.float32)
        self.L.oc_signed_dist_3d(_I(w), _I(h), _I(d), _ptr(mask), _ptr(np.ascontiguousarray(delta)), _ptr(out))
        return out

    def sdf_u8_2d(self, w, h, mask, delta, max_len):
        out = np.empty((h, w), dtype=np.uint8)
        self.L.oc_sdf_u8_2d(_I(w), _I(h), _ptr(mask), _ptr(np.ascontiguousarray(delta)), _F(max_len), _ptr(out))
        return out

    def sdf_u8_3d(self, w, h, d, mask, delta, max_len):
        out = np.empty((d, h, w), dtype=np.uint8)
        self.L.oc_sdf_u8_3d(_I(w), _I(h), _I(d), _ptr(mask), _ptr(np.ascontiguousarray(delta)), _F(max_len), _ptr(out))
        return out

    def sdf_delta_rgba_2d(self, w, h, delta, max_len):
        out = np.empty((h, w, 4), dtype=np.uint8)
        self.L.oc_sdf_delta_rgba_2d(_I(w), _I(h), _ptr(np.ascontiguousarray(delta)), _F(max_len), _ptr(out))
        return out

    def sdf_rg_rgba_2d(self, w, h, mask, delta, max_len):
        out = np.empty((h, w, 4), dtype=np.uint8)
        self.L.oc_sdf_rg_rgba_2d(_I(w), _I(h), _ptr(mask), _ptr(np.ascontiguousarray(delta)), _F(max_len), _ptr(out))
        return out

    def check_border_deltas_2d(self, w, h, mask, delta):
        f = self.L.oc_check_border_deltas_2d
        f.restype = ctypes.c_long
        return int(f(_I(w), _I(h), _ptr(np.ascontiguousarray(mask, np.uint32)),
                     _ptr(np.ascontiguousarray(delta, np.int16))))

    def check_border_deltas_3d(self, w, h, d, mask, delta):
        f = self.L.oc_check_border_deltas_3d
        f.restype = ctypes.c_long
        return int(f(_I(w), _I(h), _I(d), _ptr(np.ascontiguousarray(mask, np.uint32)),
                     _ptr(np.ascontiguousarray(delta, np.int16))))

    def border_image(self, w, h, d, mask, delta, max_len, image_type):
        """the border method's image (type 0 sdf, 1 sdf-delta, 2 sdf-rg, 3
        sdf-angles); d = 0 for a 2D field"""
        dims = 2 if d == 0 else 3
        dd = max(d, 1)
        shape = (dd, h, w) if image_type == 0 else (dd, h, w, 4)
        out = np.empty(shape, dtype=np.uint8)
        self.L.oc_border_image(_I(w), _I(h), _I(dd), _I(dims), _ptr(np.ascontiguousarray(mask, np.uint32)),
                               _ptr(np.ascontiguousarray(delta, np.int16)), _F(max_len), _I(image_type), _ptr(out))
        return out[0] if d == 0 else out

    def distance_u8(self, dims, delta, max_len):
        delta = np.ascontiguousarray(delta, np.int16)
        n = delta.size // dims
        out = np.empty(n, dtype=np.uint8)
        self.L.oc_distance_u8(ctypes.c_size_t(n), _I(dims), _ptr(delta), _F(max_len), _ptr(out))
        return out

    def scalar_u8(self, values, max_len):
        v = np.ascontiguousarray(values)
        out = np.empty(v.size, dtype=np.uint8)
        fn = {np.dtype(np.float32): self.L.oc_scalar_u8_f32, np.dtype(np.int32): self.L.oc_scalar_u8_i32,
              np.dtype(np.int16): self.L.oc_scalar_u8_i16}[v.dtype]
        fn(ctypes.c_size_t(v.size), _ptr(v), _F(max_len), _ptr(out))
        return out

    def sqrt_f32(self, values):
        v = np.ascontiguousarray(values, np.float32)
        out = np.empty_like(v)
        self.L.oc_sqrt_f32(ctypes.c_size_t(v.size), _ptr(v), _ptr(out))
        return out

    def two_field_image_2d(self, w, h, d0, d1, max_len, image_type):
        out = np.empty((h, w) if image_type == 0 else (h, w, 4), dtype=np.uint8)
        self.L.oc_two_field_image_2d(_I(w), _I(h), _ptr(np.ascontiguousarray(d0, np.int16)),
                                     _ptr(np.ascontiguousarray(d1, np.int16)), _F(max_len), _I(image_type), _ptr(out))
        return out

    def two_field_image_3d(self, w, h, d, d0, d1, max_len, image_type):
        out = np.empty((d, h, w) if image_type == 0 else (d, h, w, 4), dtype=np.uint8)
        self.L.oc_two_field_image_3d(_I(w), _I(h), _I(d), _ptr(np.ascontiguousarray(d0, np.int16)),
                                     _ptr(np.ascontiguousarray(d1, np.int16)), _F(max_len), _I(image_type), _ptr(out))
        return out

    def mask_from_triangles(self, w, h, d, vertices, indices, bbox, mask=None):
        """ORs the surface voxels of the triangles into mask (a new one if None)"""
        if mask is None:
            mask = empty_mask(w, h, d)
        vertices = np.ascontiguousarray(vertices, np.float32)
        idx = None if indices is None else np.ascontiguousarray(indices, np.int32)
        n = (vertices.size // 9) if idx is None else idx.size // 3
        self.L.oc_mask_from_triangles(_I(n), _ptr(idx), _ptr(vertices),
                                      _ptr(np.ascontiguousarray(bbox, np.float32)), _I(w), _I(h), _I(d), _ptr(mask))
        return mask

    def danielsson_2d(self, w, h, delta, cw=1):
        out = np.ascontiguousarray(delta, dtype=np.int16).copy()
        self.L.oc_danielsson_2d(_I(w), _I(h), _ptr(out), _I(cw))
        return out

    def danielsson_3d(self, w, h, d, delta, cw=1):
        out = np.ascontiguousarray(delta, dtype=np.int16).copy()
        self.L.oc_danielsson_3d(_I(w), _I(h), _I(d), _ptr(out), _I(cw))
        return out

    def chamfer(self, w, h, dist):
        """(result, divisor) of Chamfer on an int16 or int32 field"""
        out = np.ascontiguousarray(dist).copy()
        fn = self.L.oc_chamfer_i16 if out.dtype == np.int16 else self.L.oc_chamfer_i32
        r = fn(_I(w), _I(h), _ptr(out))
        return out, r

    def init_occ(self, w, h, mask, scale=1.0):
        out = np.empty((h, w), dtype=np.float32)
        self.L.oc_init_occ(_I(w), _I(h), _ptr(mask), _ptr(out), _F(scale))
        return out

    def shadow_sweep(self, gx, gy, w, h, occ):
        occ = np.ascontiguousarray(occ, dtype=np.float32)
        out = np.empty_like(occ)
        self.L.oc_shadow_sweep(_F(gx), _F(gy), _I(w), _I(h), _ptr(occ), _ptr(out))
        return out

    def dist_2d(self, delta):
        delta = np.ascontiguousarray(delta, dtype=np.int16)
        n = delta.size // 2
        out = np.empty(delta.shape[:-1], dtype=np.float32)
        self.L.oc_dist_2d(_I(n), _ptr(delta), _ptr(out))
        return out

    def check_deltas_2d(self, w, h, delta, cw=1):
        return self.L.oc_check_deltas_2d(_I(w), _I(h), _ptr(np.ascontiguousarray(delta)), _I(cw))

    def check_deltas_3d(self, w, h, d, delta, cw=1):
        return self.L.oc_check_deltas_3d(_I(w), _I(h), _I(d), _ptr(np.ascontiguousarray(delta)), _I(cw))

    def seed_nearest_2d(self, w, h, seeds, threads=1):
        seeds = np.ascontiguousarray(seeds, dtype=np.float32)
        idx = np.empty((h, w), dtype=np.int32)
        dist = np.empty((h, w), dtype=np.float64)
        self.L.oc_seed_nearest_2d_mt(_I(w), _I(h), _I(seeds.shape[0]), _ptr(seeds), _ptr(idx), _ptr(dist),
                                     _I(threads))
        return idx, dist

    def glyph_sdf_batch(self, widths, heights, offsets, data, threshold=128, less_than=False, threads=1):
        """signed byte SDF of every ragged glyph mask (threshold, border init, JumpFlood cw = 2, encode)"""
        w = np.ascontiguousarray(widths, dtype=np.int32)
        h = np.ascontiguousarray(heights, dtype=np.int32)
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        data = np.ascontiguousarray(data, dtype=np.uint8)
        out = np.empty_like(data)
        self.L.oc_glyph_sdf_batch(_I(len(w)), _ptr(w), _ptr(h), _ptr(off), _ptr(data), _I(threshold),
                                  _I(int(less_than)), _ptr(out), _I(threads))
        return out

    def seed_grad_2d(self, w, h, seeds, idx, weight):
        seeds = np.ascontiguousarray(seeds, dtype=np.float32)
        grad = np.empty((seeds.shape[0], 2), dtype=np.float64)
        self.L.oc_seed_grad_2d(_I(w), _I(h), _I(seeds.shape[0]), _ptr(seeds),
                               _ptr(np.ascontiguousarray(idx, dtype=np.int32)),
                               _ptr(np.ascontiguousarray(weight, dtype=np.float32)), _ptr(grad))
        return grad


class Reference:
    """The reference library compiled from its own sources (oracle/_ref)."""

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.L = ctypes.CDLL(path)

    def set_threads(self, n):
        """threads of the OpenMP loops (the -fopenmp build); returns the count in effect"""
        return int(self.L.ref_set_threads(_I(n)))

    def mask_add_points_2d(self, w, h, n, seed, mask):
        self.L.ref_mask_add_points_2d(_I(w), _I(h), _I(n), _U(seed), _ptr(mask))

    def mask_add_points_3d(self, w, h, d, n, seed, mask):
        self.L.ref_mask_add_points_3d(_I(w), _I(h), _I(d), _I(n), _U(seed), _ptr(mask))

    def mask_add_block_2d(self, w, h, sides, mask):
        self.L.ref_mask_add_block_2d(_I(w), _I(h), _I(sides), _ptr(mask))

    def mask_add_block_3d(self, w, h, d, sides, mask):
        self.L.ref_mask_add_block_3d(_I(w), _I(h), _I(d), _I(sides), _ptr(mask))

    def init_deltas_2d(self, w, h, mask, invert=False):
        out = np.empty((h, w, 2), dtype=np.int16)
        self.L.ref_init_deltas_2d(_I(w), _I(h), _ptr(mask), _ptr(out), _I(int(invert)))
        return out

    def init_deltas_3d(self, w, h, d, mask, invert=False):
        out = np.empty((d, h, w, 3), dtype=np.int16)
        self.L.ref_init_deltas_3d(_I(w), _I(h), _I(d), _ptr(mask), _ptr(out), _I(int(invert)))
        return out

    def init_distances_f32(self, w, h, mask, max_d=np.finfo(np.float32).max):
        out = np.empty((h, w), dtype=np.float32)
        self.L.ref_init_distances_f32(_I(w), _I(h), _ptr(mask), _ptr(out), _F(max_d))
        return out

    def border_2d(self, w, h, mask):
        out = np.empty((h, w, 2), dtype=np.int16)
        self.L.ref_border_2d(_I(w), _I(h), _ptr(mask), _ptr(out))
        return out

    def border_3d(self, w, h, d, mask):
        out = np.empty((d, h, w, 3), dtype=np.int16)
        self.L.ref_border_3d(_I(w), _I(h), _I(d), _ptr(mask), _ptr(out))
        return out

    def jfa_2d(self, w, h, delta, cw=1):
        out = np.ascontiguousarray(delta, dtype=np.int16).copy()
        self.L.ref_jfa_2d(_I(w), _I(h), _ptr(out), _I(cw))
        return out

    def jfa_3d(self, w, h, d, delta, cw=1):
        out = np.ascontiguousarray(delta, dtype=np.int16).copy()
        self.L.ref_jfa_3d(_I(w), _I(h), _I(d), _ptr(out), _I(cw))
        return out

    def fast_sweep_2d(self, w, h, delta, cw=1):
        out = np.ascontiguousarray(delta, dtype=np.int16).copy()
        self.L.ref_fast_sweep_2d(_I(w), _I(h), _ptr(out), _I(cw))
        return out

    def fast_sweep_3d(self, w, h, d, delta, cw=1):
        out = np.ascontiguousarray(delta, dtype=np.int16).copy()
        self.L.ref_fast_sweep_3d(_I(w), _I(h), _I(d), _ptr(out), _I(cw))
        return out

    def init_dirw_2d(self, w, h, mask):
        out = np.empty((4, h, w), dtype=np.float32)
        self.L.ref_init_dirw_2d(_I(w), _I(h), _ptr(mask), _ptr(out))
        return out

    def init_dirw_3d(self, w, h, d, mask):
        out = np.empty((8, d, h, w), dtype=np.float32)
        self.L.ref_init_dirw_3d(_I(w), _I(h), _I(d), _ptr(mask), _ptr(out))
        return out

    def occlusion_sweep_2d(self, w, h, dirw):
        out = np.ascontiguousarray(dirw, dtype=np.float32).copy()
        self.L.ref_occlusion_sweep_2d(_I(w), _I(h), _ptr(out))
        return out

    def occlusion_sweep_3d(self, w, h, d, dirw):
        out = np.ascontiguousarray(dirw, dtype=np.float32).copy()
        self.L.ref_occlusion_sweep_3d(_I(w), _I(h), _I(d), _ptr(out))
        return out

    def occlusion_sweep_3d_from(self, w, h, d, dirw_in, dirw_out):
        src = np.ascontiguousarray(dirw_in, dtype=np.float32)
        out = np.ascontiguousarray(dirw_out, dtype=np.float32).copy()
        self.L.ref_occlusion_sweep_3d_from(_I(w), _I(h), _I(d), _ptr(src), _ptr(out))
        return out

    def encode_ao(self, dirw, dirw_mode=False):
        d = np.ascontiguousarray(dirw, dtype=np.float32)
        planes = d.shape[0]
        n = d.size // planes
        out = np.empty((n, 4) if dirw_mode else n, dtype=np.uint8)
        fn = self.L.ref_encode_dirw if dirw_mode else self.L.ref_encode_ao
        fn(ctypes.c_size_t(n), _I(planes), _ptr(d), _ptr(out))
        return out.reshape(d.shape[1:] + ((4,) if dirw_mode else ()))

    def felzenszwalb_2d(self, w, h, dist):
        out = np.ascontiguousarray(dist, dtype=np.float32).copy()
        self.L.ref_felzenszwalb_2d(_I(w), _I(h), _ptr(out))
        return out

    def danielsson_2d(self, w, h, delta, cw=1):
        out = np.ascontiguousarray(delta, dtype=np.int16).copy()
        self.L.ref_danielsson_2d(_I(w), _I(h), _ptr(out), _I(cw))
        return out

    def danielsson_3d(self, w, h, d, delta, cw=1):
        out = np.ascontiguousarray(delta, dtype=np.int16).copy()
        self.L.ref_danielsson_3d(_I(w), _I(h), _I(d), _ptr(out), _I(cw))
        return out

    def chamfer(self, w, h, dist):
        out = np.ascontiguousarray(dist).copy()
        fn = self.L.ref_chamfer_i16 if out.dtype == np.int16 else self.L.ref_chamfer_i32
        r = fn(_I(w), _I(h), _ptr(out))
        return out, r

    def init_occ(self, w, h, mask, scale=1.0):
        out = np.empty((h, w), dtype=np.float32)
        self.L.ref_init_occ(_I(w), _I(h), _ptr(mask), _ptr(out), _F(scale))
        return out

    def shadow_sweep(self, gx, gy, w, h, occ):
        occ = np.ascontiguousarray(occ, dtype=np.float32)
        out = np.empty_like(occ)
        self.L.ref_shadow_sweep(_F(gx), _F(gy), _I(w), _I(h), _ptr(occ), _ptr(out))
        return out

    def brute_force_2d(self, w, h, delta, cw=1):
        out = np.ascontiguousarray(delta, dtype=np.int16).copy()
        self.L.ref_brute_force_2d(_I(w), _I(h), _ptr(out), _I(cw))
        return out

    def brute_force_3d(self, w, h, d, delta, cw=1):
        out = np.ascontiguousarray(delta, dtype=np.int16).copy()
        self.L.ref_brute_force_3d(_I(w), _I(h), _I(d), _ptr(out), _I(cw))
        return out


def load_oracle():
    return Oracle()


def host_threads():
    """threads for the oracle's parallel loops: the host's cores, at most 16"""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return max(1, min(16, n))


def load_reference(omp=False):
    """The compiled reference (its -fopenmp build with omp=True), or None where
    it was not built."""
    try:
        return Reference(REF_OMP_SO if omp else REF_SO)
    except (FileNotFoundError, OSError):
        return None


class MeshReference:
    """The reference's MeshSupport.cpp compiled from its sources (oracle/_ref)."""

    def __init__(self, path=MESH_REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.L = ctypes.CDLL(path)

    def read_obj(self, path):
        npos, nidx = ctypes.c_int(), ctypes.c_int()
        if not self.L.ref_read_obj(path.encode(), ctypes.byref(npos), ctypes.byref(nidx), None, None):
            raise IOError(path)
        pos = np.empty((npos.value, 3), np.float32)
        idx = np.empty(nidx.value, np.int32)
        self.L.ref_read_obj(path.encode(), ctypes.byref(npos), ctypes.byref(nidx), _ptr(pos), _ptr(idx))
        return pos, idx

    def find_bounds(self, positions):
        positions = np.ascontiguousarray(positions, np.float32)
        out = np.empty(6, np.float32)
        self.L.ref_find_bounds(_I(positions.size // 3), _ptr(positions), _ptr(out))
        return out

    def mask_from_triangles(self, w, h, d, vertices, indices, bbox):
        mask = empty_mask(w, h, d)
        idx = np.ascontiguousarray(indices, np.int32)
        self.L.ref_mask_from_triangles(_I(idx.size // 3), _ptr(idx), _ptr(np.ascontiguousarray(vertices, np.float32)),
                                       _ptr(np.ascontiguousarray(bbox, np.float32)), _I(w), _I(h), _I(d), _ptr(mask))
        return mask


def load_mesh_reference():
    """The compiled mesh support, or None where it was not built."""
    try:
        return MeshReference()
    except (FileNotFoundError, OSError):
        return None
