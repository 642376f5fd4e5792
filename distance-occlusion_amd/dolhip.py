"""Python face of libdolhip (include/dol_hip.h) for tests, benchmarks and
torch users.

Host-array methods take and return numpy arrays with the reference's layouts
(DistanceOcclusionLib.h:16-57): bitmasks as flat uint32 words, 2D deltas as
int16 (h, w, 2), 3D deltas as int16 (d, h, w, 3).  `*_dev` methods take
torch tensors (or raw device addresses) already on the GPU and a stream
(default: torch's current stream) and return without synchronising.

There is no CPU fallback: loading fails if libdolhip.so was not built, and
every compute call raises DolHipError when no gfx950 device is usable.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdolhip.so")

OK = 0
ERR_INVALID_ARG = 1
ERR_OUT_OF_MEMORY = 2
ERR_DEVICE = 3
ERR_NO_DEVICE = 4

MAX_DELTA2 = 32767
MAX_DELTA3 = 25000
EDT_INF = 0x7FFFFFFF

_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_SZ = ctypes.c_size_t


class DolHipError(RuntimeError):
    def __init__(self, code, name, msg):
        super().__init__(f"{name} failed ({code}): {msg}")
        self.code = code


# name -> (argtypes, restype); _P for every pointer and the stream
_SIGS = {
    "dolhip_last_error": ([], ctypes.c_char_p),
    "dolhip_version": ([], ctypes.c_char_p),
    "dolhip_device_count": ([], _I),
    "dolhip_profile_enable": ([_I], _I),
    "dolhip_profile_query": ([ctypes.c_char_p, _P, _P], _I),
    "dolhip_mask_from_bytes": ([_I, _I, _P, _I, _I, _P], _I),
    "dolhip_mask_from_bytes_dev": ([_I, _I, _P, _I, _I, _P, _P], _I),
    "dolhip_init_deltas_2d": ([_I, _I, _P, _P, _I], _I),
    "dolhip_init_deltas_2d_dev": ([_I, _I, _P, _P, _I, _P], _I),
    "dolhip_init_deltas_3d": ([_I, _I, _I, _P, _P, _I], _I),
    "dolhip_init_deltas_3d_dev": ([_I, _I, _I, _P, _P, _I, _P], _I),
    "dolhip_init_border_deltas_2d": ([_I, _I, _P, _P], _I),
    "dolhip_init_border_deltas_2d_dev": ([_I, _I, _P, _P, _P], _I),
    "dolhip_init_border_deltas_3d": ([_I, _I, _I, _P, _P], _I),
    "dolhip_init_border_deltas_3d_dev": ([_I, _I, _I, _P, _P, _P], _I),
    "dolhip_init_distances_f32": ([_I, _I, _P, _P, _F], _I),
    "dolhip_init_distances_f32_dev": ([_I, _I, _P, _P, _F, _P], _I),
    "dolhip_init_distances_i32": ([_I, _I, _P, _P, ctypes.c_int32], _I),
    "dolhip_init_distances_i16": ([_I, _I, _P, _P, ctypes.c_int16], _I),
    "dolhip_jump_flood_2d": ([_I, _I, _P, _I], _I),
    "dolhip_jump_flood_2d_workspace": ([_I, _I], _SZ),
    "dolhip_jump_flood_2d_dev": ([_I, _I, _P, _I, _P, _P], _I),
    "dolhip_jump_flood_2d_noalloc": ([_I, _I, _P, _I, _P], _I),
    "dolhip_jump_flood_3d": ([_I, _I, _I, _P, _I], _I),
    "dolhip_jump_flood_3d_workspace": ([_I, _I, _I], _SZ),
    "dolhip_jump_flood_3d_dev": ([_I, _I, _I, _P, _I, _P, _P], _I),
    "dolhip_fast_sweep_2d": ([_I, _I, _P, _I], _I),
    "dolhip_fast_sweep_2d_dev": ([_I, _I, _P, _I, _P], _I),
    "dolhip_fast_sweep_3d": ([_I, _I, _I, _P, _I], _I),
    "dolhip_fast_sweep_3d_dev": ([_I, _I, _I, _P, _I, _P], _I),
    "dolhip_brute_force_2d": ([_I, _I, _P, _I], _I),
    "dolhip_brute_force_2d_dev": ([_I, _I, _P, _I, _P], _I),
    "dolhip_brute_force_3d": ([_I, _I, _I, _P, _I], _I),
    "dolhip_brute_force_3d_dev": ([_I, _I, _I, _P, _I, _P], _I),
    "dolhip_encode_border_image_2d": ([_I, _I, _P, _P, _F, _I, _P], _I),
    "dolhip_encode_border_image_2d_dev": ([_I, _I, _P, _P, _F, _I, _P, _P], _I),
    "dolhip_encode_border_image_3d": ([_I, _I, _I, _P, _P, _F, _I, _P], _I),
    "dolhip_encode_border_image_3d_dev": ([_I, _I, _I, _P, _P, _F, _I, _P, _P], _I),
    "dolhip_encode_distance_u8_2d": ([_I, _I, _P, _F, _P], _I),
    "dolhip_encode_distance_u8_2d_dev": ([_I, _I, _P, _F, _P, _P], _I),
    "dolhip_encode_distance_u8_3d": ([_I, _I, _I, _P, _F, _P], _I),
    "dolhip_encode_distance_u8_3d_dev": ([_I, _I, _I, _P, _F, _P, _P], _I),
    "dolhip_encode_scalar_u8_f32": ([_I, _I, _P, _F, _P], _I),
    "dolhip_encode_scalar_u8_f32_dev": ([_I, _I, _P, _F, _P, _P], _I),
    "dolhip_encode_scalar_u8_i32": ([_I, _I, _P, _F, _P], _I),
    "dolhip_encode_scalar_u8_i32_dev": ([_I, _I, _P, _F, _P, _P], _I),
    "dolhip_encode_scalar_u8_i16": ([_I, _I, _P, _F, _P], _I),
    "dolhip_encode_scalar_u8_i16_dev": ([_I, _I, _P, _F, _P, _P], _I),
    "dolhip_sqrt_f32": ([_I, _I, _P, _P], _I),
    "dolhip_sqrt_f32_dev": ([_I, _I, _P, _P, _P], _I),
    "dolhip_danielsson_2d": ([_I, _I, _P, _I], _I),
    "dolhip_danielsson_2d_dev": ([_I, _I, _P, _I, _P], _I),
    "dolhip_danielsson_3d": ([_I, _I, _I, _P, _I], _I),
    "dolhip_danielsson_3d_dev": ([_I, _I, _I, _P, _I, _P], _I),
    "dolhip_chamfer_i16": ([_I, _I, _P, _P], _I),
    "dolhip_chamfer_i16_dev": ([_I, _I, _P, _P, _P], _I),
    "dolhip_chamfer_i32": ([_I, _I, _P, _P], _I),
    "dolhip_chamfer_i32_dev": ([_I, _I, _P, _P, _P], _I),
    "dolhip_init_occ": ([_I, _I, _P, _P, _F], _I),
    "dolhip_init_occ_dev": ([_I, _I, _P, _P, _F, _P], _I),
    "dolhip_shadow_sweep": ([_F, _F, _I, _I, _P, _P], _I),
    "dolhip_shadow_sweep_dev": ([_F, _F, _I, _I, _P, _P, _P], _I),
    "dolhip_init_dirw_2d": ([_I, _I, _P, _P], _I),
    "dolhip_init_dirw_2d_dev": ([_I, _I, _P, _P, _P], _I),
    "dolhip_init_dirw_3d": ([_I, _I, _I, _P, _P], _I),
    "dolhip_init_dirw_3d_dev": ([_I, _I, _I, _P, _P, _P], _I),
    "dolhip_occlusion_sweep_2d": ([_I, _I, _P], _I),
    "dolhip_occlusion_sweep_2d_dev": ([_I, _I, _P, _P], _I),
    "dolhip_occlusion_sweep_3d": ([_I, _I, _I, _P], _I),
    "dolhip_occlusion_sweep_3d_dev": ([_I, _I, _I, _P, _P], _I),
    "dolhip_occlusion_sweep_3d_from": ([_I, _I, _I, _P, _P], _I),
    "dolhip_occlusion_sweep_3d_from_dev": ([_I, _I, _I, _P, _P, _P], _I),
    "dolhip_encode_ao_2d": ([_I, _I, _P, _P], _I),
    "dolhip_encode_ao_2d_dev": ([_I, _I, _P, _P, _P], _I),
    "dolhip_encode_ao_3d": ([_I, _I, _I, _P, _P], _I),
    "dolhip_encode_ao_3d_dev": ([_I, _I, _I, _P, _P, _P], _I),
    "dolhip_encode_dirw_2d": ([_I, _I, _P, _P], _I),
    "dolhip_encode_dirw_2d_dev": ([_I, _I, _P, _P, _P], _I),
    "dolhip_encode_dirw_3d": ([_I, _I, _I, _P, _P], _I),
    "dolhip_encode_dirw_3d_dev": ([_I, _I, _I, _P, _P, _P], _I),
    "dolhip_felzenszwalb_2d": ([_I, _I, _P], _I),
    "dolhip_felzenszwalb_2d_workspace": ([_I, _I], _SZ),
    "dolhip_felzenszwalb_2d_dev": ([_I, _I, _P, _P, _P], _I),
    "dolhip_edt_2d": ([_I, _I, _P, _I, _P, _P], _I),
    "dolhip_edt_2d_workspace": ([_I, _I], _SZ),
    "dolhip_edt_2d_dev": ([_I, _I, _P, _I, _P, _P, _P, _P], _I),
    "dolhip_edt_3d": ([_I, _I, _I, _P, _I, _P, _P], _I),
    "dolhip_edt_3d_workspace": ([_I, _I, _I], _SZ),
    "dolhip_edt_3d_dev": ([_I, _I, _I, _P, _I, _P, _P, _P, _P], _I),
    "dolhip_signed_edt_2d": ([_I, _I, _P, _P], _I),
    "dolhip_signed_edt_2d_dev": ([_I, _I, _P, _P, _P, _P], _I),
    "dolhip_signed_edt_3d": ([_I, _I, _I, _P, _P], _I),
    "dolhip_signed_edt_3d_dev": ([_I, _I, _I, _P, _P, _P, _P], _I),
    "dolhip_dist_2d_dev": ([_I, _P, _P, _P], _I),
    "dolhip_signed_distance_2d": ([_I, _I, _P, _P, _P], _I),
    "dolhip_signed_distance_2d_dev": ([_I, _I, _P, _P, _P, _P], _I),
    "dolhip_signed_distance_3d": ([_I, _I, _I, _P, _P, _P], _I),
    "dolhip_signed_distance_3d_dev": ([_I, _I, _I, _P, _P, _P, _P], _I),
    "dolhip_encode_sdf_u8_2d": ([_I, _I, _P, _P, _F, _P], _I),
    "dolhip_encode_sdf_u8_2d_dev": ([_I, _I, _P, _P, _F, _P, _P], _I),
    "dolhip_encode_sdf_u8_3d": ([_I, _I, _I, _P, _P, _F, _P], _I),
    "dolhip_encode_sdf_u8_3d_dev": ([_I, _I, _I, _P, _P, _F, _P, _P], _I),
    "dolhip_encode_sdf_delta_rgba_2d": ([_I, _I, _P, _F, _P], _I),
    "dolhip_encode_sdf_rg_rgba_2d": ([_I, _I, _P, _P, _F, _P], _I),
    "dolhip_signed_border_2d_dev": ([_I, _I, _P, _P, _P, _P], _I),
    "dolhip_signed_border_3d_dev": ([_I, _I, _I, _P, _P, _P, _P], _I),
    "dolhip_mask_from_triangles": ([_I, _P, _P, _P, _I, _I, _I, _P], _I),
    "dolhip_mask_from_triangles_dev": ([_I, _P, _P, _P, _I, _I, _I, _P, _P], _I),
    "dolhip_two_field_2d_workspace": ([_I, _I], _SZ),
    "dolhip_two_field_3d_workspace": ([_I, _I, _I], _SZ),
    "dolhip_two_field_deltas_2d": ([_I, _I, _P, _P, _P, _I], _I),
    "dolhip_two_field_deltas_2d_dev": ([_I, _I, _P, _P, _P, _I, _P, _P], _I),
    "dolhip_two_field_deltas_3d": ([_I, _I, _I, _P, _P, _P, _I], _I),
    "dolhip_two_field_deltas_3d_dev": ([_I, _I, _I, _P, _P, _P, _I, _P, _P], _I),
    "dolhip_encode_two_field_2d": ([_I, _I, _P, _P, _F, _I, _P], _I),
    "dolhip_encode_two_field_2d_dev": ([_I, _I, _P, _P, _F, _I, _P, _P], _I),
    "dolhip_encode_two_field_3d": ([_I, _I, _I, _P, _P, _F, _I, _P], _I),
    "dolhip_encode_two_field_3d_dev": ([_I, _I, _I, _P, _P, _F, _I, _P, _P], _I),
    "dolhip_batch_create": ([_I, _P, _P, _P, _I, _I, _P], _I),
    "dolhip_batch_destroy": ([_P], _I),
    "dolhip_batch_bytes": ([_P], ctypes.c_int64),
    "dolhip_batch_sdf_u8_dev": ([_P, _P, _P, _I, _P], _I),
    "dolhip_glyph_sdf_u8_dev": ([_I, _I, _P, _I, _I, _P, _P, _P], _I),
    "dolhip_batch_per_image_dev": ([_P, _P, _P, _P], _I),
    "dolhip_batch_config": ([_P, _P, _P], _I),
    "dolhip_batch_set_variant": ([_P, _I], _I),
    "dolhip_batch_tune": ([_P, _P, _P, _P], _I),
    "dolhip_mask_add_points_2d": ([_I, _I, _I, ctypes.c_uint32, _P], None),
    "dolhip_mask_add_points_3d": ([_I, _I, _I, _I, ctypes.c_uint32, _P], None),
    "dolhip_mask_add_block_2d": ([_I, _I, _I, _P], None),
    "dolhip_mask_add_block_3d": ([_I, _I, _I, _I, _P], None),
    "dolhip_scene_shapes_2d": ([_I, _I, _I, ctypes.c_uint64, _P], None),
    "dolhip_scene_shapes_3d": ([_I, _I, _I, _I, _I, ctypes.c_uint64, _P], None),
    "dolhip_scene_glyphs": ([_I, _I, _I, ctypes.c_uint64, _P, _P, _P, _P], ctypes.c_int64),
    "dolhip_scene_seeds": ([_I, _I, _I, ctypes.c_uint64, _P, _P], None),
    "dolhip_seed_workspace": ([_I, _I, _I], _SZ),
    "dolhip_seed_nearest_2d_dev": ([_I, _I, _I, _P, _P, _P, _P, _P], _I),
    "dolhip_seed_grad_2d_dev": ([_I, _I, _I, _P, _P, _P, _P, _P, _P], _I),
}


def _addr(x):
    """numpy array / torch tensor / int / None -> c_void_p."""
    if x is None:
        return None
    if isinstance(x, int):
        return ctypes.c_void_p(x)
    if isinstance(x, np.ndarray):
        return ctypes.c_void_p(x.ctypes.data)
    if hasattr(x, "data_ptr"):
        return ctypes.c_void_p(x.data_ptr())
    raise TypeError(f"cannot take the address of {type(x)}")


def _stream(stream):
    if stream is None:
        try:
            import torch

            if torch.cuda.is_available():
                return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        except ImportError:
            pass
        return None
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(stream.cuda_stream)


def mask_stride(w):
    return (w + 31) // 32


class DolHip:
    def __init__(self, path=LIB_PATH):
        if not os.path.exists(path):
            raise ImportError(
                f"{path} is not built: run python distance-occlusion_amd/build.py "
                "(there is no CPU fallback)")
        self.path = path
        # PyTorch ships its own HIP runtime under the same soname as ROCm's
        # (libamdhip64.so.7); whichever is loaded first serves the whole
        # process.  Load PyTorch's first so both share the runtime PyTorch
        # initialises the device with.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        self.L = ctypes.CDLL(path)
        self.missing = []
        for name, (args, res) in _SIGS.items():
            try:
                fn = getattr(self.L, name)
            except AttributeError:
                self.missing.append(name)
                continue
            fn.argtypes = args
            fn.restype = res

    # ---- plumbing
    def last_error(self):
        return self.L.dolhip_last_error().decode()

    def _call(self, name, *args):
        rc = getattr(self.L, name)(*args)
        if rc != OK:
            raise DolHipError(rc, name, self.last_error())
        return rc

    def device_count(self):
        return self.L.dolhip_device_count()

    def profile_enable(self, on=True):
        self._call("dolhip_profile_enable", int(on))

    def profile_query(self, prefix):
        avg = ctypes.c_double(0.0)
        cnt = ctypes.c_int(0)
        self._call("dolhip_profile_query", prefix.encode(), ctypes.byref(avg), ctypes.byref(cnt))
        return avg.value, cnt.value

    # ---- host-array entry points
    def mask_from_bytes(self, w, rows, data, threshold=128, less_than=True):
        data = np.ascontiguousarray(data, dtype=np.uint8)
        assert data.size == w * rows
        mask = np.zeros(mask_stride(w) * rows, dtype=np.uint32)
        self._call("dolhip_mask_from_bytes", w, rows, _addr(data), threshold, int(less_than), _addr(mask))
        return mask

    def init_deltas_2d(self, w, h, mask, invert=False):
        out = np.empty((h, w, 2), dtype=np.int16)
        self._call("dolhip_init_deltas_2d", w, h, _addr(np.ascontiguousarray(mask, np.uint32)), _addr(out), int(invert))
        return out

    def init_deltas_3d(self, w, h, d, mask, invert=False):
        out = np.empty((d, h, w, 3), dtype=np.int16)
        self._call("dolhip_init_deltas_3d", w, h, d, _addr(np.ascontiguousarray(mask, np.uint32)), _addr(out), int(invert))
        return out

    def init_border_deltas_2d(self, w, h, mask):
        out = np.empty((h, w, 2), dtype=np.int16)
        self._call("dolhip_init_border_deltas_2d", w, h, _addr(np.ascontiguousarray(mask, np.uint32)), _addr(out))
        return out

    def init_border_deltas_3d(self, w, h, d, mask):
        out = np.empty((d, h, w, 3), dtype=np.int16)
        self._call("dolhip_init_border_deltas_3d", w, h, d, _addr(np.ascontiguousarray(mask, np.uint32)), _addr(out))
        return out

    def init_distances_f32(self, w, h, mask, max_d=float(np.finfo(np.float32).max)):
        out = np.empty((h, w), dtype=np.float32)
        self._call("dolhip_init_distances_f32", w, h, _addr(np.ascontiguousarray(mask, np.uint32)), _addr(out), max_d)
        return out

    def init_distances_i32(self, w, h, mask, max_d=2**31 - 1):
        out = np.empty((h, w), dtype=np.int32)
        self._call("dolhip_init_distances_i32", w, h, _addr(np.ascontiguousarray(mask, np.uint32)), _addr(out), max_d)
        return out

    def init_distances_i16(self, w, h, mask, max_d=32767):
        out = np.empty((h, w), dtype=np.int16)
        self._call("dolhip_init_distances_i16", w, h, _addr(np.ascontiguousarray(mask, np.uint32)), _addr(out), max_d)
        return out

    def jump_flood_2d(self, w, h, delta, cw=1):
        out = np.ascontiguousarray(delta, dtype=np.int16).copy()
        self._call("dolhip_jump_flood_2d", w, h, _addr(out), cw)
        return out

    def jump_flood_3d(self, w, h, d, delta, cw=1):
        out = np.ascontiguousarray(delta, dtype=np.int16).copy()
        self._call("dolhip_jump_flood_3d", w, h, d, _addr(out), cw)
        return out

    def fast_sweep_2d(self, w, h, delta, cw=1):
        out = np.ascontiguousarray(delta, dtype=np.int16).copy()
        self._call("dolhip_fast_sweep_2d", w, h, _addr(out), cw)
        return out

    def fast_sweep_3d(self, w, h, d, delta, cw=1):
        out = np.ascontiguousarray(delta, dtype=np.int16).copy()
        self._call("dolhip_fast_sweep_3d", w, h, d, _addr(out), cw)
        return out

    def brute_force_2d(self, w, h, delta, cw=1):
        out = np.ascontiguousarray(delta, dtype=np.int16).copy()
        self._call("dolhip_brute_force_2d", w, h, _addr(out), cw)
        return out

    def brute_force_3d(self, w, h, d, delta, cw=1):
        out = np.ascontiguousarray(delta, dtype=np.int16).copy()
        self._call("dolhip_brute_force_3d", w, h, d, _addr(out), cw)
        return out

    def border_image(self, w, h, d, mask, delta, max_len, image_type):
        """the border method's image: type 0 sdf (bytes), 1 sdf-delta, 2 sdf-rg,
        3 sdf-angles (RGBA); d = 0 for a 2D field"""
        mask = np.ascontiguousarray(mask, np.uint32)
        delta = np.ascontiguousarray(delta, np.int16)
        dd = max(d, 1)
        out = np.empty((dd, h, w) if image_type == 0 else (dd, h, w, 4), dtype=np.uint8)
        if d == 0:
            self._call("dolhip_encode_border_image_2d", w, h, _addr(mask), _addr(delta), max_len, image_type, _addr(out))
            return out[0]
        self._call("dolhip_encode_border_image_3d", w, h, d, _addr(mask), _addr(delta), max_len, image_type, _addr(out))
        return out

    def distance_u8(self, w, h, d, delta, max_len):
        delta = np.ascontiguousarray(delta, np.int16)
        out = np.empty((max(d, 1), h, w), dtype=np.uint8)
        if d == 0:
            self._call("dolhip_encode_distance_u8_2d", w, h, _addr(delta), max_len, _addr(out))
            return out[0]
        self._call("dolhip_encode_distance_u8_3d", w, h, d, _addr(delta), max_len, _addr(out))
        return out

    def scalar_u8(self, w, h, values, max_len):
        v = np.ascontiguousarray(values)
        name = {np.dtype(np.float32): "f32", np.dtype(np.int32): "i32", np.dtype(np.int16): "i16"}[v.dtype]
        out = np.empty((h, w), dtype=np.uint8)
        self._call(f"dolhip_encode_scalar_u8_{name}", w, h, _addr(v), max_len, _addr(out))
        return out

    def sqrt_f32(self, w, h, values):
        v = np.ascontiguousarray(values, np.float32)
        out = np.empty((h, w), dtype=np.float32)
        self._call("dolhip_sqrt_f32", w, h, _addr(v), _addr(out))
        return out

    def danielsson_2d(self, w, h, delta, cw=1):
        out = np.ascontiguousarray(delta, dtype=np.int16).copy()
        self._call("dolhip_danielsson_2d", w, h, _addr(out), cw)
        return out

    def danielsson_3d(self, w, h, d, delta, cw=1):
        out = np.ascontiguousarray(delta, dtype=np.int16).copy()
        self._call("dolhip_danielsson_3d", w, h, d, _addr(out), cw)
        return out

    def chamfer(self, w, h, dist):
        """Chamfer in place on int16 or int32 distances; returns (out, multiplier)."""
        dist = np.asarray(dist)
        if dist.dtype not in (np.int16, np.int32):
            raise TypeError("Chamfer takes int16 or int32 distances")
        out = np.ascontiguousarray(dist).copy()
        mult = ctypes.c_int(0)
        name = "dolhip_chamfer_i16" if out.dtype == np.int16 else "dolhip_chamfer_i32"
        self._call(name, w, h, _addr(out), ctypes.addressof(mult))
        return out, mult.value

    def init_occ(self, w, h, mask, scale=1.0):
        out = np.empty((h, w), dtype=np.float32)
        self._call("dolhip_init_occ", w, h, _addr(np.ascontiguousarray(mask, np.uint32)), _addr(out), scale)
        return out

    def shadow_sweep(self, gx, gy, w, h, occ):
        occ = np.ascontiguousarray(occ, dtype=np.float32)
        out = np.empty((h, w), dtype=np.float32)
        self._call("dolhip_shadow_sweep", gx, gy, w, h, _addr(occ), _addr(out))
        return out

    def init_dirw_2d(self, w, h, mask):
        out = np.empty((4, h, w), dtype=np.float32)
        self._call("dolhip_init_dirw_2d", w, h, _addr(np.ascontiguousarray(mask, np.uint32)), _addr(out))
        return out

    def init_dirw_3d(self, w, h, d, mask):
        out = np.empty((8, d, h, w), dtype=np.float32)
        self._call("dolhip_init_dirw_3d", w, h, d, _addr(np.ascontiguousarray(mask, np.uint32)), _addr(out))
        return out

    def occlusion_sweep_2d(self, w, h, dirw):
        out = np.ascontiguousarray(dirw, dtype=np.float32).copy()
        self._call("dolhip_occlusion_sweep_2d", w, h, _addr(out))
        return out

    def occlusion_sweep_3d(self, w, h, d, dirw):
        out = np.ascontiguousarray(dirw, dtype=np.float32).copy()
        self._call("dolhip_occlusion_sweep_3d", w, h, d, _addr(out))
        return out

    def occlusion_sweep_3d_from(self, w, h, d, dirw_in, dirw_out):
        src = np.ascontiguousarray(dirw_in, dtype=np.float32)
        out = np.ascontiguousarray(dirw_out, dtype=np.float32).copy()
        self._call("dolhip_occlusion_sweep_3d_from", w, h, d, _addr(src), _addr(out))
        return out

    def encode_ao(self, dirw, rgba=False):
        """AO image of a (4, h, w) or (8, d, h, w) field: bytes, or RGBA (EncodeDirW)."""
        src = np.ascontiguousarray(dirw, dtype=np.float32)
        dims = tuple(reversed(src.shape[1:]))
        out = np.empty(src.shape[1:] + ((4,) if rgba else ()), dtype=np.uint8)
        kind = "dirw" if rgba else "ao"
        self._call(f"dolhip_encode_{kind}_{len(dims)}d", *dims, _addr(src), _addr(out))
        return out

    def felzenszwalb_2d(self, w, h, dist):
        out = np.ascontiguousarray(dist, dtype=np.float32).copy()
        self._call("dolhip_felzenszwalb_2d", w, h, _addr(out))
        return out

    def edt_2d(self, w, h, mask, invert=False, want_d2=True, want_dist=True):
        d2 = np.empty((h, w), dtype=np.int32) if want_d2 else None
        fl = np.empty((h, w), dtype=np.float32) if want_dist else None
        self._call("dolhip_edt_2d", w, h, _addr(np.ascontiguousarray(mask, np.uint32)), int(invert), _addr(d2), _addr(fl))
        return d2, fl

    def edt_3d(self, w, h, d, mask, invert=False, want_d2=True, want_dist=True):
        d2 = np.empty((d, h, w), dtype=np.int32) if want_d2 else None
        fl = np.empty((d, h, w), dtype=np.float32) if want_dist else None
        self._call("dolhip_edt_3d", w, h, d, _addr(np.ascontiguousarray(mask, np.uint32)), int(invert), _addr(d2), _addr(fl))
        return d2, fl

    def signed_edt_2d(self, w, h, mask):
        out = np.empty((h, w), dtype=np.float32)
        self._call("dolhip_signed_edt_2d", w, h, _addr(np.ascontiguousarray(mask, np.uint32)), _addr(out))
        return out

    def signed_edt_3d(self, w, h, d, mask):
        out = np.empty((d, h, w), dtype=np.float32)
        self._call("dolhip_signed_edt_3d", w, h, d, _addr(np.ascontiguousarray(mask, np.uint32)), _addr(out))
        return out

    def signed_distance_2d(self, w, h, mask, delta):
        out = np.empty((h, w), dtype=np.float32)
        self._call("dolhip_signed_distance_2d", w, h, _addr(np.ascontiguousarray(mask, np.uint32)),
                   _addr(np.ascontiguousarray(delta, np.int16)), _addr(out))
        return out

    def signed_distance_3d(self, w, h, d, mask, delta):
        out = np.empty((d, h, w), dtype=np.float32)
        self._call("dolhip_signed_distance_3d", w, h, d, _addr(np.ascontiguousarray(mask, np.uint32)),
                   _addr(np.ascontiguousarray(delta, np.int16)), _addr(out))
        return out

    def encode_sdf_u8_2d(self, w, h, mask, delta, max_len):
        out = np.empty((h, w), dtype=np.uint8)
        self._call("dolhip_encode_sdf_u8_2d", w, h, _addr(np.ascontiguousarray(mask, np.uint32)),
                   _addr(np.ascontiguousarray(delta, np.int16)), max_len, _addr(out))
        return out

    def encode_sdf_u8_3d(self, w, h, d, mask, delta, max_len):
        out = np.empty((d, h, w), dtype=np.uint8)
        self._call("dolhip_encode_sdf_u8_3d", w, h, d, _addr(np.ascontiguousarray(mask, np.uint32)),
                   _addr(np.ascontiguousarray(delta, np.int16)), max_len, _addr(out))
        return out

    def encode_sdf_delta_rgba_2d(self, w, h, delta, max_len):
        out = np.empty((h, w, 4), dtype=np.uint8)
        self._call("dolhip_encode_sdf_delta_rgba_2d", w, h, _addr(np.ascontiguousarray(delta, np.int16)), max_len, _addr(out))
        return out

    def encode_sdf_rg_rgba_2d(self, w, h, mask, delta, max_len):
        out = np.empty((h, w, 4), dtype=np.uint8)
        self._call("dolhip_encode_sdf_rg_rgba_2d", w, h, _addr(np.ascontiguousarray(mask, np.uint32)),
                   _addr(np.ascontiguousarray(delta, np.int16)), max_len, _addr(out))
        return out

    def two_field_deltas_2d(self, w, h, mask, cw=1):
        """exterior and interior JumpFlood fields of the tool's -S mode"""
        d0 = np.empty((h, w, 2), dtype=np.int16)
        d1 = np.empty((h, w, 2), dtype=np.int16)
        self._call("dolhip_two_field_deltas_2d", w, h, _addr(np.ascontiguousarray(mask, np.uint32)),
                   _addr(d0), _addr(d1), cw)
        return d0, d1

    def two_field_deltas_3d(self, w, h, d, mask, cw=1):
        d0 = np.empty((d, h, w, 3), dtype=np.int16)
        d1 = np.empty((d, h, w, 3), dtype=np.int16)
        self._call("dolhip_two_field_deltas_3d", w, h, d, _addr(np.ascontiguousarray(mask, np.uint32)),
                   _addr(d0), _addr(d1), cw)
        return d0, d1

    def encode_two_field_2d(self, w, h, d0, d1, max_len, image_type):
        out = np.empty((h, w) if image_type == 0 else (h, w, 4), dtype=np.uint8)
        self._call("dolhip_encode_two_field_2d", w, h, _addr(np.ascontiguousarray(d0, np.int16)),
                   _addr(np.ascontiguousarray(d1, np.int16)), max_len, image_type, _addr(out))
        return out

    def encode_two_field_3d(self, w, h, d, d0, d1, max_len, image_type):
        out = np.empty((d, h, w) if image_type == 0 else (d, h, w, 4), dtype=np.uint8)
        self._call("dolhip_encode_two_field_3d", w, h, d, _addr(np.ascontiguousarray(d0, np.int16)),
                   _addr(np.ascontiguousarray(d1, np.int16)), max_len, image_type, _addr(out))
        return out

    def mask_from_triangles(self, w, h, d, vertices, indices, bbox, mask=None):
        """ORs the surface voxels of the triangles into mask (a new zero one if None)"""
        if mask is None:
            mask = np.zeros(mask_stride(w) * h * d, np.uint32)
        vertices = np.ascontiguousarray(vertices, np.float32)
        idx = None if indices is None else np.ascontiguousarray(indices, np.int32)
        n = (vertices.size // 9) if idx is None else idx.size // 3
        self._call("dolhip_mask_from_triangles", n, _addr(idx), _addr(vertices),
                   _addr(np.ascontiguousarray(bbox, np.float32)), w, h, d, _addr(mask))
        return mask

    # ---- host generators
    def mask_add_points_2d(self, w, h, n, seed, mask):
        self.L.dolhip_mask_add_points_2d(w, h, n, seed, _addr(mask))

    def mask_add_points_3d(self, w, h, d, n, seed, mask):
        self.L.dolhip_mask_add_points_3d(w, h, d, n, seed, _addr(mask))

    def mask_add_block_2d(self, w, h, sides, mask):
        self.L.dolhip_mask_add_block_2d(w, h, sides, _addr(mask))

    def mask_add_block_3d(self, w, h, d, sides, mask):
        self.L.dolhip_mask_add_block_3d(w, h, d, sides, _addr(mask))

    def scene_shapes_2d(self, w, h, n=2000, seed=2):
        mask = np.zeros(mask_stride(w) * h, dtype=np.uint32)
        self.L.dolhip_scene_shapes_2d(w, h, n, seed, _addr(mask))
        return mask

    def scene_shapes_3d(self, w, h, d, n_ellipsoids=300, n_boxes=100, seed=3):
        mask = np.zeros(mask_stride(w) * h * d, dtype=np.uint32)
        self.L.dolhip_scene_shapes_3d(w, h, d, n_ellipsoids, n_boxes, seed, _addr(mask))
        return mask

    def scene_glyphs(self, n=20000, min_size=16, max_size=256, seed=4):
        """(widths, heights, offsets, bytes) of a ragged glyph set."""
        w = np.empty(n, dtype=np.int32)
        h = np.empty(n, dtype=np.int32)
        off = np.empty(n, dtype=np.int64)
        total = self.L.dolhip_scene_glyphs(n, min_size, max_size, seed, _addr(w), _addr(h), _addr(off), None)
        data = np.empty(total, dtype=np.uint8)
        self.L.dolhip_scene_glyphs(n, min_size, max_size, seed, _addr(w), _addr(h), _addr(off), _addr(data))
        return w, h, off, data

    def scene_seeds(self, w, h, n=4096, seed=5, weights=True):
        seeds = np.empty((n, 2), dtype=np.float32)
        wt = np.empty((h, w), dtype=np.float32) if weights else None
        self.L.dolhip_scene_seeds(w, h, n, seed, _addr(seeds), _addr(wt))
        return seeds, wt

    # ---- device entry points (torch tensors or raw addresses)
    def workspace(self, name, *dims):
        return getattr(self.L, f"dolhip_{name}_workspace")(*dims)

    def call_dev(self, name, *args, stream=None):
        """dolhip_<name>_dev with tensors/addresses converted; the stream is appended."""
        conv = []
        for a in args:
            if a is None or isinstance(a, (np.ndarray,)) or hasattr(a, "data_ptr"):
                conv.append(_addr(a))
            else:
                conv.append(a)
        return self._call(f"dolhip_{name}_dev", *conv, _stream(stream))


class Batch:
    """A ragged set of glyph masks (dolhip_batch_*); arrays are numpy."""

    def __init__(self, lib, widths, heights, offsets, threshold=128, less_than=False):
        self.lib = lib
        self.w = np.ascontiguousarray(widths, dtype=np.int32)
        self.h = np.ascontiguousarray(heights, dtype=np.int32)
        self.off = np.ascontiguousarray(offsets, dtype=np.int64)
        self.ptr = ctypes.c_void_p()
        lib._call("dolhip_batch_create", len(self.w), _addr(self.w), _addr(self.h), _addr(self.off),
                  threshold, int(less_than), ctypes.byref(self.ptr))

    def close(self):
        if self.ptr:
            self.lib.L.dolhip_batch_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nbytes(self):
        return self.lib.L.dolhip_batch_bytes(self.ptr)

    def variants(self):
        return self.lib.L.dolhip_batch_set_variant(self.ptr, -1)

    def set_variant(self, v):
        self.lib._call("dolhip_batch_set_variant", self.ptr, v)

    def config(self):
        cells = ctypes.c_int(0)
        waves = ctypes.c_int(0)
        self.lib._call("dolhip_batch_config", self.ptr, ctypes.byref(cells), ctypes.byref(waves))
        return cells.value, waves.value

    def run(self, data, out, graph=False, stream=None):
        self.lib._call("dolhip_batch_sdf_u8_dev", self.ptr, _addr(data), _addr(out), int(graph), _stream(stream))

    def run_per_image(self, data, out, stream=None):
        self.lib._call("dolhip_batch_per_image_dev", self.ptr, _addr(data), _addr(out), _stream(stream))

    def tune(self, data, out, stream=None):
        self.lib._call("dolhip_batch_tune", self.ptr, _addr(data), _addr(out), _stream(stream))


_LIB = None


def load(path=LIB_PATH):
    """The process-wide library instance."""
    global _LIB
    if _LIB is None:
        _LIB = DolHip(path)
    return _LIB
