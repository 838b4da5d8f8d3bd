"""voxelhashmt_amd -- MI355X-native voxel-hash TSDF fusion (libvhm.so).

Thin ctypes view of the C-ABI in include/vhm.h, used by the tests, the
benchmark and Python callers.  The product is the HIP library; this module
only marshals arguments.  Loading fails loudly when libvhm.so is missing --
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libvhm.so"

VHM_HOST, VHM_DEVICE = 0, 1
PRESET_REFERENCE_CPU, PRESET_REFERENCE_CUDA = 0, 1
FLAG_PREFILTERED = 1
STATUS = {0: "VHM_OK", 1: "VHM_ERR_INVALID_ARGUMENT", 2: "VHM_ERR_HIP", 3: "VHM_ERR_CAPACITY",
          4: "VHM_ERR_KEY_RANGE", 5: "VHM_ERR_NO_DEVICE", 6: "VHM_ERR_BUFFER_TOO_SMALL", 7: "VHM_ERR_IO"}
BLOCK_BYTES = 147456  # 512 voxels x 6 directions x 3 slots x 16 B
VOXEL_DTYPE = np.dtype([("weight", "<f4"), ("surface_coord", "<f4"), ("plane_offset", "<f4"),
                        ("packed_normal", "<u4")])


class Intrinsics(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("fx", C.c_float), ("fy", C.c_float),
                ("cx", C.c_float), ("cy", C.c_float)]


class Pose(C.Structure):
    _fields_ = [("R", C.c_float * 9), ("t", C.c_float * 3)]


class Params(C.Structure):
    _fields_ = [(n, t) for n, t in [
        ("voxel_size", C.c_float), ("truncation_distance", C.c_float),
        ("hash_log2_buckets", C.c_uint32), ("max_blocks", C.c_uint32),
        ("max_width", C.c_int32), ("max_height", C.c_int32),
        ("filter_sigma_spatial", C.c_float), ("filter_sigma_range", C.c_float),
        ("filter_sigma_angular", C.c_float), ("filter_disc_base", C.c_float),
        ("filter_disc_rel", C.c_float), ("filter_min_cos", C.c_float),
        ("max_weight", C.c_float), ("min_view_angle_weight", C.c_float),
        ("normal_sigma_spatial", C.c_float), ("normal_sigma_range", C.c_float),
        ("normal_sigma_angular", C.c_float), ("disc_base", C.c_float), ("disc_rel", C.c_float),
        ("layer_merge_scale", C.c_float), ("layer_extract_scale", C.c_float),
        ("layer_normal_cos", C.c_float)]]


class Primitive(C.Structure):
    _fields_ = [("kind", C.c_int32), ("c", C.c_float * 3), ("r", C.c_float), ("h", C.c_float * 3)]


class Scene(C.Structure):
    _fields_ = [("room_min", C.c_float * 3), ("room_max", C.c_float * 3), ("num_primitives", C.c_int32),
                ("primitives", C.POINTER(Primitive))]


class Noise(C.Structure):
    _fields_ = [("a", C.c_float), ("b", C.c_float), ("c", C.c_float), ("dropout", C.c_float),
                ("max_depth", C.c_float), ("seed", C.c_uint64), ("enabled", C.c_int32)]


_lib = None

# name -> (restype, argtypes)
_SIGS = {
    "vhm_default_params": (None, [C.POINTER(Params), C.c_int, C.c_float, C.c_float]),
    "vhm_query_device": (C.c_int, [C.c_char_p, C.c_size_t]),
    "vhm_create": (C.c_int, [C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "vhm_destroy": (C.c_int, [C.c_void_p]),
    "vhm_last_error": (C.c_char_p, [C.c_void_p]),
    "vhm_reset": (C.c_int, [C.c_void_p]),
    "vhm_integrate_f32": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(Intrinsics), C.POINTER(Pose),
                                    C.c_int, C.c_uint32]),
    "vhm_integrate_u16": (C.c_int, [C.c_void_p, C.c_void_p, C.c_float, C.c_int, C.POINTER(Intrinsics),
                                    C.POINTER(Pose), C.c_int, C.c_uint32]),
    "vhm_filter_depth_f32": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(Intrinsics), C.c_void_p,
                                       C.c_int]),
    "vhm_filter_depth_u16": (C.c_int, [C.c_void_p, C.c_void_p, C.c_float, C.c_int, C.POINTER(Intrinsics),
                                       C.c_void_p, C.c_int]),
    "vhm_allocate_active_blocks": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(Intrinsics),
                                             C.POINTER(Pose), C.c_int, C.c_void_p, C.c_size_t,
                                             C.POINTER(C.c_size_t)]),
    "vhm_synchronize": (C.c_int, [C.c_void_p]),
    "vhm_block_count": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t)]),
    "vhm_observed_voxel_count": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t)]),
    "vhm_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "vhm_block_keys": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "vhm_download_blocks": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "vhm_observed_block_keys": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "vhm_extract_mesh": (C.c_int, [C.c_void_p, C.c_float, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "vhm_extract_mesh_to": (C.c_int, [C.c_void_p, C.c_float, C.c_void_p, C.c_size_t, C.c_int,
                                      C.POINTER(C.c_size_t)]),
    "vhm_extract_mesh_incremental": (C.c_int, [C.c_void_p, C.c_float, C.c_void_p, C.c_void_p, C.c_size_t,
                                               C.POINTER(C.c_size_t), C.c_void_p, C.c_void_p, C.c_size_t, C.c_int,
                                               C.POINTER(C.c_size_t)]),
    "vhm_incremental_mesh_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                            C.POINTER(C.c_size_t)]),
    "vhm_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "vhm_kernel_times": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "vhm_synth_render": (C.c_int, [C.POINTER(Scene), C.POINTER(Intrinsics), C.POINTER(Pose), C.c_int, C.c_uint32,
                                   C.POINTER(Noise), C.c_void_p, C.c_int]),
    "vhm_philox4x32_10": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "vhm_denoiser_num_params": (C.c_int, []),
    "vhm_denoiser_create": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "vhm_denoiser_destroy": (C.c_int, [C.c_void_p]),
    "vhm_denoiser_last_error": (C.c_char_p, [C.c_void_p]),
    "vhm_denoiser_set_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "vhm_denoiser_get_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "vhm_denoiser_clear": (C.c_int, [C.c_void_p]),
    "vhm_denoiser_count": (C.c_int, [C.c_void_p]),
    "vhm_denoiser_append": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    "vhm_denoiser_forward": (C.c_int, [C.c_void_p, C.c_size_t, C.c_size_t, C.c_void_p]),
    "vhm_denoiser_train_step": (C.c_int, [C.c_void_p, C.c_size_t, C.c_size_t, C.c_float, C.POINTER(C.c_float),
                                          C.c_void_p]),
    "vhm_denoiser_apply": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t, C.c_float, C.c_float, C.c_float,
                                     C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "vhm_denoiser_samples": (C.c_int, [C.c_void_p, C.c_size_t, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p]),
    "vhm_mfma_layout_probe": (C.c_int, [C.c_void_p]),
    "vhm_batch_create": (C.c_int, [C.c_int, C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "vhm_batch_destroy": (C.c_int, [C.c_void_p]),
    "vhm_batch_last_error": (C.c_char_p, [C.c_void_p]),
    "vhm_batch_volume": (C.c_void_p, [C.c_void_p, C.c_int]),
    "vhm_batch_set_sequence": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_float,
                                         C.POINTER(Intrinsics), C.POINTER(Pose), C.c_size_t]),
    "vhm_batch_prepare": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "vhm_batch_step": (C.c_int, [C.c_void_p, C.c_int]),
    "vhm_batch_synchronize": (C.c_int, [C.c_void_p]),
    "vhm_batch_steps_done": (C.c_int, [C.c_void_p]),
    "vhm_batch_rewind": (C.c_int, [C.c_void_p]),
    "vhm_batch_timed_step": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_uint64)]),
    "vhm_batch_mesh_output": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                        C.c_void_p, C.c_size_t]),
    "vhm_extract_mesh_incremental_noalloc": (C.c_int, [C.c_void_p, C.c_float, C.c_void_p, C.c_void_p, C.c_size_t,
                                                       C.POINTER(C.c_size_t), C.c_void_p, C.c_void_p, C.c_size_t,
                                                       C.POINTER(C.c_size_t)]),
    "vhm_mfma_layout_probe_k32": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "vhm_mesh_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "vhm_block_digests": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "vhm_batch_mesh_items": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_size_t,
                                       C.POINTER(C.c_size_t), C.c_void_p, C.c_size_t]),
    "vhm_batch_autotune": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int),
                                     C.POINTER(C.c_int)]),
    "vhm_batch_autotune_tiles": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int), C.c_int,
                                           C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                           C.POINTER(C.c_int)]),
    "vhm_device_memory": (C.c_int, [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "vhm_integrate_point_cloud": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.POINTER(Pose)]),
    # include/vhm_io.h (host only)
    "vhm_io_last_error": (C.c_char_p, []),
    "vhm_demo_intrinsics": (None, [C.c_int32, C.c_int32, C.POINTER(Intrinsics)]),
    "vhm_demo_pose": (None, [C.c_int32, C.c_int32, C.POINTER(Pose)]),
    "vhm_demo_trace": (C.c_int, [C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "vhm_demo_renderer_create": (C.c_int, [C.c_uint32, C.POINTER(C.c_void_p)]),
    "vhm_demo_renderer_destroy": (None, [C.c_void_p]),
    "vhm_demo_render": (C.c_int, [C.c_void_p, C.POINTER(Intrinsics), C.POINTER(Pose), C.c_int32, C.c_uint32,
                                  C.c_void_p]),
    "vhm_write_ply": (C.c_int, [C.c_char_p, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "vhm_dataset_frame_ids": (C.c_int, [C.c_char_p, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "vhm_dataset_intrinsics": (C.c_int, [C.c_char_p, C.POINTER(Intrinsics), C.POINTER(C.c_float)]),
    "vhm_load_pose": (C.c_int, [C.c_char_p, C.POINTER(Pose)]),
    "vhm_quaternion_to_matrix": (None, [C.c_float, C.c_float, C.c_float, C.c_float, C.POINTER(C.c_float)]),
    "vhm_png_size": (C.c_int, [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "vhm_load_depth_png": (C.c_int, [C.c_char_p, C.POINTER(Intrinsics), C.c_float, C.c_void_p,
                                     C.POINTER(C.c_size_t)]),
    "vhm_load_depth_png_u16": (C.c_int, [C.c_char_p, C.POINTER(Intrinsics), C.c_float, C.c_void_p,
                                         C.POINTER(C.c_size_t)]),
    "vhm_load_dataset_frame": (C.c_int, [C.c_char_p, C.c_int32, C.c_char_p, C.POINTER(Intrinsics), C.c_float,
                                         C.POINTER(Pose), C.c_void_p, C.POINTER(C.c_size_t)]),
}


def exported_symbols():
    """Names the C-ABI must export (every function declared in include/*.h)."""
    return sorted(_SIGS)


def lib():
    """Loads libvhm.so (build it with `python -m voxelhashmt_amd.build`)."""
    global _lib
    if _lib is None:
        path = os.environ.get("VHM_LIBRARY", str(LIB_PATH))
        if not Path(path).exists():
            raise RuntimeError(f"libvhm.so not built ({path}); run `python -m voxelhashmt_amd.build` "
                               "(there is no CPU fallback)")
        # One HSA runtime per process: torch's ROCm wheel brings its own HIP
        # and HSA runtimes, and an HSA runtime loaded before it (ours, via
        # libvhm.so) makes torch find no GPU.  Loading torch first lets
        # libvhm.so's HIP runtime bind to the HSA runtime torch loaded.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(path)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class VhmError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{STATUS.get(code, code)}: {message}")
        self.code = code


def _ptr(a) -> int:
    """Address of a numpy array (host) or a torch tensor (its device pointer)."""
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return a.ctypes.data
    return a.data_ptr()


def _kind(a) -> int:
    if isinstance(a, np.ndarray):
        return VHM_HOST
    return VHM_DEVICE if a.is_cuda else VHM_HOST


def make_intrinsics(width, height, fx, fy, cx, cy) -> Intrinsics:
    return Intrinsics(int(width), int(height), float(fx), float(fy), float(cx), float(cy))


def make_pose(R, t) -> Pose:
    p = Pose()
    R = np.asarray(R, dtype=np.float32).reshape(9)
    t = np.asarray(t, dtype=np.float32).reshape(3)
    for i in range(9):
        p.R[i] = float(R[i])
    for i in range(3):
        p.t[i] = float(t[i])
    return p


def pose_array(pose: Pose) -> np.ndarray:
    """12 floats: R row-major then t."""
    return np.array(list(pose.R) + list(pose.t), dtype=np.float32)


def default_params(voxel_size: float, truncation: float, preset: int = PRESET_REFERENCE_CPU, **overrides) -> Params:
    p = Params()
    lib().vhm_default_params(C.byref(p), preset, voxel_size, truncation)
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def device_available() -> tuple[bool, str]:
    buf = C.create_string_buffer(256)
    ok = lib().vhm_query_device(buf, 256)
    return bool(ok), buf.value.decode()


class Volume:
    """One GPU volume (a vhm_ctx): the drop-in for a VoxelHashTSDF driven by
    TsdfIntegrationBackend's GPU path."""

    def __init__(self, params: Params):
        self.params = params
        self._owned = True
        self._h = C.c_void_p()
        rc = lib().vhm_create(C.byref(params), C.byref(self._h))
        if rc != 0:
            raise VhmError(rc, "vhm_create failed")

    @classmethod
    def _borrow(cls, handle: int, params: Params, owner) -> "Volume":
        """A view of a volume owned by a Batch (not destroyed by close())."""
        v = cls.__new__(cls)
        v.params, v._owned, v._owner = params, False, owner
        v._h = C.c_void_p(handle)
        return v

    def close(self):
        if self._h:
            if self._owned:
                lib().vhm_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise VhmError(rc, lib().vhm_last_error(self._h).decode())

    def integrate(self, depth, K: Intrinsics, T_wc: Pose, stride: int = 1, depth_scale: float = 0.001,
                  prefiltered: bool = False):
        flags = FLAG_PREFILTERED if prefiltered else 0
        if str(getattr(depth, "dtype", "")) in ("uint16", "torch.uint16", "torch.int16"):
            self._check(lib().vhm_integrate_u16(self._h, _ptr(depth), depth_scale, _kind(depth), C.byref(K),
                                                C.byref(T_wc), stride, flags))
        else:
            self._check(lib().vhm_integrate_f32(self._h, _ptr(depth), _kind(depth), C.byref(K), C.byref(T_wc),
                                                stride, flags))

    def integrate_point_cloud(self, points, T_wc: Pose):
        """integratePointCloud (voxel_hash_tsdf.h:75-97): n x 3 float32
        camera-space points (numpy, or a torch CUDA tensor)."""
        if isinstance(points, np.ndarray):
            points = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        n = int(points.shape[0])
        self._check(lib().vhm_integrate_point_cloud(self._h, _ptr(points), n, _kind(points), C.byref(T_wc)))

    def filter_depth(self, depth, K: Intrinsics, depth_scale: float = 0.001) -> np.ndarray:
        out = np.zeros((K.height, K.width), dtype=np.float32)
        if str(getattr(depth, "dtype", "")) in ("uint16", "torch.uint16", "torch.int16"):
            self._check(lib().vhm_filter_depth_u16(self._h, _ptr(depth), depth_scale, _kind(depth), C.byref(K),
                                                   out.ctypes.data, VHM_HOST))
        else:
            self._check(lib().vhm_filter_depth_f32(self._h, _ptr(depth), _kind(depth), C.byref(K),
                                                   out.ctypes.data, VHM_HOST))
        return out

    def allocate_active_blocks(self, depth, K: Intrinsics, T_wc: Pose, stride: int = 1) -> np.ndarray:
        n = C.c_size_t()
        cap = 1 << int(self.params.hash_log2_buckets)  # a frame activates at most one key per bucket
        out = np.zeros((cap, 3), dtype=np.int32)
        self._check(lib().vhm_allocate_active_blocks(self._h, _ptr(depth), _kind(depth), C.byref(K),
                                                     C.byref(T_wc), stride, out.ctypes.data, cap, C.byref(n)))
        return out[: n.value].copy()

    def synchronize(self):
        self._check(lib().vhm_synchronize(self._h))

    def block_count(self) -> int:
        n = C.c_size_t()
        self._check(lib().vhm_block_count(self._h, C.byref(n)))
        return n.value

    def observed_voxel_count(self) -> int:
        n = C.c_size_t()
        self._check(lib().vhm_observed_voxel_count(self._h, C.byref(n)))
        return n.value

    def mesh_items_staged(self) -> int:
        """(block, direction) items marching cubes has staged since creation/reset."""
        n = C.c_uint64()
        self._check(lib().vhm_mesh_stats(self._h, C.byref(n)))
        return n.value

    def stats(self) -> dict:
        u, v, f = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self._check(lib().vhm_stats(self._h, C.byref(u), C.byref(v), C.byref(f)))
        return {"updates": u.value, "visited": v.value, "frames": f.value}

    def block_keys(self) -> np.ndarray:
        n = self.block_count()
        out = np.zeros((max(n, 1), 3), dtype=np.int32)
        m = C.c_size_t()
        self._check(lib().vhm_block_keys(self._h, out.ctypes.data, n, C.byref(m)))
        return out[:n]

    def block_digests(self, keys: np.ndarray) -> np.ndarray:
        """Per key, the device digest of the block's AoS words (sum w_i P^i mod
        2^64, P = 0x100000001B3; 0 for an unallocated key)."""
        k = np.ascontiguousarray(keys, np.int32).reshape(-1, 3)
        out = np.zeros(len(k), np.uint64)
        if len(k):
            self._check(lib().vhm_block_digests(self._h, k.ctypes.data, len(k), out.ctypes.data))
        return out

    def observed_block_keys(self) -> np.ndarray:
        """Keys of blocks with any observed voxel (pool order)."""
        n = C.c_size_t()
        self._check(lib().vhm_observed_block_keys(self._h, None, 0, C.byref(n)))
        out = np.zeros((n.value, 3), np.int32)
        if n.value:
            m = C.c_size_t()
            self._check(lib().vhm_observed_block_keys(self._h, out.ctypes.data, n.value, C.byref(m)))
        return out

    def reset(self):
        """Clears the volume, keeping its reservations."""
        self._check(lib().vhm_reset(self._h))

    def download_blocks(self, keys: np.ndarray) -> np.ndarray:
        """Blocks in the reference's AoS layout: (n, 512, 6, 3) of VOXEL_DTYPE."""
        keys = np.ascontiguousarray(keys, dtype=np.int32).reshape(-1, 3)
        out = np.zeros((len(keys), 512, 6, 3), dtype=VOXEL_DTYPE)
        if len(keys):
            self._check(lib().vhm_download_blocks(self._h, keys.ctypes.data, len(keys), out.ctypes.data))
        return out

    def extract_mesh(self, iso_level: float = 0.0, device: bool = False):
        """Triangles (n, 3, 3) float32: a numpy array, or with device=True a
        torch tensor on the current GPU written in place by the emit pass
        (vhm_extract_mesh_to, no host copy)."""
        if device:
            return self._extract_mesh_device(iso_level)
        n = C.c_size_t()
        self._check(lib().vhm_extract_mesh(self._h, iso_level, None, 0, C.byref(n)))
        out = np.zeros((max(n.value, 1), 3, 3), dtype=np.float32)
        m = C.c_size_t()
        if n.value:
            self._check(lib().vhm_extract_mesh(self._h, iso_level, out.ctypes.data, n.value, C.byref(m)))
        return out[: n.value]

    def _extract_mesh_device(self, iso_level: float):
        import torch
        # capacity: the last mesh's size, else a guess from the block count;
        # a mesh larger than the guess is extracted once more into a buffer
        # of its size
        cap = max(1 << 20, getattr(self, "_tri_hint", 0), 48 * self.block_count())
        for _ in range(2):
            out = torch.empty((cap, 3, 3), dtype=torch.float32, device="cuda")
            n = C.c_size_t()
            self._check(lib().vhm_extract_mesh_to(self._h, iso_level, out.data_ptr(), cap, VHM_DEVICE, C.byref(n)))
            if n.value <= cap:
                self._tri_hint = n.value
                return out[: n.value]
            cap = n.value
        raise VhmError(3, "device mesh extraction: mesh grew between two calls")

    def extract_mesh_incremental(self, iso_level: float = 0.0, device_tris=None, device_items=None):
        """Re-meshes the (block, direction) pairs changed since the last
        incremental call.  Returns (block_keys (m,3), dirs (m,) direction bit
        masks, triangles (n,3,3), items (n,): d * m + i for direction d of
        listed block i).  With device_tris (a torch CUDA float tensor of
        n_max*9 elements) and device_items (int32, n_max) the triangles stay
        on the GPU (each item's triangles one contiguous run, the runs in the
        order the items finished) and (m, n) is returned."""
        nb, nt = C.c_size_t(), C.c_size_t()
        if device_tris is not None:
            cap = device_tris.numel() // 9
            self._check(lib().vhm_extract_mesh_incremental(
                self._h, iso_level, None, None, 0, C.byref(nb), device_tris.data_ptr(),
                None if device_items is None else device_items.data_ptr(), cap, VHM_DEVICE, C.byref(nt)))
            return nb.value, nt.value
        cap_b = max(self.block_count(), 1)
        keys = np.zeros((cap_b, 3), np.int32)
        dirs = np.zeros(cap_b, np.uint8)
        cap_t = 1 << 20
        tris = np.zeros((cap_t, 3, 3), np.float32)
        items = np.zeros(cap_t, np.uint32)
        self._check(lib().vhm_extract_mesh_incremental(self._h, iso_level, keys.ctypes.data, dirs.ctypes.data, cap_b,
                                                       C.byref(nb), tris.ctypes.data, items.ctypes.data, cap_t,
                                                       VHM_HOST, C.byref(nt)))
        if nt.value > cap_t:  # fetch everything from the library's copy
            tris = np.zeros((nt.value, 3, 3), np.float32)
            items = np.zeros(nt.value, np.uint32)
            n = C.c_size_t()
            self._check(lib().vhm_incremental_mesh_copy(self._h, tris.ctypes.data, items.ctypes.data, nt.value,
                                                        C.byref(n)))
        return keys[: nb.value].copy(), dirs[: nb.value].copy(), tris[: nt.value].copy(), items[: nt.value].copy()

    TIMING_STAGES = ("filter", "allocate", "integrate", "mesh")

    def set_timing(self, enabled: bool, stages=None):
        """Per-stage GPU timing with HIP events (vhm_set_timing); stages: the
        names to time (default all), each costing two events per launch."""
        mode = int(bool(enabled))
        if enabled and stages is not None:
            mode = 0x100
            for s in stages:
                mode |= 1 << self.TIMING_STAGES.index(s)
        self._check(lib().vhm_set_timing(self._h, mode))

    def kernel_times(self) -> dict:
        ms = (C.c_double * 4)()
        cnt = (C.c_uint64 * 4)()
        self._check(lib().vhm_kernel_times(self._h, ms, cnt))
        names = ["filter", "allocate", "integrate", "mesh"]
        return {n: {"ms": ms[i], "launches": cnt[i]} for i, n in enumerate(names)}


class Denoiser:
    """Learned per-block denoiser on the GPU (vhm_denoiser_*): a device sample
    set, BF16 MFMA forward/backward, Adam, write-back with fallback."""

    NUM_PARAMS = 43270

    def __init__(self, capacity: int, batch_capacity: int = 4096):
        self._h = C.c_void_p()
        rc = lib().vhm_denoiser_create(int(capacity), int(batch_capacity), C.byref(self._h))
        if rc != 0:
            raise VhmError(rc, "vhm_denoiser_create failed")
        self.capacity, self.batch_capacity = capacity, batch_capacity

    def close(self):
        if self._h:
            lib().vhm_denoiser_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise VhmError(rc, lib().vhm_denoiser_last_error(self._h).decode())

    def set_params(self, params: np.ndarray):
        p = np.ascontiguousarray(params, np.float32).reshape(-1)
        assert p.size == self.NUM_PARAMS
        self._check(lib().vhm_denoiser_set_params(self._h, p.ctypes.data))

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.NUM_PARAMS, np.float32)
        self._check(lib().vhm_denoiser_get_params(self._h, p.ctypes.data))
        return p

    def clear(self):
        self._check(lib().vhm_denoiser_clear(self._h))

    @property
    def count(self) -> int:
        return lib().vhm_denoiser_count(self._h)

    def append(self, input_vol: Volume, keys: np.ndarray, target_vol: Volume | None = None):
        k = np.ascontiguousarray(keys, np.int32).reshape(-1, 3)
        self._check(lib().vhm_denoiser_append(self._h, input_vol._h, target_vol._h if target_vol else None,
                                              k.ctypes.data, len(k)))

    def forward(self, offset: int = 0, n: int | None = None, fetch: bool = True):
        n = self.count - offset if n is None else n
        y = np.zeros((n, 512, 6), np.float32) if fetch else None
        self._check(lib().vhm_denoiser_forward(self._h, offset, n, y.ctypes.data if fetch else None))
        return y

    def train_step(self, offset: int, n: int, lr: float, want_grad: bool = False):
        loss = C.c_float()
        g = np.zeros(self.NUM_PARAMS, np.float32) if want_grad else None
        self._check(lib().vhm_denoiser_train_step(self._h, offset, n, lr, C.byref(loss),
                                                  g.ctypes.data if want_grad else None))
        return (loss.value, g) if want_grad else loss.value

    def apply(self, vol: Volume, offset: int = 0, n: int | None = None, min_weight: float = 0.0,
              max_delta: float = 0.5, block_max_delta: float = 0.25) -> dict:
        n = self.count - offset if n is None else n
        a, b = C.c_uint64(), C.c_uint64()
        self._check(lib().vhm_denoiser_apply(self._h, vol._h, offset, n, min_weight, max_delta, block_max_delta,
                                             C.byref(a), C.byref(b)))
        return {"cells_replaced": a.value, "blocks_kept_plain": b.value}

    def samples(self, offset: int, n: int):
        x = np.zeros((n, 10, 10, 10, 16), np.uint16)
        t = np.zeros((n, 512, 6), np.float32)
        m = np.zeros((n, 512, 6), np.float32)
        self._check(lib().vhm_denoiser_samples(self._h, offset, n, x.ctypes.data, t.ctypes.data, m.ctypes.data))
        return x, t, m


class Batch:
    """Batched multi-scene fusion (vhm_batch_*): S volumes advance one frame
    per step, one launch per stage for all scenes, the step replayed as a
    captured hipGraph."""

    def __init__(self, params_list):
        self.params = list(params_list)
        arr = (Params * len(self.params))(*self.params)
        self._h = C.c_void_p()
        rc = lib().vhm_batch_create(len(self.params), arr, C.byref(self._h))
        if rc != 0:
            raise VhmError(rc, "vhm_batch_create failed")
        self._keep = []  # device frame tensors must outlive the steps
        self.volumes = [Volume._borrow(lib().vhm_batch_volume(self._h, i), p, self)
                        for i, p in enumerate(self.params)]

    def close(self):
        if self._h:
            lib().vhm_batch_destroy(self._h)
            self._h = C.c_void_p()
            self._keep = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise VhmError(rc, lib().vhm_batch_last_error(self._h).decode())

    def set_sequence(self, scene: int, frames, K: Intrinsics, poses, depth_scale: float = 0.001,
                     tri_capacity: int = 1 << 20):
        """frames: device int16/uint16 tensor (n, H, W)."""
        if not getattr(frames, "is_cuda", False):
            raise ValueError("batch frames must be a device tensor")
        n = int(frames.shape[0])
        P = (Pose * n)(*poses[:n])
        self._keep.append(frames)
        self._check(lib().vhm_batch_set_sequence(self._h, scene, frames.data_ptr(), n, depth_scale, C.byref(K), P,
                                                 tri_capacity))

    def prepare(self, integrate_grid: int = 0, mesh_grid: int = 0, use_graph: bool = True):
        self._check(lib().vhm_batch_prepare(self._h, integrate_grid, mesh_grid, int(use_graph)))

    def step(self, n: int = 1):
        self._check(lib().vhm_batch_step(self._h, n))

    def synchronize(self):
        self._check(lib().vhm_batch_synchronize(self._h))

    def rewind(self):
        self._check(lib().vhm_batch_rewind(self._h))

    @property
    def steps_done(self) -> int:
        return lib().vhm_batch_steps_done(self._h)

    def mesh_output(self, scene: int, fetch: bool = True):
        """(blocks re-meshed, triangle total, triangles (n,3,3) or None) of
        the last step's incremental mesh of one scene."""
        blocks, tris = C.c_uint32(), C.c_uint32()
        self._check(lib().vhm_batch_mesh_output(self._h, scene, C.byref(blocks), C.byref(tris), None, 0))
        out = None
        if fetch:
            out = np.zeros((tris.value, 3, 3), np.float32)
            if tris.value:
                self._check(lib().vhm_batch_mesh_output(self._h, scene, None, None, out.ctypes.data, tris.value))
        return blocks.value, tris.value, out

    def mesh_items(self, scene: int):
        """(block keys (m,3), direction masks (m,), triangle items (n,)) of the
        last step's incremental mesh of one scene (as extract_mesh_incremental)."""
        blocks, tris, _ = self.mesh_output(scene, fetch=False)
        keys = np.zeros((max(blocks, 1), 3), np.int32)
        dirs = np.zeros(max(blocks, 1), np.uint8)
        items = np.zeros(max(tris, 1), np.uint32)
        nb = C.c_size_t()
        self._check(lib().vhm_batch_mesh_items(self._h, scene, keys.ctypes.data, dirs.ctypes.data, blocks,
                                               C.byref(nb), items.ctypes.data, tris))
        return keys[:blocks], dirs[:blocks], items[:tris]

    def timed_step(self) -> dict:
        """One step with direct launches, timed per stage (synchronous)."""
        ms = (C.c_float * 3)()
        u, r = C.c_uint64(), C.c_uint64()
        self._check(lib().vhm_batch_timed_step(self._h, ms, C.byref(u), C.byref(r)))
        return {"front_ms": ms[0], "integrate_ms": ms[1], "mesh_ms": ms[2], "updates": u.value,
                "blocks_remeshed": r.value}

    def autotune(self, candidates=(2, 4, 8, 16)) -> tuple[int, int]:
        c = (C.c_int * len(candidates))(*candidates)
        gi, gm = C.c_int(), C.c_int()
        self._check(lib().vhm_batch_autotune(self._h, c, len(candidates), C.byref(gi), C.byref(gm)))
        return gi.value, gm.value

    def autotune_tiles(self, filter_widths=(32, 16, 64), alloc_widths=(16, 32, 64),
                       chunks=(0, 16, 64)) -> tuple[int, int, int]:
        """Filter / allocation tile widths and the integrator's chunk
        (vhm_batch_autotune_tiles); consumes max of the list lengths steps."""
        f = (C.c_int * len(filter_widths))(*filter_widths)
        a = (C.c_int * len(alloc_widths))(*alloc_widths)
        c = (C.c_int * len(chunks))(*chunks)
        fo, ao, co = C.c_int(), C.c_int(), C.c_int()
        self._check(lib().vhm_batch_autotune_tiles(self._h, f, len(filter_widths), a, len(alloc_widths), c,
                                                   len(chunks), C.byref(fo), C.byref(ao), C.byref(co)))
        return fo.value, ao.value, co.value


def mfma_layout_probe() -> np.ndarray:
    out = np.zeros(256, np.float32)
    rc = lib().vhm_mfma_layout_probe(out.ctypes.data)
    if rc != 0:
        raise VhmError(rc, "vhm_mfma_layout_probe failed")
    return out.reshape(64, 4)
