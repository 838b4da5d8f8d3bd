"""Host-side companions of the fusion path (include/vhm_io.h), in libvhm.so.

* The reference demo's synthetic scene (src/main.cpp:177-335, 400-432):
  `demo_intrinsics`, `demo_pose`, `demo_trace`, `DemoRenderer`.
* Its binary PLY writer (main.cpp:337-398): `write_ply`, and `read_ply` to
  read such a file back.
* Its dataset loader (include/tsdfmc/dataset_io.h): `dataset_frame_ids`,
  `dataset_intrinsics`, `load_pose`, `quaternion_to_matrix`, `png_size`,
  `load_depth_png`, `load_depth_png_u16`, `load_dataset_frame`.

Errors raise `voxelhashmt_amd.VhmError` carrying the reference's exception
text.  Nothing here uses the GPU.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import Intrinsics, Pose, VhmError, lib

DEMO_NOISE = 1


def _check(rc: int):
    if rc != 0:
        raise VhmError(rc, lib().vhm_io_last_error().decode())


def _b(path) -> bytes:
    return str(path).encode()


def demo_intrinsics(width: int = 128, height: int = 96) -> Intrinsics:
    K = Intrinsics()
    lib().vhm_demo_intrinsics(width, height, C.byref(K))
    return K


def demo_pose(frame_index: int, frame_count: int) -> Pose:
    T = Pose()
    lib().vhm_demo_pose(frame_index, frame_count, C.byref(T))
    return T


def demo_trace(origin, direction):
    """Nearest hit parameter of the demo scene along a ray, or None."""
    o = (C.c_float * 3)(*origin)
    d = (C.c_float * 3)(*direction)
    t = C.c_float(0.0)
    return float(t.value) if lib().vhm_demo_trace(o, d, C.byref(t)) else None


class DemoRenderer:
    """renderSyntheticDepth with the demo's persistent std::mt19937 (seed 42)."""

    def __init__(self, seed: int = 42):
        self._h = C.c_void_p()
        _check(lib().vhm_demo_renderer_create(seed, C.byref(self._h)))

    def close(self):
        if self._h:
            lib().vhm_demo_renderer_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, K: Intrinsics, T_wc: Pose, frame_index: int, noise: bool = True) -> np.ndarray:
        out = np.zeros((K.height, K.width), np.float32)
        _check(lib().vhm_demo_render(self._h, C.byref(K), C.byref(T_wc), frame_index,
                                     DEMO_NOISE if noise else 0, out.ctypes.data))
        return out


def write_ply(path, triangles: np.ndarray) -> int:
    """Writes n x 9 float32 triangles; returns the faces written."""
    t = np.ascontiguousarray(triangles, dtype=np.float32).reshape(-1, 9)
    n = C.c_size_t(0)
    _check(lib().vhm_write_ply(_b(path), t.ctypes.data, len(t), C.byref(n)))
    return n.value


def read_ply(path) -> tuple[np.ndarray, np.ndarray]:
    """Reads a file written by write_ply (or the reference demo): returns
    (triangles n x 9, per-face unit normals n x 3), in file order."""
    data = Path(path).read_bytes()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    header = data[:end].decode().splitlines()
    nv = int(next(l.split()[2] for l in header if l.startswith("element vertex")))
    nf = int(next(l.split()[2] for l in header if l.startswith("element face")))
    verts = np.frombuffer(data, np.float32, nv * 6, end).reshape(-1, 6)
    faces = np.frombuffer(data, np.uint8, nf * 13, end + nv * 24).reshape(-1, 13)
    if nf and (not np.all(faces[:, 0] == 3) or
               not np.array_equal(faces[:, 1:].copy().view("<i4").reshape(-1), np.arange(3 * nf, dtype=np.int32))):
        raise ValueError(f"{path}: not an unindexed triangle PLY")
    tris = verts[:, :3].reshape(nf, 9)
    normals = verts[0::3, 3:6] if nf else np.zeros((0, 3), np.float32)
    return tris.copy(), normals.copy()


def dataset_frame_ids(root) -> list[int]:
    n = C.c_size_t(0)
    _check(lib().vhm_dataset_frame_ids(_b(root), None, 0, C.byref(n)))
    ids = np.zeros(n.value, np.int32)
    _check(lib().vhm_dataset_frame_ids(_b(root), ids.ctypes.data, len(ids), C.byref(n)))
    return [int(i) for i in ids]


def dataset_intrinsics(path) -> tuple[Intrinsics, float]:
    K = Intrinsics()
    s = C.c_float(0.0)
    _check(lib().vhm_dataset_intrinsics(_b(path), C.byref(K), C.byref(s)))
    return K, float(s.value)


def load_pose(path) -> Pose:
    T = Pose()
    _check(lib().vhm_load_pose(_b(path), C.byref(T)))
    return T


def quaternion_to_matrix(w: float, x: float, y: float, z: float) -> np.ndarray:
    R = (C.c_float * 9)()
    lib().vhm_quaternion_to_matrix(w, x, y, z, R)
    return np.array(R, np.float32).reshape(3, 3)


def png_size(path) -> tuple[int, int]:
    w, h = C.c_int32(0), C.c_int32(0)
    _check(lib().vhm_png_size(_b(path), C.byref(w), C.byref(h)))
    return w.value, h.value


def load_depth_png(path, K: Intrinsics, depth_scale: float) -> tuple[np.ndarray, int]:
    out = np.zeros((K.height, K.width), np.float32)
    valid = C.c_size_t(0)
    _check(lib().vhm_load_depth_png(_b(path), C.byref(K), depth_scale, out.ctypes.data, C.byref(valid)))
    return out, valid.value


def load_depth_png_u16(path, K: Intrinsics, depth_scale: float) -> tuple[np.ndarray, int]:
    out = np.zeros((K.height, K.width), np.uint16)
    valid = C.c_size_t(0)
    _check(lib().vhm_load_depth_png_u16(_b(path), C.byref(K), depth_scale, out.ctypes.data, C.byref(valid)))
    return out, valid.value


def load_dataset_frame(root, frame_id: int, pose_filename: str, K: Intrinsics,
                       depth_scale: float) -> tuple[Pose, np.ndarray, int]:
    T = Pose()
    out = np.zeros((K.height, K.width), np.float32)
    valid = C.c_size_t(0)
    _check(lib().vhm_load_dataset_frame(_b(root), frame_id, pose_filename.encode(), C.byref(K), depth_scale,
                                        C.byref(T), out.ctypes.data, C.byref(valid)))
    return T, out, valid.value
