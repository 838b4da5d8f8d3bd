"""ctypes bindings of the CPU checkers (test infrastructure only).

  * liboracle.so  -- the restated reference CPU pipeline (oracle/cpu_ref.hpp),
                     float (parity) and double (FP64 report) instantiations.
  * _ref/libvhm_ref.so -- the reference's own CPU path compiled from its
                     sources (oracle/Makefile `ref`); present only where the
                     reference tree was available at build time.
  * _ref/libdemo_ref.so -- the reference demo's and dataset loader's own
                     functions (oracle/demo_harness.cpp), likewise.
  * _ref/tsdf_voxel_hash_demo -- the reference's CPU-only demo executable.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parents[1] / "oracle"
ORACLE_LIB = ORACLE_DIR / "liboracle.so"
REF_LIB = ORACLE_DIR / "_ref" / "libvhm_ref.so"
DEMO_REF_LIB = ORACLE_DIR / "_ref" / "libdemo_ref.so"
REF_DEMO_EXE = ORACLE_DIR / "_ref" / "tsdf_voxel_hash_demo"


class OracleParams(C.Structure):
    _fields_ = [("voxel_size", C.c_float), ("truncation", C.c_float), ("filter_radius", C.c_int),
                ("filter_sigma_spatial", C.c_float), ("filter_sigma_range", C.c_float),
                ("filter_sigma_angular", C.c_float), ("filter_disc_base", C.c_float),
                ("filter_disc_rel", C.c_float), ("filter_min_cos", C.c_float), ("max_weight", C.c_float),
                ("min_view_angle_weight", C.c_float), ("normal_sigma_spatial", C.c_float),
                ("normal_sigma_range", C.c_float), ("normal_sigma_angular", C.c_float),
                ("disc_base", C.c_float), ("disc_rel", C.c_float), ("layer_merge_scale", C.c_float),
                ("layer_extract_scale", C.c_float), ("layer_normal_cos", C.c_float)]


_oracle = None
_ref = None
_demo_ref = None

VP = C.c_void_p
FP = C.POINTER(C.c_float)


def oracle():
    global _oracle
    if _oracle is None:
        if not ORACLE_LIB.exists():
            raise RuntimeError(f"{ORACLE_LIB} missing: run `make -C oracle`")
        L = C.CDLL(str(ORACLE_LIB))
        sig = {
            "vor_sizeof_params": (C.c_size_t, []),
            "vor_default_params": (None, [C.POINTER(OracleParams)]),
            "vor_create": (VP, [C.POINTER(OracleParams), C.c_int]),
            "vor_destroy": (None, [VP]),
            "vor_filter": (None, [C.POINTER(OracleParams), C.c_int, VP, C.c_int, C.c_int, VP, VP]),
            "vor_integrate": (None, [VP, VP, C.c_int, C.c_int, VP, VP, C.c_int, C.c_int]),
            "vor_integrate_point_cloud": (None, [VP, VP, C.c_size_t, VP]),
            "vor_prepare": (C.c_size_t, [VP, VP, C.c_int, C.c_int, VP, VP, C.c_int, VP, C.c_size_t]),
            "vor_integrate_blocks": (None, [VP, VP, C.c_size_t, VP, C.c_int, C.c_int, VP, VP]),
            "vor_block_count": (C.c_size_t, [VP]),
            "vor_observed_voxel_count": (C.c_size_t, [VP]),
            "vor_update_count": (C.c_uint64, [VP]),
            "vor_block_keys": (C.c_size_t, [VP, VP, C.c_size_t]),
            "vor_dump_blocks_soa": (None, [VP, VP, C.c_size_t, VP, VP, VP, VP]),
            "vor_dump_blocks": (None, [VP, VP, C.c_size_t, VP]),
            "vor_load_blocks": (None, [VP, VP, C.c_size_t, VP]),
            "vor_extract_mesh": (C.c_size_t, [VP, C.c_float, VP, C.c_size_t]),
            "vor_pack_normal": (C.c_uint32, [VP]),
            "vor_unpack_normal": (None, [C.c_uint32, VP]),
            "vor_observation_weight": (C.c_float, [C.c_float, C.c_float, C.c_float, C.c_float]),
            "vor_dominant_direction": (C.c_int, [VP]),
            "vor_complete_cube": (C.c_int, [VP, VP]),
            "vor_polygonise": (C.c_int, [VP, VP, C.c_float, VP]),
            "vor_sample_bilinear": (C.c_int, [C.POINTER(OracleParams), VP, C.c_int, C.c_int, VP, C.c_float,
                                              C.c_float, VP]),
            "vor_surface_normal": (C.c_int, [C.POINTER(OracleParams), VP, C.c_int, C.c_int, VP, C.c_float,
                                             C.c_float, VP]),
            "vor_integrate_sample": (None, [VP, C.c_int, C.c_float, VP, VP, C.c_float, C.c_float, C.c_float]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype, f.argtypes = r, a
        assert L.vor_sizeof_params() == C.sizeof(OracleParams)
        _oracle = L
    return _oracle


def ref_available() -> bool:
    return REF_LIB.exists()


def ref():
    global _ref
    if _ref is None:
        if not REF_LIB.exists():
            raise RuntimeError(f"{REF_LIB} missing: run `make -C oracle ref` where the reference tree exists")
        L = C.CDLL(str(REF_LIB))
        sig = {
            "vref_volume_create": (VP, [C.c_float, C.c_float]),
            "vref_volume_destroy": (None, [VP]),
            "vref_integrate_point_cloud": (None, [VP, VP, C.c_size_t, VP]),
            "vref_integrate": (None, [VP, VP, C.c_int, C.c_int, VP, VP, C.c_int, C.c_int]),
            "vref_filter": (None, [VP, C.c_int, C.c_int, VP, VP]),
            "vref_pixel_normal": (C.c_int, [VP, C.c_int, C.c_int, VP, C.c_int, C.c_int, VP]),
            "vref_prepare": (C.c_size_t, [VP, VP, C.c_int, C.c_int, VP, VP, C.c_int, VP, C.c_size_t]),
            "vref_integrate_blocks": (None, [VP, VP, C.c_size_t, VP, C.c_int, C.c_int, VP, VP]),
            "vref_block_count": (C.c_size_t, [VP]),
            "vref_observed_voxel_count": (C.c_size_t, [VP]),
            "vref_block_keys": (C.c_size_t, [VP, VP, C.c_size_t]),
            "vref_dump_blocks": (None, [VP, VP, C.c_size_t, VP]),
            "vref_load_blocks": (None, [VP, VP, C.c_size_t, VP]),
            "vref_extract_mesh": (C.c_size_t, [VP, C.c_float, VP, C.c_size_t]),
            "vref_sizeof_voxel": (C.c_size_t, []),
            "vref_gaussian_weight": (C.c_float, [C.c_float, C.c_float]),
            "vref_expf_array": (None, [VP, VP, C.c_size_t]),
            "vref_pack_normal": (C.c_uint32, [VP]),
            "vref_unpack_normal": (None, [C.c_uint32, VP]),
            "vref_observation_weight": (C.c_float, [C.c_float, C.c_float, C.c_float, C.c_float]),
            "vref_dominant_direction": (C.c_int, [VP]),
            "vref_sample_bilinear": (C.c_int, [VP, VP, C.c_int, C.c_int, VP, C.c_float, C.c_float, VP]),
            "vref_estimate_normal": (C.c_int, [VP, VP, C.c_int, C.c_int, VP, C.c_float, C.c_float, VP]),
            "vref_complete_cube": (C.c_int, [VP, VP]),
            "vref_polygonise": (C.c_int, [VP, VP, C.c_float, VP]),
            "vref_integrate_sample": (None, [VP, C.c_int, C.c_float, VP, VP, C.c_float, C.c_float, C.c_float]),
            "vref_mc_tables": (None, [VP, VP]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype, f.argtypes = r, a
        _ref = L
    return _ref


def demo_ref_available() -> bool:
    return DEMO_REF_LIB.exists()


def demo_ref():
    """The reference demo / dataset-loader functions (oracle/demo_harness.cpp)."""
    global _demo_ref
    if _demo_ref is None:
        if not DEMO_REF_LIB.exists():
            raise RuntimeError(f"{DEMO_REF_LIB} missing: run `make -C oracle ref` where the reference tree exists")
        L = C.CDLL(str(DEMO_REF_LIB))
        CP = C.c_char_p
        sig = {
            "ref_demo_look_at": (None, [FP, FP, FP, FP]),
            "ref_demo_trace": (C.c_int, [FP, FP, FP]),
            "ref_demo_rng_create": (VP, [C.c_uint32]),
            "ref_demo_rng_destroy": (None, [VP]),
            "ref_demo_render": (None, [VP, C.c_int, C.c_int, FP, FP, FP, C.c_int, VP]),
            "ref_write_ply": (C.c_int, [CP, VP, C.c_size_t]),
            "ref_parse_intrinsics_file": (C.c_int, [CP, FP, C.POINTER(C.c_int), FP, C.POINTER(C.c_int), CP,
                                                    C.c_size_t]),
            "ref_load_pose": (C.c_int, [CP, FP, FP, CP, C.c_size_t]),
            "ref_quaternion_to_matrix": (None, [C.c_float, C.c_float, C.c_float, C.c_float, FP]),
            "ref_list_frame_ids": (C.c_int, [CP, VP, C.c_size_t, C.POINTER(C.c_size_t), CP, C.c_size_t]),
            "ref_find_intrinsics_path": (C.c_int, [CP, CP, C.c_size_t, CP, C.c_size_t]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype, f.argtypes = r, a
        _demo_ref = L
    return _demo_ref


def p(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def k4(K) -> np.ndarray:
    return np.array([K.fx, K.fy, K.cx, K.cy], dtype=np.float32)


def oracle_params(vp, filter_sigma_spatial=None, filter_sigma_range=None) -> OracleParams:
    """Oracle constants mirroring a libvhm Params (same field values)."""
    o = OracleParams()
    oracle().vor_default_params(C.byref(o))
    for name, _ in OracleParams._fields_:
        if name == "truncation":
            o.truncation = vp.truncation_distance
        elif name == "filter_radius":
            o.filter_radius = 2
        else:
            setattr(o, name, getattr(vp, name))
    return o


class OracleVolume:
    """Restated reference CPU volume (float or double)."""

    def __init__(self, params: OracleParams, fp64: bool = False):
        self.params = params
        self.h = oracle().vor_create(C.byref(params), int(fp64))
        self.fp64 = fp64

    def __del__(self):
        if getattr(self, "h", None):
            oracle().vor_destroy(self.h)
            self.h = None

    def integrate(self, depth_m: np.ndarray, K, pose12: np.ndarray, stride=1, apply_filter=True):
        d = np.ascontiguousarray(depth_m, np.float32)
        oracle().vor_integrate(self.h, p(d), K.width, K.height, p(k4(K)), p(pose12), stride, int(apply_filter))

    def integrate_point_cloud(self, points_c: np.ndarray, pose12: np.ndarray):
        pts = np.ascontiguousarray(points_c, np.float32).reshape(-1, 3)
        oracle().vor_integrate_point_cloud(self.h, p(pts), len(pts), p(np.ascontiguousarray(pose12, np.float32)))

    def prepare(self, depth_m, K, pose12, stride=1) -> np.ndarray:
        d = np.ascontiguousarray(depth_m, np.float32)
        cap = 1 << 22
        out = np.zeros((cap, 3), np.int32)
        n = oracle().vor_prepare(self.h, p(d), K.width, K.height, p(k4(K)), p(pose12), stride, p(out), cap)
        return out[:n].copy()

    def block_count(self):
        return oracle().vor_block_count(self.h)

    def observed_voxel_count(self):
        return oracle().vor_observed_voxel_count(self.h)

    def updates(self):
        return oracle().vor_update_count(self.h)

    def block_keys(self) -> np.ndarray:
        n = self.block_count()
        out = np.zeros((max(n, 1), 3), np.int32)
        oracle().vor_block_keys(self.h, p(out), n)
        return out[:n]

    def dump_blocks(self, keys: np.ndarray) -> np.ndarray:
        keys = np.ascontiguousarray(keys, np.int32)
        from voxelhashmt_amd import VOXEL_DTYPE
        out = np.zeros((len(keys), 512, 6, 3), VOXEL_DTYPE)
        oracle().vor_dump_blocks(self.h, p(keys), len(keys), p(out))
        return out

    def dump_blocks_soa(self, keys: np.ndarray):
        keys = np.ascontiguousarray(keys, np.int32)
        shape = (len(keys), 512, 6, 3)
        w, sc, po = (np.zeros(shape, np.float64) for _ in range(3))
        pn = np.zeros(shape, np.uint32)
        oracle().vor_dump_blocks_soa(self.h, p(keys), len(keys), p(w), p(sc), p(po), p(pn))
        return w, sc, po, pn

    def extract_mesh(self, iso=0.0) -> np.ndarray:
        n = oracle().vor_extract_mesh(self.h, iso, None, 0)
        out = np.zeros((max(n, 1), 3, 3), np.float32)
        oracle().vor_extract_mesh(self.h, iso, p(out), n)
        return out[:n]


class RefVolume:
    """The reference's own CPU volume (VoxelHashTSDF + TsdfIntegrationBackend(kCpu))."""

    def __init__(self, voxel_size: float, truncation: float):
        self.h = ref().vref_volume_create(voxel_size, truncation)

    def __del__(self):
        if getattr(self, "h", None):
            ref().vref_volume_destroy(self.h)
            self.h = None

    def integrate(self, depth_m, K, pose12, stride=1, apply_filter=True):
        d = np.ascontiguousarray(depth_m, np.float32)
        ref().vref_integrate(self.h, p(d), K.width, K.height, p(k4(K)), p(pose12), stride, int(apply_filter))

    def integrate_point_cloud(self, points_c: np.ndarray, pose12: np.ndarray):
        pts = np.ascontiguousarray(points_c, np.float32).reshape(-1, 3)
        ref().vref_integrate_point_cloud(self.h, p(pts), len(pts), p(np.ascontiguousarray(pose12, np.float32)))

    def prepare(self, depth_m, K, pose12, stride=1) -> np.ndarray:
        d = np.ascontiguousarray(depth_m, np.float32)
        cap = 1 << 22
        out = np.zeros((cap, 3), np.int32)
        n = ref().vref_prepare(self.h, p(d), K.width, K.height, p(k4(K)), p(pose12), stride, p(out), cap)
        return out[:n].copy()

    def block_count(self):
        return ref().vref_block_count(self.h)

    def observed_voxel_count(self):
        return ref().vref_observed_voxel_count(self.h)

    def block_keys(self) -> np.ndarray:
        n = self.block_count()
        out = np.zeros((max(n, 1), 3), np.int32)
        ref().vref_block_keys(self.h, p(out), n)
        return out[:n]

    def dump_blocks(self, keys) -> np.ndarray:
        keys = np.ascontiguousarray(keys, np.int32)
        from voxelhashmt_amd import VOXEL_DTYPE
        out = np.zeros((len(keys), 512, 6, 3), VOXEL_DTYPE)
        ref().vref_dump_blocks(self.h, p(keys), len(keys), p(out))
        return out

    def extract_mesh(self, iso=0.0) -> np.ndarray:
        n = ref().vref_extract_mesh(self.h, iso, None, 0)
        out = np.zeros((max(n, 1), 3, 3), np.float32)
        ref().vref_extract_mesh(self.h, iso, p(out), n)
        return out[:n]


def ref_filter(depth_m, K) -> np.ndarray:
    d = np.ascontiguousarray(depth_m, np.float32)
    out = np.zeros_like(d)
    ref().vref_filter(p(d), K.width, K.height, p(k4(K)), p(out))
    return out


def oracle_filter(params: OracleParams, depth_m, K, fp64=False) -> np.ndarray:
    d = np.ascontiguousarray(depth_m, np.float32)
    out = np.zeros_like(d)
    oracle().vor_filter(C.byref(params), int(fp64), p(d), K.width, K.height, p(k4(K)), p(out))
    return out


def sorted_keys(keys: np.ndarray) -> np.ndarray:
    keys = np.asarray(keys, np.int32).reshape(-1, 3)
    return keys[np.lexsort((keys[:, 2], keys[:, 1], keys[:, 0]))]


def mesh_digest(tris: np.ndarray) -> str:
    """sha256 of the canonical (sorted) triangle multiset."""
    import hashlib
    return hashlib.sha256(canonical_mesh(tris).tobytes()).hexdigest()


def canonical_mesh(tris: np.ndarray) -> np.ndarray:
    """Triangles as a sorted multiset (rows of 9 float bit patterns)."""
    t = np.ascontiguousarray(tris, np.float32).reshape(-1, 9).view(np.uint32)
    if len(t) == 0:
        return t
    order = np.lexsort(t.T[::-1])
    return t[order]
