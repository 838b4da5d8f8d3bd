"""Host-side companions of the fusion path (include/vhm_io.h): the demo
scene, the PLY writer, the dataset loader and the PNG decoder.

Pinned three ways:
* against the reference's own functions compiled from its sources
  (oracle/_ref/libdemo_ref.so: makeLookAtPose, traceScene,
  renderSyntheticDepth, writeMeshAsPly, the dataset_io.h parsers), bit for
  bit and message for message;
* against closed forms (ray-sphere and ray-box distances computed in
  double precision, rotations of known quaternions, orthonormality);
* against the PNG specification through an independent encoder
  (tests/pngenc.py) over every filter type, bit depth, alpha and Adam7.
"""
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
import pngenc
import voxelhashmt_amd as vhm
from voxelhashmt_amd import hostio as H

needs_ref = pytest.mark.skipif(not O.demo_ref_available(), reason="reference demo harness not built here")
F3 = C.c_float * 3
F4 = C.c_float * 4
F9 = C.c_float * 9


def _u32(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


# ------------------------------------------------------------ demo scene ---

@needs_ref
def test_demo_pose_look_at_matches_reference_and_circle():
    R = O.demo_ref()
    target = F3(0.08, -0.02, 0.0)
    for n in (1, 7, 28):
        for i in range(n):
            T = H.demo_pose(i, n)
            a = 2.0 * np.pi * i / max(1, n)
            eye = np.array([1.15 * np.cos(a), 0.18 + 0.10 * np.sin(a / 2), 1.15 * np.sin(a)])
            assert np.allclose(np.array(T.t), eye, atol=2e-6)
            Rr, tr = F9(), F3()
            R.ref_demo_look_at(F3(*T.t), target, Rr, tr)
            assert np.array_equal(_u32(list(T.R)), _u32(list(Rr)))
            M = np.array(T.R, np.float64).reshape(3, 3)
            assert np.allclose(M.T @ M, np.eye(3), atol=1e-6)
            fwd = np.array([0.08, -0.02, 0.0]) - eye
            assert np.allclose(M[:, 2], fwd / np.linalg.norm(fwd), atol=1e-6)


def _rays(rng, n):
    o = rng.uniform(-1.5, 1.5, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    # half of the rays aimed near the objects
    aim = rng.uniform([-0.5, -0.3, -0.3], [0.45, 0.2, 0.2], (n // 2, 3))
    v = aim - o[: n // 2]
    d[: n // 2] = (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)
    d[:8] = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1],
                      [1e-7, 1, 0], [0, 1e-7, -1]], np.float32)
    return o, d


@needs_ref
def test_trace_matches_reference():
    R = O.demo_ref()
    o, d = _rays(np.random.default_rng(5), 4000)
    hits = 0
    for i in range(len(o)):
        ours = H.demo_trace(o[i], d[i])
        t = C.c_float(0.0)
        ref = R.ref_demo_trace(F3(*o[i]), F3(*d[i]), C.byref(t))
        assert (ours is not None) == bool(ref)
        if ref:
            hits += 1
            assert np.float32(ours).view(np.uint32) == np.float32(t.value).view(np.uint32)
    assert hits > 1000


def test_trace_closed_forms():
    c, r = np.array([-0.18, 0.02, 0.0]), 0.34
    rng = np.random.default_rng(9)
    for _ in range(200):  # from outside, aimed at the sphere's centre, clear of the box
        u = rng.normal(size=3)
        u[0] = -abs(u[0]) - 0.5
        u /= np.linalg.norm(u)
        o = c + u * rng.uniform(0.6, 2.0)
        d = (c - o) / np.linalg.norm(c - o)
        t = H.demo_trace(o.astype(np.float32), d.astype(np.float32))
        assert t is not None and abs(t - (np.linalg.norm(o - c) - r)) < 2e-6 * (1 + np.linalg.norm(o - c))
    # straight down onto the box top (y = 0.14) away from the sphere
    for x, z in [(0.35, 0.1), (0.4, -0.15), (0.3, 0.0)]:
        t = H.demo_trace(np.float32([x, 1.0, z]), np.float32([0, -1, 0]))
        assert t is not None and abs(t - (1.0 - 0.14)) < 1e-6
    # parallel to an axis outside the box's slab, away from the sphere: miss
    assert H.demo_trace(np.float32([0.3, 0.5, 1.0]), np.float32([0, 0, -1])) is None
    # origin inside the box: exits through the far face
    t = H.demo_trace(np.float32([0.3, 0.0, 0.0]), np.float32([1, 0, 0]))
    assert t is not None and abs(t - 0.12) < 1e-6


@needs_ref
@pytest.mark.parametrize("w,h", [(128, 96), (37, 23)])
def test_render_sequence_matches_reference(w, h):
    R = O.demo_ref()
    K = H.demo_intrinsics(w, h)
    assert (K.fx, K.fy, K.cx, K.cy) == (np.float32(w * np.float32(1.05)), np.float32(h * np.float32(1.15)),
                                        np.float32((w - 1) * 0.5), np.float32((h - 1) * 0.5))
    ours = H.DemoRenderer(42)
    rng = R.ref_demo_rng_create(42)
    try:
        for i in range(4):
            T = H.demo_pose(i, 4)
            a = ours.render(K, T, i)
            b = np.zeros((h, w), np.float32)
            R.ref_demo_render(rng, w, h, F4(K.fx, K.fy, K.cx, K.cy), F9(*T.R), F3(*T.t), i, b.ctypes.data)
            assert np.array_equal(_u32(a), _u32(b)), f"frame {i}"
            assert 0.2 < (a > 0).mean() < 0.9
    finally:
        R.ref_demo_rng_destroy(rng)


def test_render_noise_free_against_double_precision_ray_cast():
    K = H.demo_intrinsics(64, 48)
    T = H.demo_pose(3, 10)
    depth = H.DemoRenderer(42).render(K, T, 3, noise=False)
    Rm = np.array(T.R, np.float64).reshape(3, 3)
    t = np.array(T.t, np.float64)
    u, v = np.meshgrid(np.arange(64), np.arange(48))
    dc = np.stack([(u - K.cx) / K.fx, (v - K.cy) / K.fy, np.ones_like(u, float)], -1)
    dw = dc @ Rm.T
    dw /= np.linalg.norm(dw, axis=-1, keepdims=True)
    best = np.full(u.shape, np.inf)
    # sphere
    oc = t - np.array([-0.18, 0.02, 0.0])
    b = dw @ oc
    disc = b * b - (oc @ oc - 0.34 ** 2)
    s = np.sqrt(np.maximum(disc, 0))
    ts = np.where(-b - s > 1e-4, -b - s, -b + s)
    best = np.where((disc >= 0) & (ts > 1e-4), np.minimum(best, ts), best)
    # box
    lo, hi = np.array([0.02, -0.26, -0.20]), np.array([0.42, 0.14, 0.18])
    with np.errstate(divide="ignore", invalid="ignore"):
        t0 = (lo - t) / dw
        t1 = (hi - t) / dw
    tn = np.max(np.minimum(t0, t1), -1)
    tf = np.min(np.maximum(t0, t1), -1)
    tn = np.maximum(tn, 0)
    tb = np.where(tn > 1e-4, tn, tf)
    best = np.where((tf >= tn) & (tf > 1e-4), np.minimum(best, tb), best)
    with np.errstate(invalid="ignore"):
        z = np.where(np.isfinite(best), (best[..., None] * dw) @ Rm[:, 2], 0)  # camera z of the hit
    z = np.where((z > 0.1) & (z < 4.0), z, 0)
    assert np.array_equal(depth > 0, z > 0)
    err = np.abs(depth - z)[depth > 0]
    # float32 rounding; grazing rays (discriminant near 0) lose more
    assert err.max() < 5e-5 and np.median(err) < 1e-6


# ------------------------------------------------------------------- PLY ---

def _triangles(rng, n):
    t = rng.normal(size=(n, 9)).astype(np.float32)
    t[::17, 3:6] = t[::17, 0:3]                       # repeated vertex: zero area
    t[5::23, 6:9] = t[5::23, 3:6]
    t[7::31] = 0.0                                    # all zero
    t[11::29, 3:9] = t[11::29, 0:3].repeat(2, 0).reshape(-1, 6) + np.float32(1e-5)  # tiny, |n| ~ 0
    t[13::37, 3:6] = t[13::37, 0:3] + np.float32(1e-4)
    t[13::37, 6:9] = t[13::37, 0:3] + np.float32([1e-4, 0, 1e-4])  # |n| near 1e-8
    return t


@needs_ref
@pytest.mark.parametrize("n", [0, 1, 600])
def test_ply_bytes_match_reference(tmp_path, n):
    tris = _triangles(np.random.default_rng(n), n)
    a, b = tmp_path / "ours.ply", tmp_path / "ref.ply"
    written = H.write_ply(a, tris)
    assert O.demo_ref().ref_write_ply(str(b).encode(), tris.ctypes.data, n) == 1
    assert a.read_bytes() == b.read_bytes()
    if n > 100:
        assert 0 < written < n  # the degenerate ones were dropped


def test_ply_format_and_round_trip(tmp_path):
    tris = _triangles(np.random.default_rng(3), 300)
    path = tmp_path / "m.ply"
    n = H.write_ply(path, tris)
    data = path.read_bytes()
    head = data[: data.index(b"end_header\n") + 11].decode().splitlines()
    assert head[:3] == ["ply", "format binary_little_endian 1.0", "comment generated by tsdf_voxel_hash_demo"]
    assert f"element vertex {3 * n}" in head and f"element face {n}" in head
    assert len(data) == len("\n".join(head)) + 1 + n * (3 * 24 + 13)
    back, normals = H.read_ply(path)
    cr = np.cross(tris[:, 3:6].astype(np.float64) - tris[:, 0:3], tris[:, 6:9].astype(np.float64) - tris[:, 0:3])
    kept = np.linalg.norm(cr, axis=1) > 2e-8
    assert len(back) == n
    # every clearly non-degenerate triangle survives, in order
    assert np.array_equal(back[np.isin(_u32(back).view("V36").ravel(), _u32(tris[kept]).view("V36").ravel())],
                          tris[kept])
    nb = np.cross(back[:, 3:6].astype(np.float64) - back[:, 0:3], back[:, 6:9].astype(np.float64) - back[:, 0:3])
    good = np.linalg.norm(nb, axis=1) > 1e-6
    nb = nb[good] / np.linalg.norm(nb[good], axis=1, keepdims=True)
    assert np.allclose(normals[good], nb, atol=1e-5)


def test_ply_unwritable_path_reports_reference_message(tmp_path):
    bad = tmp_path / "missing_dir" / "x.ply"
    with pytest.raises(vhm.VhmError, match="Failed to write mesh to"):
        H.write_ply(bad, np.zeros((1, 9), np.float32))


# -------------------------------------------------------------- datasets ---

def _write(path, text):
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_text(text)
    return path


@needs_ref
@pytest.mark.parametrize("text", ["525 525.5 319.5 239.25 640 480 0.001", "600 601 320 240 1280 720",
                                  "  575.8\n575.8 319.5\t239.5 640 480 0.0002 trailing words",
                                  "1e3 1e3 2.5e2 1.75e2 500 350 1e-3"])
def test_intrinsics_match_reference(tmp_path, text):
    p = _write(tmp_path / "intrinsic.txt", text)
    K, scale = H.dataset_intrinsics(p)
    k4, wh, s, has = F4(), (C.c_int * 2)(), C.c_float(0), C.c_int(0)
    assert O.demo_ref().ref_parse_intrinsics_file(str(p).encode(), k4, wh, C.byref(s), C.byref(has), None, 0) == 0
    assert has.value == 1
    assert _u32([K.fx, K.fy, K.cx, K.cy]).tolist() == _u32(list(k4)).tolist()
    assert (K.width, K.height) == tuple(wh)
    assert np.float32(scale).view(np.uint32) == np.float32(s.value).view(np.uint32)


@needs_ref
def test_intrinsics_without_resolution_take_the_first_depth_png(tmp_path):
    _write(tmp_path / "intrinsic.txt", "500 500 159.5 119.5 0.00025")
    for fid in (4, 2, 9):
        (tmp_path / str(fid)).mkdir()
    (tmp_path / "2" / "depth.png").write_bytes(pngenc.encode(np.zeros((7, 11), np.uint16)))
    (tmp_path / "4" / "depth.png").write_bytes(pngenc.encode(np.zeros((5, 3), np.uint16)))
    K, scale = H.dataset_intrinsics(tmp_path / "intrinsic.txt")
    assert (K.width, K.height) == (11, 7) and K.fx == 500 and scale == np.float32(0.00025)
    k4, wh, s, has = F4(), (C.c_int * 2)(), C.c_float(0), C.c_int(0)
    assert O.demo_ref().ref_parse_intrinsics_file(str(tmp_path / "intrinsic.txt").encode(), k4, wh, C.byref(s),
                                                  C.byref(has), None, 0) == 0
    assert has.value == 0 and s.value == scale


@needs_ref
def test_intrinsics_from_params_txt_and_error_messages(tmp_path):
    (tmp_path / "3").mkdir()
    _write(tmp_path / "5" / "params.txt", "400 400 79.5 59.5 160 120 0.001")
    K, _ = H.dataset_intrinsics(tmp_path / "intrinsic.txt")
    assert (K.width, K.height, K.fx) == (160, 120, 400)
    out, err = C.create_string_buffer(4096), C.create_string_buffer(4096)
    assert O.demo_ref().ref_find_intrinsics_path(str(tmp_path / "intrinsic.txt").encode(), out, 4096, err, 4096) == 0
    assert out.value.decode() == str(tmp_path / "5" / "params.txt")
    empty = tmp_path / "empty"
    empty.mkdir()
    with pytest.raises(vhm.VhmError) as e:
        H.dataset_intrinsics(empty / "intrinsic.txt")
    assert O.demo_ref().ref_find_intrinsics_path(str(empty / "intrinsic.txt").encode(), out, 4096, err, 4096) == -1
    assert str(e.value).endswith(err.value.decode())
    for bad in ["1 2 3 4", "1 2 3 4 5 6 7 8", "abc"]:
        p = _write(tmp_path / "b" / "intrinsic.txt", bad)
        with pytest.raises(vhm.VhmError) as e:
            H.dataset_intrinsics(p)
        k4, wh, s, has = F4(), (C.c_int * 2)(), C.c_float(0), C.c_int(0)
        assert O.demo_ref().ref_parse_intrinsics_file(str(p).encode(), k4, wh, C.byref(s), C.byref(has), err,
                                                      4096) == -1
        assert str(e.value).endswith(err.value.decode()), (bad, str(e.value), err.value)


@needs_ref
@pytest.mark.parametrize("values", [
    "1 0 0 0 0.5 -0.25 2", "0.7071068 0 0 0.7071068 1 2 3", "2 0.4 -1 0.3 0 0 0", "0 0 0 0 1 1 1",
    "0.9 0.1 0.2 0.3 -1.5 0.25 4.75e-1",
    " ".join(str(x) for x in [0.1, 0.2, 0.97, 1.5, -0.98, 0.05, 0.1, -2.0, 0.1, 0.97, -0.2, 0.25]),
    " ".join(str(x) for x in [1, 0, 0, 0.1, 0, 0, -1, 0.2, 0, 1, 0, 0.3, 0, 0, 0, 1])])
def test_pose_matches_reference(tmp_path, values):
    p = _write(tmp_path / "pose.txt", values)
    T = H.load_pose(p)
    Rr, tr, err = F9(), F3(), C.create_string_buffer(1024)
    assert O.demo_ref().ref_load_pose(str(p).encode(), Rr, tr, err, 1024) == 0
    assert _u32(list(T.R)).tolist() == _u32(list(Rr)).tolist()
    assert _u32(list(T.t)).tolist() == _u32(list(tr)).tolist()


@needs_ref
def test_pose_errors_match_reference(tmp_path):
    err = C.create_string_buffer(1024)
    for name, text in [("eight.txt", "1 2 3 4 5 6 7 8"), ("empty.txt", ""), ("words.txt", "pose here")]:
        p = _write(tmp_path / name, text)
        with pytest.raises(vhm.VhmError) as e:
            H.load_pose(p)
        assert O.demo_ref().ref_load_pose(str(p).encode(), F9(), F3(), err, 1024) == -1
        assert str(e.value).endswith(err.value.decode())
    with pytest.raises(vhm.VhmError) as e:
        H.load_pose(tmp_path / "nope.txt")
    assert O.demo_ref().ref_load_pose(str(tmp_path / "nope.txt").encode(), F9(), F3(), err, 1024) == -1
    assert str(e.value).endswith(err.value.decode())


@needs_ref
def test_quaternion_matrix_matches_reference_and_closed_forms():
    rng = np.random.default_rng(11)
    for q in list(rng.normal(size=(500, 4)).astype(np.float32)) + [np.float32([0, 0, 0, 0]),
                                                                   np.float32([1e-9, 0, 0, 0])]:
        ours = H.quaternion_to_matrix(*q)
        Rr = F9()
        O.demo_ref().ref_quaternion_to_matrix(*[float(x) for x in q], Rr)
        assert np.array_equal(_u32(ours.ravel()), _u32(list(Rr)))
        if np.linalg.norm(q) > 1e-3:
            assert np.allclose(ours.astype(np.float64).T @ ours, np.eye(3), atol=3e-6)
            assert abs(np.linalg.det(ours.astype(np.float64)) - 1) < 3e-6
    s = np.float32(np.sqrt(0.5))
    assert np.allclose(H.quaternion_to_matrix(s, 0, 0, s), [[0, -1, 0], [1, 0, 0], [0, 0, 1]], atol=1e-7)
    assert np.allclose(H.quaternion_to_matrix(0, 1, 0, 0), np.diag([1, -1, -1]), atol=0)
    assert np.array_equal(H.quaternion_to_matrix(0, 0, 0, 0), np.eye(3, dtype=np.float32))


@needs_ref
def test_frame_ids_match_reference(tmp_path):
    for name in ["3", "10", "007", "abc", "1a", "0"]:
        (tmp_path / name).mkdir()
    (tmp_path / "5").write_text("a file, not a frame")
    ids = H.dataset_frame_ids(tmp_path)
    assert ids == [0, 3, 7, 10]
    buf, n, err = (C.c_int * 16)(), C.c_size_t(0), C.create_string_buffer(256)
    assert O.demo_ref().ref_list_frame_ids(str(tmp_path).encode(), buf, 16, C.byref(n), err, 256) == 0
    assert list(buf)[: n.value] == ids


# ------------------------------------------------------------------- PNG ---

def _K(w, h):
    return vhm.make_intrinsics(w, h, 100.0, 100.0, w / 2, h / 2)


@pytest.mark.parametrize("bits,ctype,interlace,shape", [
    (16, 0, False, (9, 13)), (16, 0, True, (13, 7)), (16, 0, True, (1, 1)), (16, 0, True, (3, 2)),
    (8, 0, False, (6, 10)), (8, 0, True, (17, 9)), (4, 0, False, (5, 11)), (2, 0, True, (6, 9)),
    (1, 0, False, (4, 19)), (16, 4, False, (5, 6)), (8, 4, True, (9, 9))])
def test_png_decoder_round_trip(tmp_path, bits, ctype, interlace, shape):
    rng = np.random.default_rng(bits * 7 + shape[0])
    h, w = shape
    gray = rng.integers(0, 1 << bits, size=(h, w))
    samples = gray if ctype == 0 else np.stack([gray, rng.integers(0, 1 << bits, size=(h, w))], -1)
    for f in (0, 1, 2, 3, 4, None):
        p = tmp_path / f"x{f}.png"
        p.write_bytes(pngenc.encode(samples, bits, ctype, interlace, filters=f, chunk_split=7))
        assert H.png_size(p) == (w, h)
        raw, valid = H.load_depth_png_u16(p, _K(w, h), 0.001)
        expect = (gray * (65535 // ((1 << bits) - 1))).astype(np.uint16)
        assert np.array_equal(raw, expect), f"filter {f}"
        assert valid == int((expect > 0).sum())


def test_png_depth_scaling_rule(tmp_path):
    raw = np.array([[0, 1, 1000, 65535], [7, 0, 2, 3]], np.uint16)
    p = tmp_path / "d.png"
    pngenc.write_depth_png(p, raw)
    for scale in (0.001, 0.0002, 1.0, -0.001, 0.0):
        d, valid = H.load_depth_png(p, _K(4, 2), scale)
        expect = raw.astype(np.float32) * np.float32(scale)
        expect[(raw == 0) | (expect <= 0)] = 0
        assert np.array_equal(d, expect) and valid == int((expect > 0).sum())
        _, v16 = H.load_depth_png_u16(p, _K(4, 2), scale)
        assert v16 == valid


def test_png_errors(tmp_path):
    good = pngenc.encode(np.ones((4, 5), np.uint16) * 9)
    p = tmp_path / "a.png"
    p.write_bytes(good)
    with pytest.raises(vhm.VhmError) as e:
        H.load_depth_png(p, _K(6, 4), 0.001)
    assert str(e.value).endswith(f"Depth image resolution does not match dataset intrinsics metadata: {p} "
                                 f"(detail=5x4 vs 6x4)")
    rgb = tmp_path / "rgb.png"
    rgb.write_bytes(pngenc.encode(np.zeros((2, 6), np.uint8), 8, 2))  # IHDR says RGB
    with pytest.raises(vhm.VhmError, match="Failed to convert depth PNG to 16-bit grayscale"):
        H.load_depth_png(rgb, _K(2, 2), 0.001)
    bad = bytearray(good)
    bad[40] ^= 0xFF  # inside IDAT
    p2 = tmp_path / "crc.png"
    p2.write_bytes(bytes(bad))
    with pytest.raises(vhm.VhmError, match="CRC"):
        H.load_depth_png(p2, _K(5, 4), 0.001)
    p3 = tmp_path / "short.png"
    p3.write_bytes(good[:-12])
    with pytest.raises(vhm.VhmError, match="truncated"):
        H.load_depth_png(p3, _K(5, 4), 0.001)
    with pytest.raises(vhm.VhmError, match="Failed to open depth PNG"):
        H.load_depth_png(tmp_path / "none.png", _K(5, 4), 0.001)


def test_dataset_frame_loader(tmp_path):
    raw = np.arange(12, dtype=np.uint16).reshape(3, 4) * 100
    d = tmp_path / "17"
    d.mkdir()
    pngenc.write_depth_png(d / "depth.png", raw)
    _write(d / "pose.txt", "1 0 0 0 0.5 0.25 -1")
    T, depth, valid = H.load_dataset_frame(tmp_path, 17, "pose.txt", _K(4, 3), 0.001)
    assert list(T.t) == [0.5, 0.25, -1.0] and list(T.R) == [1, 0, 0, 0, 1, 0, 0, 0, 1]
    assert valid == 11 and depth[2, 3] == np.float32(1100) * np.float32(0.001)
    with pytest.raises(vhm.VhmError, match="Failed to open metadata file"):
        H.load_dataset_frame(tmp_path, 17, "speckle_pose.txt", _K(4, 3), 0.001)


# ------------------------------------------------------------ demo CLI ---

DEMO = Path(__file__).resolve().parents[1] / "voxelhashmt_amd" / "vhm_demo"


def test_demo_cli_argument_errors_match_reference():
    """parseArgs (main.cpp:83-152) messages; no GPU work is reached."""
    for args, msg in [(["--frames"], "Missing value for argument: --frames"),
                      (["--frame-step", "0"], "Invalid integer for --frame-step"),
                      (["--voxel-size", "abc"], "Invalid float for --voxel-size"),
                      (["--backend", "tpu"], "Invalid value for --backend (expected auto|cpu|gpu)"),
                      (["--bogus", "1"], "Unknown argument: --bogus"),
                      (["--synthetic", "--backend", "CPU"], "integrates on the GPU only")]:
        r = subprocess.run([str(DEMO), *args], capture_output=True, text=True, timeout=60)
        assert r.returncode == 1 and msg in r.stderr, (args, r.stderr)
    r = subprocess.run([str(DEMO), "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "Use --synthetic" in r.stdout


def test_demo_cli_dataset_root_errors(tmp_path):
    r = subprocess.run([str(DEMO), "--dataset-root", str(tmp_path / "nope")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and f'Dataset directory not found: "{tmp_path / "nope"}"' in r.stderr
    (tmp_path / "intrinsic.txt").write_text("500 500 63.5 47.5 128 96 0.001")
    r = subprocess.run([str(DEMO), "--dataset-root", str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and f'No frame directories found under "{tmp_path}"' in r.stderr
    (tmp_path / "4").mkdir()
    r = subprocess.run([str(DEMO), "--dataset-root", str(tmp_path), "--frame-start", "5"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1 and "No frames selected with the current start/end/step options." in r.stderr
