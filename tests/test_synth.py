"""Synthetic depth generator: Philox4x32-10 known-answer vectors (Random123's
published KAT set) and the ray caster against closed-form intersections."""
import ctypes as C
import math

import numpy as np

import voxelhashmt_amd as vhm
from voxelhashmt_amd import synth


def _philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    vhm.lib().vhm_philox4x32_10(c, k, o)
    return list(o)


def test_philox_known_answers():
    assert _philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert _philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert _philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def _K():
    return vhm.make_intrinsics(64, 48, 50.0, 50.0, 31.5, 23.5)


def test_room_wall_depth_closed_form():
    """Camera at the origin looking down +z at a wall z = 2: depth is 2 m at
    every pixel (the ray parameter equals camera z)."""
    sc = synth.SceneSpec((-50.0, -50.0, -50.0), (50.0, 50.0, 2.0), [])
    pose = vhm.make_pose(np.eye(3), [0, 0, 0])
    f = synth.render(sc, _K(), [pose], None)[0]
    assert (f == 2000).all()


def test_sphere_depth_closed_form():
    sc = synth.SceneSpec((-50.0, -50.0, -50.0), (50.0, 50.0, 10.0), [(0, (0.0, 0.0, 3.0), 0.5, (0, 0, 0))])
    pose = vhm.make_pose(np.eye(3), [0, 0, 0])
    K = _K()
    f = synth.render(sc, K, [pose], None)[0]
    for (v, u) in [(23, 31), (20, 30), (25, 34)]:
        d = np.array([(u - K.cx) / K.fx, (v - K.cy) / K.fy, 1.0])
        a, b, c = d @ d, -2 * d[2] * 3.0, 9.0 - 0.25
        s = (-b - math.sqrt(b * b - 4 * a * c)) / (2 * a)
        assert abs(int(f[v, u]) - round(s * 1000)) <= 0
    assert f[0, 0] == 10000  # misses the sphere: back wall


def test_box_depth_closed_form():
    sc = synth.SceneSpec((-50.0, -50.0, -50.0), (50.0, 50.0, 10.0), [(1, (0.0, 0.0, 4.0), 0.0, (0.3, 0.3, 0.5))])
    pose = vhm.make_pose(np.eye(3), [0, 0, 0])
    f = synth.render(sc, _K(), [pose], None)[0]
    assert f[23, 31] == 3500  # front face z = 3.5


def test_noise_statistics():
    """sigma(z) = 0.0012 + 0.0019 (z - 0.4)^2 and ~1 % dropouts."""
    sc = synth.SceneSpec((-50.0, -50.0, -50.0), (50.0, 50.0, 3.0), [])
    K = vhm.make_intrinsics(320, 240, 300.0, 300.0, 159.5, 119.5)
    poses = [vhm.make_pose(np.eye(3), [0, 0, 0])] * 4
    f = synth.render(sc, K, poses, synth.sensor_noise(5)).astype(np.float64)
    valid = f > 0
    drop = 1 - valid.mean()
    assert 0.008 < drop < 0.012
    err = (f[valid] - 3000.0) / 1000.0
    sigma = 0.0012 + 0.0019 * (3.0 - 0.4) ** 2
    assert abs(err.std() - sigma) < 0.03 * sigma and abs(err.mean()) < 0.05 * sigma
