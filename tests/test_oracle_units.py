"""Pins of the CPU oracle's single functions by closed forms, invariants and
brute force (independent of the restated code), and -- where the reference's
own CPU build is present -- by direct comparison with it."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O

ref_only = pytest.mark.skipif(not O.ref_available(), reason="reference CPU build (oracle/_ref) not present")
RNG = np.random.default_rng(1234)


def _pack(n):
    a = np.asarray(n, np.float32)
    return O.oracle().vor_pack_normal(a.ctypes.data)


def _unpack(p):
    out = np.zeros(3, np.float32)
    O.oracle().vor_unpack_normal(p, out.ctypes.data)
    return out


def test_octahedral_known_codes():
    # lround(0.5 * 65535) = 32768 (half away from zero); axes map to corners/centre
    assert _pack([0, 0, 1]) == 32768 | (32768 << 16)
    assert _pack([1, 0, 0]) == 65535 | (32768 << 16)
    assert _pack([0, -1, 0]) == 32768 | (0 << 16)
    assert _pack([0, 0, -1]) == 65535 | (65535 << 16)
    assert _pack([0, 0, 0]) == 0
    assert np.allclose(_unpack(32768 | (32768 << 16)), [0, 0, 1], atol=1e-4)


def test_octahedral_round_trip():
    v = RNG.normal(size=(2000, 3)).astype(np.float32)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    err = max(np.abs(_unpack(_pack(n)) - n).max() for n in v)
    assert err < 1e-4  # 16-bit octahedral quantisation: ~2/65535 per axis before renormalisation


@ref_only
def test_octahedral_equals_reference():
    v = RNG.normal(size=(3000, 3)).astype(np.float32)
    for n in v:
        assert O.oracle().vor_pack_normal(n.ctypes.data) == O.ref().vref_pack_normal(n.ctypes.data)
    for p in RNG.integers(0, 2**32, size=3000, dtype=np.uint64):
        a, b = np.zeros(3, np.float32), np.zeros(3, np.float32)
        O.oracle().vor_unpack_normal(int(p), a.ctypes.data)
        O.ref().vref_unpack_normal(int(p), b.ctypes.data)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _weight_closed_form(d, sd, trunc, vaw):
    if d <= 0 or trunc <= 1e-6:
        return 0.0
    dw = 1.0 / (1.0 + (d / 1.5) ** 2)
    r = abs(sd) / trunc
    robust = 0.35 / r if r > 0.35 else 1.0
    band = 1.0 - 0.5 * min(max(r, 0.0), 1.0)
    return max(0.05, dw * robust * band) * min(max(vaw, 0.15), 1.0)


def test_observation_weight_closed_form():
    for _ in range(3000):
        d, sd, tr, vaw = RNG.uniform(0.1, 6), RNG.uniform(-0.05, 0.2), RNG.choice([0.02, 0.04]), RNG.uniform(0, 1)
        got = O.oracle().vor_observation_weight(d, sd, tr, vaw)
        want = _weight_closed_form(np.float32(d), np.float32(sd), np.float32(tr), np.float32(vaw))
        assert abs(got - want) <= 2e-6 * want
    assert O.oracle().vor_observation_weight(0.0, 0.0, 0.04, 1.0) == 0.0


@ref_only
def test_observation_weight_equals_reference():
    for _ in range(3000):
        a = RNG.uniform([0.1, -0.05, 0.01, 0.0], [6, 0.2, 0.05, 1.0]).astype(np.float32)
        assert O.oracle().vor_observation_weight(*map(float, a)) == O.ref().vref_observation_weight(*map(float, a))


def test_dominant_direction_ties_and_signs():
    cases = {(1, 1, 0): 0, (0, 1, 1): 2, (-1, -1, -1): 1, (0, -2, 1): 3, (0.1, 0.2, -0.3): 5, (1e-7, 0, 0): -1,
             (0, 0, 0.5): 4}
    for n, d in cases.items():
        assert O.oracle().vor_dominant_direction(np.array(n, np.float32).ctypes.data) == d


@ref_only
def test_complete_cube_all_masks_equal_reference():
    for mask in range(256):
        for _ in range(4):
            v = RNG.uniform(-1, 1, 8).astype(np.float32)
            ok = np.array([(mask >> c) & 1 for c in range(8)], np.uint8)
            a, b = v.copy(), v.copy()
            oa, ob = ok.copy(), ok.copy()
            ra = O.oracle().vor_complete_cube(a.ctypes.data, oa.ctypes.data)
            rb = O.ref().vref_complete_cube(b.ctypes.data, ob.ctypes.data)
            assert ra == rb and ra == (bin(mask).count("1") >= 6)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)) and np.array_equal(oa, ob)


def test_complete_cube_closed_form_single_gap():
    # one missing corner: value = mean of its three valid edge neighbours
    nbrs = {0: (1, 3, 4), 6: (5, 7, 2)}
    for c, nb in nbrs.items():
        v = RNG.uniform(-1, 1, 8).astype(np.float32)
        ok = np.ones(8, np.uint8)
        ok[c] = 0
        assert O.oracle().vor_complete_cube(v.ctypes.data, ok.ctypes.data) == 1
        want = (np.float32(v[nb[0]]) + np.float32(v[nb[1]]) + np.float32(v[nb[2]])) / np.float32(3)
        assert abs(v[c] - want) <= 1e-6


CORNERS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], np.float32)
EDGES = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]


def test_polygonise_vertices_on_cut_edges():
    """Every emitted vertex is the linear zero crossing of a cut edge."""
    for _ in range(2000):
        vals = RNG.uniform(-1, 1, 8).astype(np.float32)
        tris = np.zeros(45, np.float32)
        n = O.oracle().vor_polygonise(CORNERS.ctypes.data, vals.ctypes.data, 0.0, tris.ctypes.data)
        case = sum(1 << c for c in range(8) if vals[c] < 0)
        cut = [e for e in EDGES if (vals[e[0]] < 0) != (vals[e[1]] < 0)]
        assert (n == 0) == (len(cut) == 0)
        for v in tris[: 9 * n].reshape(-1, 3):
            best = min(np.abs(v - (CORNERS[a] + (CORNERS[b] - CORNERS[a]) * (-vals[a] / (vals[b] - vals[a])))).max()
                       for a, b in cut)
            assert best < 1e-5, case


@ref_only
def test_polygonise_equals_reference():
    for _ in range(2000):
        vals = RNG.uniform(-1, 1, 8).astype(np.float32)
        pos = (CORNERS * 0.01 + RNG.uniform(-2, 2, 3)).astype(np.float32)
        a, b = np.zeros(45, np.float32), np.zeros(45, np.float32)
        na = O.oracle().vor_polygonise(pos.ctypes.data, vals.ctypes.data, 0.0, a.ctypes.data)
        nb = O.ref().vref_polygonise(pos.ctypes.data, vals.ctypes.data, 0.0, b.ctypes.data)
        assert na == nb and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _intr(w, h, f):
    import voxelhashmt_amd as vhm
    return vhm.make_intrinsics(w, h, f, f, (w - 1) / 2, (h - 1) / 2)


def test_filter_constant_depth_is_fixed_point():
    K = _intr(40, 30, 35.0)
    d = np.full((30, 40), 1.25, np.float32)
    op = O.OracleParams()
    O.oracle().vor_default_params(C.byref(op))
    out = O.oracle_filter(op, d, K)
    assert np.abs(out - d).max() <= 2 * np.spacing(np.float32(1.25))
    d[5, 7] = 0.0  # holes stay holes
    assert O.oracle_filter(op, d, K)[5, 7] == 0.0


def _allocation_brute_force(depth, K, pose12, voxel, trunc):
    """Independent numpy statement of the allocation rule: every valid pixel's
    surface point and +-trunc points along the camera ray, 27-neighbourhood of
    their blocks (all in float32, as the reference evaluates it)."""
    f32 = np.float32
    R, t = pose12[:9].reshape(3, 3).astype(f32), pose12[9:].astype(f32)
    v, u = np.nonzero(depth > 0)
    d = depth[v, u].astype(f32)
    pc = np.stack([(u.astype(f32) - f32(K.cx)) * d / f32(K.fx), (v.astype(f32) - f32(K.cy)) * d / f32(K.fy), d], 1)
    pw = np.stack([(R[i, 0] * pc[:, 0] + R[i, 1] * pc[:, 1]) + R[i, 2] * pc[:, 2] for i in range(3)], 1) + t
    r = pw - t
    n = np.sqrt((r[:, 0] * r[:, 0] + r[:, 1] * r[:, 1]) + r[:, 2] * r[:, 2]).astype(f32)
    ray = r / n[:, None]
    keys = set()
    for p in (pw, pw - ray * f32(trunc), pw + ray * f32(trunc)):
        vox = np.floor(p / f32(voxel)).astype(np.int64)
        blk = np.floor_divide(vox, 8)
        for b in {tuple(x) for x in blk.tolist()}:
            for dz in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        keys.add((b[0] + dx, b[1] + dy, b[2] + dz))
    return O.sorted_keys(np.array(sorted(keys), np.int32))


def test_allocation_matches_brute_force():
    import voxelhashmt_amd as vhm
    from voxelhashmt_amd import synth
    w = synth.tiny(n_frames=2, seed=13)
    frames = synth.render(w.scene, w.K, w.poses, w.noise).astype(np.float32) * np.float32(0.001)
    op = O.oracle_params(vhm.default_params(w.voxel_size, w.truncation))
    ov = O.OracleVolume(op)
    for i in range(2):
        pose = vhm.pose_array(w.poses[i])
        got = O.sorted_keys(ov.prepare(frames[i], w.K, pose))
        assert np.array_equal(got, _allocation_brute_force(frames[i], w.K, pose, w.voxel_size, w.truncation))


def test_single_fronto_parallel_observation():
    """One exact plane at z = 1 m seen head-on: every updated voxel stores the
    plane in slot 0 of direction -z with surface coordinate -1, plane_offset
    n.p = -1 (where a normal is estimated) and a weight within the closed-form
    bounds."""
    import voxelhashmt_amd as vhm
    K = _intr(64, 48, 60.0)
    depth = np.ones((48, 64), np.float32)
    op = O.OracleParams()
    O.oracle().vor_default_params(C.byref(op))
    op.voxel_size, op.truncation = 0.02, 0.06
    ov = O.OracleVolume(op)
    pose = vhm.pose_array(vhm.make_pose(np.eye(3), [0, 0, 0]))
    ov.integrate(depth, K, pose, apply_filter=False)
    keys = ov.block_keys()
    w, sc, po, pn = ov.dump_blocks_soa(keys)
    upd = w > 0
    assert upd[..., :5, :].sum() == 0 and upd[..., 5, 1:].sum() == 0  # only direction 5 (-z), slot 0
    assert upd.sum() == ov.updates() > 0
    assert np.allclose(sc[upd], -1.0, atol=2e-6)  # -p.z of a point on z = 1
    # n.p = -1 where the normal is estimated; near the image border the normal
    # falls back to minus the view ray, giving n.p = -|p| < -1
    assert (po[upd] <= -1.0 + 2e-6).all() and np.mean(np.abs(po[upd] + 1.0) < 2e-6) > 0.6
    # weights: closed form with view-angle weight |n.ray| evaluated per voxel, bounded by the head-on case
    assert (w[upd] <= _weight_closed_form(1.0, 0.0, 0.06, 1.0) + 1e-6).all()
    assert (w[upd] >= 0.05 * 0.15 - 1e-7).all()
