"""integratePointCloud (voxel_hash_tsdf.h:75-97, 952-983).

CPU: the oracle's restatement against the reference's own integratePointCloud
compiled from its sources (oracle/_ref/libvhm_ref.so), bit for bit: block
key sets (a block is created for every sample, even one whose weight is
zero), every voxel layer, observed-voxel counts and the mesh, over clouds
back-projected from depth frames, random clouds with degenerate points, and
successive calls.

GPU (marked gpu): vhm_integrate_point_cloud against the oracle on the same
clouds, bit for bit, including a cloud large enough that many points update
the same voxel (the order of those updates matters and is the reference's).
"""
import numpy as np
import pytest

import oracle_lib as O
import voxelhashmt_amd as vhm
from voxelhashmt_amd import synth


def depth_cloud(depth_m: np.ndarray, K) -> np.ndarray:
    """Camera-space points of the valid pixels (backProject, math_types.h:228-232)."""
    h, w = depth_m.shape
    v, u = np.mgrid[0:h, 0:w]
    d = depth_m.astype(np.float32)
    m = d > 0
    x = (u[m].astype(np.float32) - np.float32(K.cx)) * d[m] / np.float32(K.fx)
    y = (v[m].astype(np.float32) - np.float32(K.cy)) * d[m] / np.float32(K.fy)
    return np.stack([x, y, d[m]], axis=1).astype(np.float32)


def random_cloud(rng, n) -> np.ndarray:
    p = rng.normal(size=(n, 3)).astype(np.float32)
    p[:, 2] = np.abs(p[:, 2]) + 0.3
    p[::97] = 0.0                                  # range <= 1e-6: skipped
    p[1::89] = np.float32(3e-7)                    # just above / below the cut
    p[2::83] = p[3::83][: len(p[2::83])]           # duplicated points
    return p


def cases():
    w = synth.tiny(n_frames=3, seed=5, width=80, height=60)
    frames = synth.render(w.scene, w.K, w.poses, w.noise).astype(np.float32) * np.float32(0.001)
    out = []
    for i in range(2):
        out.append(("frame", depth_cloud(frames[i], w.K), vhm.pose_array(w.poses[i]), w.voxel_size, w.truncation))
    rng = np.random.default_rng(3)
    pose = np.array([0.36, -0.48, 0.8, 0.8, 0.6, 0.0, -0.48, 0.64, 0.6, 0.1, -0.2, 0.3], np.float32)
    out.append(("random", random_cloud(rng, 3000), pose, 0.02, 0.06))
    # dense: many points along nearly the same ray update the same voxels
    base = np.array([0.1, -0.05, 1.2], np.float32)
    dense = base + rng.normal(scale=0.004, size=(4000, 3)).astype(np.float32)
    out.append(("dense", dense, pose, 0.01, 0.04))
    return out


def _ref_params(voxel, trunc):
    return vhm.default_params(voxel, trunc, max_blocks=40000, hash_log2_buckets=18)


@pytest.mark.skipif(not O.ref_available(), reason="reference CPU path not built here")
def test_oracle_point_cloud_matches_reference():
    for name, pts, pose, voxel, trunc in cases():
        ov = O.OracleVolume(O.oracle_params(_ref_params(voxel, trunc)))
        rv = O.RefVolume(voxel, trunc)
        for k in range(2):  # two successive calls with the second pose shifted
            p2 = pose.copy()
            p2[9:] += np.float32(0.01 * k)
            ov.integrate_point_cloud(pts, p2)
            rv.integrate_point_cloud(pts, p2)
        keys = O.sorted_keys(rv.block_keys())
        assert np.array_equal(O.sorted_keys(ov.block_keys()), keys), name
        assert ov.observed_voxel_count() == rv.observed_voxel_count() > 0, name
        assert np.array_equal(ov.dump_blocks(keys).view(np.uint32), rv.dump_blocks(keys).view(np.uint32)), name
        assert np.array_equal(O.canonical_mesh(ov.extract_mesh()), O.canonical_mesh(rv.extract_mesh())), name


@pytest.mark.gpu
def test_gpu_point_cloud_matches_oracle(gpu):
    for name, pts, pose, voxel, trunc in cases():
        print(f"[point cloud] {name}: {len(pts)} points", flush=True)
        vp = _ref_params(voxel, trunc)
        ov = O.OracleVolume(O.oracle_params(vp))
        with vhm.Volume(vp) as vol:
            for k in range(2):
                p2 = pose.copy()
                p2[9:] += np.float32(0.01 * k)
                ov.integrate_point_cloud(pts, p2)
                vol.integrate_point_cloud(pts, vhm.make_pose(p2[:9].reshape(3, 3), p2[9:]))
            keys = O.sorted_keys(ov.block_keys())
            assert np.array_equal(O.sorted_keys(vol.block_keys()), keys), name
            assert vol.observed_voxel_count() == ov.observed_voxel_count(), name
            assert vol.stats()["updates"] == ov.updates(), name
            assert np.array_equal(vol.download_blocks(keys).view(np.uint32), ov.dump_blocks(keys).view(np.uint32)), name
            assert np.array_equal(O.canonical_mesh(vol.extract_mesh()), O.canonical_mesh(ov.extract_mesh())), name


@pytest.mark.gpu
def test_gpu_point_cloud_then_frames_match_oracle(gpu):
    """A cloud followed by depth frames on the same volume (the block pool,
    dirty masks and incremental meshing stay consistent)."""
    w = synth.tiny(n_frames=2, seed=9, width=96, height=72)
    frames = synth.render(w.scene, w.K, w.poses, w.noise)
    vp = vhm.default_params(w.voxel_size, w.truncation, hash_log2_buckets=w.hash_log2_buckets,
                            max_blocks=w.max_blocks, max_width=w.K.width, max_height=w.K.height)
    ov = O.OracleVolume(O.oracle_params(vp))
    cloud = depth_cloud(frames[0].astype(np.float32) * np.float32(0.001), w.K)
    with vhm.Volume(vp) as vol:
        vol.integrate_point_cloud(cloud, w.poses[0])
        ov.integrate_point_cloud(cloud, vhm.pose_array(w.poses[0]))
        vol.integrate(frames[1], w.K, w.poses[1])
        ov.integrate(frames[1].astype(np.float32) * np.float32(0.001), w.K, vhm.pose_array(w.poses[1]))
        keys = O.sorted_keys(ov.block_keys())
        assert np.array_equal(O.sorted_keys(vol.block_keys()), keys)
        assert np.array_equal(vol.download_blocks(keys).view(np.uint32), ov.dump_blocks(keys).view(np.uint32))
        _, _, tris, _ = vol.extract_mesh_incremental()
        assert np.array_equal(O.canonical_mesh(tris), O.canonical_mesh(ov.extract_mesh()))
