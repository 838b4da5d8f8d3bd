"""GPU parity: the HIP path (libvhm.so, through its C-ABI) against the CPU
oracle -- the restated reference CPU pipeline -- and, where it was built,
the reference's own CPU code, on identical seeded inputs.

Bar (from the fusion path's contract): bit-exact filtered depth and block
key sets; voxel layers within 1e-6 relative (they are in fact compared
bit-for-bit and the count of differing words is reported); meshes equal as
multisets with vertices within 1e-5 voxel units.
"""
import numpy as np
import pytest

import oracle_lib as O
import voxelhashmt_amd as vhm
from voxelhashmt_amd import synth

pytestmark = pytest.mark.gpu

REL_TOL = 1e-6  # per-voxel layer tolerance (relative)


def _params(w, **kw):
    base = dict(filter_sigma_spatial=w.filter_sigma_spatial, filter_sigma_range=w.filter_sigma_range,
                hash_log2_buckets=w.hash_log2_buckets, max_blocks=w.max_blocks, max_width=w.K.width,
                max_height=w.K.height)
    base.update(kw)
    return vhm.default_params(w.voxel_size, w.truncation, **base)


def _metres(frames_u16):
    # dataset_io.h:349: static_cast<float>(raw) * depth_scale, raw 0 stays 0
    return frames_u16.astype(np.float32) * np.float32(0.001)


def _layers_close(a, b):
    """Element-wise |a-b| <= 1e-6 |b| on the float fields, equality on normals."""
    ok = True
    for f in ("weight", "surface_coord", "plane_offset"):
        x, y = a[f].astype(np.float64), b[f].astype(np.float64)
        ok &= bool(np.all(np.abs(x - y) <= REL_TOL * np.abs(y) + 1e-30))
    return ok and bool(np.array_equal(a["packed_normal"], b["packed_normal"]))


def _diff_words(a, b):
    return int((a.view(np.uint32) != b.view(np.uint32)).sum())


@pytest.mark.parametrize("name", ["tiny", "vga"])
def test_filter_bitexact(gpu, name):
    w = synth.tiny(n_frames=2, seed=11) if name == "tiny" else synth.plain_small(n_frames=2)
    frames = synth.render(w.scene, w.K, w.poses, w.noise)
    print(f"[filter] {w.name} sigma_s={w.filter_sigma_spatial} sigma_r={w.filter_sigma_range}", flush=True)
    vp = _params(w)
    op = O.oracle_params(vp)
    with vhm.Volume(vp) as vol:
        for i in range(len(frames)):
            g16 = vol.filter_depth(frames[i], w.K)            # uint16 input path
            g32 = vol.filter_depth(_metres(frames[i]), w.K)   # float input path
            o = O.oracle_filter(op, _metres(frames[i]), w.K)
            assert np.array_equal(g16.view(np.uint32), o.view(np.uint32)), "u16 filter differs from oracle"
            assert np.array_equal(g32.view(np.uint32), o.view(np.uint32)), "f32 filter differs from oracle"
            if O.ref_available() and w.filter_sigma_spatial == 1.5:
                r = O.ref_filter(_metres(frames[i]), w.K)
                assert np.array_equal(g32.view(np.uint32), r.view(np.uint32))


@pytest.mark.parametrize("name,stride", [("tiny", 1), ("tiny", 2), ("vga", 1)])
def test_allocation_key_set(gpu, name, stride):
    w = synth.tiny(n_frames=3, seed=5) if name == "tiny" else synth.plain_small(n_frames=2)
    frames = _metres(synth.render(w.scene, w.K, w.poses, w.noise))
    vp = _params(w)
    ov = O.OracleVolume(O.oracle_params(vp))
    print(f"[alloc] {w.name} stride={stride}", flush=True)
    with vhm.Volume(vp) as vol:
        for i in range(len(frames)):
            pose = vhm.pose_array(w.poses[i])
            g = O.sorted_keys(vol.allocate_active_blocks(frames[i], w.K, w.poses[i], stride))
            o = O.sorted_keys(ov.prepare(frames[i], w.K, pose, stride))
            assert len(g) == len(o) and np.array_equal(g, o), f"frame {i}: {len(g)} vs {len(o)} keys"
        # the persistent set is the union of the frames' sets
        assert np.array_equal(O.sorted_keys(vol.block_keys()), O.sorted_keys(ov.block_keys()))


def _run_pipeline(w, n_frames, check_ref):
    frames = synth.render(w.scene, w.K, w.poses[:n_frames], w.noise)
    vp = _params(w)
    ov = O.OracleVolume(O.oracle_params(vp))
    rv = O.RefVolume(w.voxel_size, w.truncation) if check_ref else None
    vol = vhm.Volume(vp)
    for i in range(n_frames):
        vol.integrate(frames[i], w.K, w.poses[i])  # uint16 path, filter on the GPU
        m = _metres(frames[i])
        ov.integrate(m, w.K, vhm.pose_array(w.poses[i]))
        if rv is not None:
            rv.integrate(m, w.K, vhm.pose_array(w.poses[i]))
    vol.synchronize()
    return vol, ov, rv


@pytest.mark.parametrize("seed", [3, 9])
def test_pipeline_tiny(gpu, seed):
    w = synth.tiny(n_frames=4, seed=seed)
    print(f"[pipeline] {w.name} seed={seed}", flush=True)
    vol, ov, rv = _run_pipeline(w, 4, check_ref=O.ref_available())
    keys = O.sorted_keys(ov.block_keys())
    assert np.array_equal(O.sorted_keys(vol.block_keys()), keys)
    assert vol.block_count() == ov.block_count()
    assert vol.observed_voxel_count() == ov.observed_voxel_count()
    assert vol.stats()["updates"] == ov.updates()
    g, o = vol.download_blocks(keys), ov.dump_blocks(keys)
    print(f"[pipeline] blocks={len(keys)} differing words={_diff_words(g, o)}", flush=True)
    assert _layers_close(g, o)
    assert _diff_words(g, o) == 0
    if rv is not None:
        assert _diff_words(g, rv.dump_blocks(keys)) == 0
    gm, om = O.canonical_mesh(vol.extract_mesh()), O.canonical_mesh(ov.extract_mesh())
    print(f"[pipeline] triangles gpu={len(gm)} oracle={len(om)}", flush=True)
    assert len(gm) == len(om)
    assert np.array_equal(gm, om)
    vol.close()


def test_pipeline_qvga_config_sigmas(gpu):
    """160x120 frames with the benchmark filter constants (sigma_s 4.5 px,
    sigma_r 30 mm), 1 cm voxels."""
    w = synth.tiny(n_frames=3, seed=21, width=160, height=120)
    w.filter_sigma_spatial, w.filter_sigma_range, w.voxel_size, w.truncation = 4.5, 0.03, 0.01, 0.04
    w.max_blocks, w.hash_log2_buckets = 60000, 18
    vol, ov, _ = _run_pipeline(w, 3, check_ref=False)
    keys = O.sorted_keys(ov.block_keys())
    assert np.array_equal(O.sorted_keys(vol.block_keys()), keys)
    g, o = vol.download_blocks(keys), ov.dump_blocks(keys)
    print(f"[qvga] blocks={len(keys)} differing words={_diff_words(g, o)}", flush=True)
    assert _layers_close(g, o)
    gm, om = O.canonical_mesh(vol.extract_mesh()), O.canonical_mesh(ov.extract_mesh())
    assert len(gm) == len(om)
    if not np.array_equal(gm, om):  # vertices within 1e-5 voxel units
        a = gm.view(np.float32).reshape(-1, 9)
        b = om.view(np.float32).reshape(-1, 9)
        assert np.abs(a - b).max() <= 1e-5 * w.voxel_size
    vol.close()


def test_fp64_oracle_report(gpu):
    """Reports how far the float pipeline (GPU == float oracle) sits from the
    same algorithm evaluated in double precision."""
    w = synth.tiny(n_frames=3, seed=3)
    frames = synth.render(w.scene, w.K, w.poses, w.noise)
    vp = _params(w)
    d64 = O.OracleVolume(O.oracle_params(vp), fp64=True)
    with vhm.Volume(vp) as vol:
        for i in range(3):
            vol.integrate(frames[i], w.K, w.poses[i])
            d64.integrate(_metres(frames[i]), w.K, vhm.pose_array(w.poses[i]))
        keys = O.sorted_keys(d64.block_keys())
        g = vol.download_blocks(keys)
    w64, sc64, po64, _ = d64.dump_blocks_soa(keys)
    both = (g["weight"] > 0) & (w64 > 0)
    rel_w = np.abs(g["weight"][both] - w64[both]) / np.abs(w64[both])
    rel_po = np.abs(g["plane_offset"][both] - po64[both]) / np.maximum(np.abs(po64[both]), 1e-3)
    print(f"[fp64] slots compared={both.sum()} weight rel err median={np.median(rel_w):.2e} "
          f"p99={np.quantile(rel_w, 0.99):.2e}; plane_offset rel err median={np.median(rel_po):.2e}", flush=True)
    assert np.median(rel_w) < 1e-5


def test_empty_and_invalid_frames(gpu):
    w = synth.tiny(n_frames=1)
    with vhm.Volume(_params(w)) as vol:
        vol.integrate(np.zeros((w.K.height, w.K.width), np.uint16), w.K, w.poses[0])
        assert vol.block_count() == 0 and vol.observed_voxel_count() == 0
        assert len(vol.extract_mesh()) == 0
        big = vhm.make_intrinsics(w.K.width * 4, w.K.height, w.K.fx, w.K.fy, w.K.cx, w.K.cy)
        with pytest.raises(vhm.VhmError) as e:
            vol.integrate(np.zeros((big.height, big.width), np.uint16), big, w.poses[0])
        assert e.value.code == 1


def test_capacity_overflow_is_reported(gpu):
    """A pool too small for the frame reports VHM_ERR_CAPACITY (the analogue of
    CudaActiveBlockAllocationFallback)."""
    w = synth.tiny(n_frames=1)
    frames = synth.render(w.scene, w.K, w.poses, w.noise)
    with vhm.Volume(_params(w, max_blocks=64)) as vol:
        vol.integrate(frames[0], w.K, w.poses[0])
        with pytest.raises(vhm.VhmError) as e:
            vol.synchronize()
        assert e.value.code == 3


def _packed(keys):
    """Packed 21-bit block keys (bias 2^20; x lowest), the order blocks are numbered in."""
    k = np.asarray(keys, np.int64).reshape(-1, 3) + (1 << 20)
    return k[:, 0] | (k[:, 1] << 21) | (k[:, 2] << 42)


def test_determinism(gpu):
    """Two runs over the same frames return the same bytes: block keys in pool
    order, the full extraction and the host-output incremental extraction in
    their output order.  Pool order is creation order: each frame's new
    blocks follow the earlier ones, in packed-key order."""
    w = synth.plain_small(n_frames=3)
    frames = synth.render(w.scene, w.K, w.poses, w.noise)
    outs = []
    for run in range(2):
        print(f"[determinism] run {run}", flush=True)
        with vhm.Volume(_params(w)) as vol:
            counts = []
            for i in range(3):
                vol.integrate(frames[i], w.K, w.poses[i])
                counts.append(vol.block_count())
            pool_keys = vol.block_keys()
            inc = vol.extract_mesh_incremental()
            outs.append((pool_keys, vol.extract_mesh(), inc))
            packed = _packed(pool_keys)
            start = 0
            for n in counts:
                assert np.all(np.diff(packed[start:n]) > 0), "a frame's new blocks are numbered in key order"
                start = n
    assert np.array_equal(outs[0][0], outs[1][0])
    assert outs[0][1].tobytes() == outs[1][1].tobytes()
    for a, b in zip(outs[0][2], outs[1][2]):
        assert np.asarray(a).tobytes() == np.asarray(b).tobytes()


def test_back_to_back_frames_equal_synchronised_frames(gpu):
    """A frame's deferred border pass runs on a second stream beside the next
    frame's filter and allocation.  Frames from device memory integrated back
    to back (nothing waits between them) leave the same volume, bit for bit,
    as the same frames with a synchronisation after each, which orders every
    pass one after another; 1280x720 frames of the plain-at-scale walk."""
    import torch
    w = synth.plain_scale(n_frames=8)
    frames = torch.empty((8, w.K.height, w.K.width), dtype=torch.int16, device="cuda:0")
    synth.render(w.scene, w.K, w.poses, w.noise, 0, frames)
    torch.cuda.synchronize()
    digests, keys = [], None
    for sync in (False, True):
        with vhm.Volume(_params(w)) as vol:
            for i in range(8):
                vol.integrate(frames[i], w.K, w.poses[i])
                if sync:
                    vol.synchronize()
            vol.synchronize()
            k = vol.block_keys()
            keys = k if keys is None else keys
            assert np.array_equal(k, keys)
            digests.append((vol.block_digests(keys), vol.stats()["updates"], vol.observed_voxel_count()))
            print(f"[back-to-back] sync={sync}: {len(k)} blocks, {digests[-1][1]} updates", flush=True)
    assert np.array_equal(digests[0][0], digests[1][0]) and digests[0][1:] == digests[1][1:]


def test_full_extraction_into_device_memory(gpu):
    """vhm_extract_mesh_to with a device buffer writes the bytes the host
    extraction returns, in the same order; a smaller device buffer receives
    their first `capacity` triangles."""
    import torch
    w = synth.plain_small(n_frames=3)
    frames = synth.render(w.scene, w.K, w.poses, w.noise)
    with vhm.Volume(_params(w)) as vol:
        for i in range(3):
            vol.integrate(frames[i], w.K, w.poses[i])
        host = vol.extract_mesh()
        dev = vol.extract_mesh(device=True)
        assert len(host) > 1000
        assert dev.cpu().numpy().tobytes() == host.tobytes()
        cap = len(host) // 3
        out = torch.full((cap, 3, 3), float("nan"), dtype=torch.float32, device="cuda")
        n = vhm.C.c_size_t()
        assert vhm.lib().vhm_extract_mesh_to(vol._h, 0.0, out.data_ptr(), cap, vhm.VHM_DEVICE, vhm.C.byref(n)) == 0
        assert n.value == len(host)
        assert out.cpu().numpy().tobytes() == host[:cap].tobytes()


def test_point_cloud_blocks_numbered_in_key_order(gpu):
    """A point cloud's new blocks take pool indices in packed-key order."""
    w = synth.tiny(n_frames=1, seed=4)
    vp = _params(w)
    rng = np.random.default_rng(5)
    pts = np.stack([rng.uniform(-0.8, 0.8, 4000), rng.uniform(-0.6, 0.6, 4000), rng.uniform(0.5, 1.5, 4000)], 1)
    pose = vhm.make_pose(np.eye(3, dtype=np.float32), np.zeros(3, np.float32))
    with vhm.Volume(vp) as vol:
        vol.integrate_point_cloud(pts.astype(np.float32), pose)
        packed = _packed(vol.block_keys())
        assert len(packed) > 100
        assert np.all(np.diff(packed) > 0)


def test_incremental_mesh_equals_full_extraction(gpu):
    """Keeping one mesh per (block, direction) and replacing the items each
    incremental call re-meshes gives, after every call, the full
    extraction's mesh; a call with nothing written lists nothing."""
    w = synth.tiny(n_frames=6, seed=12)
    frames = synth.render(w.scene, w.K, w.poses, w.noise)
    with vhm.Volume(_params(w)) as vol:
        per_item = {}
        for i in range(6):
            vol.integrate(frames[i], w.K, w.poses[i])
            if i % 2 == 1:
                keys, dirs, tris, items = vol.extract_mesh_incremental()
                m = len(keys)
                n_items = int(sum(bin(int(d)).count("1") for d in dirs))
                print(f"[incremental] frame {i}: {m} blocks, {n_items} (block, direction) items re-meshed, "
                      f"{len(tris)} triangles", flush=True)
                assert m > 0 and n_items < 6 * m  # some directions of listed blocks were not re-meshed
                assert np.all(items // m < 6) and np.all(dirs[items % m] & (1 << (items // m)).astype(np.uint8))
                for j, k in enumerate(keys):
                    for d in range(6):
                        if dirs[j] & (1 << d):
                            per_item[(tuple(k), d)] = tris[items == d * m + j]
                union = np.concatenate([t for t in per_item.values() if len(t)]) if per_item else np.zeros((0, 3, 3))
                assert np.array_equal(O.canonical_mesh(union), O.canonical_mesh(vol.extract_mesh()))
        keys, dirs, tris, items = vol.extract_mesh_incremental()
        assert len(keys) == 0 and len(tris) == 0  # nothing written since the last call


def test_device_renderer_matches_host(gpu):
    """The benchmark and the experiments render on the GPU; the parity tests
    render on the host.  Same scenes, poses and noise: the frames agree
    pixel for pixel except where the Box-Muller draw lands within a rounding
    of a millimetre boundary (device log/cos vs glibc: at most 1 mm, rare)."""
    import torch
    for w in (synth.tiny(n_frames=3, seed=5), synth.plain_small(n_frames=2, seed=2)):
        host = synth.render(w.scene, w.K, w.poses, w.noise)
        dev = torch.empty((len(w.poses), w.K.height, w.K.width), dtype=torch.int16, device="cuda:0")
        synth.render(w.scene, w.K, w.poses, w.noise, 0, dev)
        d = dev.cpu().numpy().view(np.uint16)
        diff = np.abs(d.astype(np.int32) - host.astype(np.int32))
        print(f"[render] {w.name}: {int((diff > 0).sum())} of {diff.size} pixels differ, max {int(diff.max())} mm",
              flush=True)
        assert diff.max() <= 1
        assert (diff > 0).mean() < 1e-4
        clean_h = synth.render(w.scene, w.K, w.poses, None)
        synth.render(w.scene, w.K, w.poses, None, 0, dev)
        assert np.array_equal(dev.cpu().numpy().view(np.uint16), clean_h)  # no noise: exact


def test_incremental_mesh_copy_after_small_buffer(gpu):
    """A host buffer smaller than the incremental mesh gets the first
    triangles; vhm_incremental_mesh_copy then returns all of them, equal to
    what a large enough buffer receives."""
    import ctypes as C
    w = synth.tiny(n_frames=4, seed=14)
    frames = synth.render(w.scene, w.K, w.poses, w.noise)
    L = vhm.lib()
    results = []
    for cap in (16, 1 << 20):
        with vhm.Volume(_params(w)) as vol:
            for i in range(4):
                vol.integrate(frames[i], w.K, w.poses[i])
            nb_cap = vol.block_count()
            keys = np.zeros((nb_cap, 3), np.int32)
            dirs = np.zeros(nb_cap, np.uint8)
            tris = np.zeros((cap, 3, 3), np.float32)
            items = np.zeros(cap, np.uint32)
            nb, nt = C.c_size_t(), C.c_size_t()
            assert L.vhm_extract_mesh_incremental(vol._h, 0.0, keys.ctypes.data, dirs.ctypes.data, nb_cap,
                                                  C.byref(nb), tris.ctypes.data, items.ctypes.data, cap,
                                                  vhm.VHM_HOST, C.byref(nt)) == 0
            n = nt.value
            assert n > 16
            full = np.zeros((n, 3, 3), np.float32)
            full_items = np.zeros(n, np.uint32)
            got = C.c_size_t()
            assert L.vhm_incremental_mesh_copy(vol._h, full.ctypes.data, full_items.ctypes.data, n, C.byref(got)) == 0
            assert got.value == n
            m = min(cap, n)
            assert np.array_equal(tris[:m].view(np.uint32), full[:m].view(np.uint32))
            results.append((full, full_items, keys[:nb.value].copy(), dirs[:nb.value].copy()))
    # the two volumes saw the same frames: same items and triangles (pool order may differ)
    (t0, i0, k0, d0), (t1, i1, k1, d1) = results
    assert np.array_equal(O.canonical_mesh(t0), O.canonical_mesh(t1))
    assert sorted(map(tuple, k0)) == sorted(map(tuple, k1))


def test_block_digests_equal_host_hash(gpu):
    """vhm_block_digests equals the polynomial hash of the bytes
    vhm_download_blocks returns, per block; an absent key hashes to 0."""
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
    from make_golden import block_hash
    w = synth.tiny(n_frames=2, seed=6)
    frames = synth.render(w.scene, w.K, w.poses, w.noise)
    with vhm.Volume(_params(w)) as vol:
        for i in range(2):
            vol.integrate(frames[i], w.K, w.poses[i])
        keys = vol.block_keys()
        probe = np.concatenate([keys, np.array([[900000, 900000, 900000]], np.int32)])
        dig = vol.block_digests(probe)
        assert np.array_equal(dig[:-1], block_hash(vol.download_blocks(keys)))
        blocks = vol.download_blocks(keys)
        nonzero = blocks.view(np.uint32).reshape(len(keys), -1).any(axis=1)
        assert dig[-1] == 0 and np.array_equal(dig[:-1] != 0, nonzero)
        assert len(set(dig[:-1][nonzero].tolist())) == int(nonzero.sum()) > 100
