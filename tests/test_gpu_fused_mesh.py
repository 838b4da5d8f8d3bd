"""Incremental marching cubes into device buffers (the fused count + emit
pass, one atomic output range per item) against the same call with host
output (the two-pass count, scan, emit path), on two volumes fed the same
frames.

The device output has its items in completion order, so the outputs are
compared per (block key, direction) item: the same items,
and for each the same triangles in the same order (cube order, then cluster
order).  Within one output each item's triangles must be one contiguous run;
the host output has its items in item order, the device output in the order
the items finished.  With a device buffer smaller than the output, only
whole items are written, an item that does not fit is marked invalid and is
listed again by the next call.  After point-cloud integration
(whose blocks carry both near-iso sign hints, so the fused pass's item
filter keeps them) and frames on top, the same per-item equality holds.
"""
import ctypes as C

import numpy as np
import pytest

import voxelhashmt_amd as vhm
from meshitems import per_item as _per_item
from voxelhashmt_amd import synth

pytestmark = pytest.mark.gpu


def _volumes(w):
    vp = vhm.default_params(w.voxel_size, w.truncation, hash_log2_buckets=w.hash_log2_buckets,
                            max_blocks=w.max_blocks, max_width=w.K.width, max_height=w.K.height)
    return vhm.Volume(vp), vhm.Volume(vp)


def _device_call(vol, dev_tris, dev_items):
    """vhm_extract_mesh_incremental with host keys / directions and device triangles / items."""
    cap_b = max(vol.block_count(), 1)
    keys = np.zeros((cap_b, 3), np.int32)
    dirs = np.zeros(cap_b, np.uint8)
    nb, nt = C.c_size_t(), C.c_size_t()
    rc = vhm.lib().vhm_extract_mesh_incremental(vol._h, 0.0, keys.ctypes.data, dirs.ctypes.data, cap_b, C.byref(nb),
                                                dev_tris.data_ptr(), dev_items.data_ptr(), dev_tris.numel() // 9,
                                                vhm.VHM_DEVICE, C.byref(nt))
    assert rc == 0, vhm.lib().vhm_last_error(vol._h)
    return keys[: nb.value], dirs[: nb.value], nb.value, nt.value


@pytest.mark.parametrize("workload", ["tiny", "plain_small"])
def test_fused_device_output_equals_two_pass_host_output(gpu, workload):
    import torch
    w = synth.tiny(n_frames=6, seed=11) if workload == "tiny" else synth.plain_small(n_frames=6)
    frames = synth.render(w.scene, w.K, w.poses, w.noise)
    a, b = _volumes(w)
    cap = 3_000_000
    dev_tris = torch.empty(cap * 9, dtype=torch.float32, device="cuda:0")
    dev_items = torch.empty(cap, dtype=torch.int32, device="cuda:0")
    for i in range(len(frames)):
        a.integrate(frames[i], w.K, w.poses[i])
        b.integrate(frames[i], w.K, w.poses[i])
        if i % 2 == 1:
            ka, da, m, n = _device_call(a, dev_tris, dev_items)
            kb, db, tris, items = b.extract_mesh_incremental()
            print(f"[fused] {workload} frame {i}: {m} blocks, {n} triangles", flush=True)
            assert (m, n) == (len(kb), len(tris)) and n > 0
            got = dev_tris[: 9 * n].cpu().numpy().reshape(-1, 3, 3)
            got_items = dev_items[:n].cpu().numpy().view(np.uint32)
            ga, gb = _per_item(ka, got, got_items), _per_item(kb, tris, items, in_order=True)
            assert ga.keys() == gb.keys()
            for k in ga:
                assert np.array_equal(ga[k].view(np.uint32), gb[k].view(np.uint32)), k
            # the re-meshed (block, direction) masks agree too
            assert dict(zip(map(tuple, ka), da)) == dict(zip(map(tuple, kb), db))
    a.close()
    b.close()


def test_fused_device_output_past_capacity(gpu):
    """An item whose run would cross the buffer's capacity writes no
    triangle: its entries below capacity carry item 0xFFFFFFFF and NaN
    coordinates, nothing past capacity is written, every other item is
    complete, and the next call (nothing integrated in between) lists the
    dropped items again and returns their triangles."""
    import torch
    w = synth.tiny(n_frames=3, seed=4)
    frames = synth.render(w.scene, w.K, w.poses, w.noise)
    a, b = _volumes(w)
    for i in range(len(frames)):
        a.integrate(frames[i], w.K, w.poses[i])
        b.integrate(frames[i], w.K, w.poses[i])
    kb, db, tris, items = b.extract_mesh_incremental()
    full = _per_item(kb, tris, items, in_order=True)
    cap = len(tris) // 3
    dev_tris = torch.full((cap * 9 + 9,), -7.0, dtype=torch.float32, device="cuda:0")
    dev_items = torch.full((cap,), -2, dtype=torch.int32, device="cuda:0")
    ka, da, m, n = _device_call(a, dev_tris[: cap * 9], dev_items)
    assert (m, n) == (len(kb), len(tris)) and n > cap
    got = dev_tris[: 9 * cap].cpu().numpy().reshape(-1, 3, 3)
    got_items = dev_items.cpu().numpy().view(np.uint32)
    assert not np.any(got_items == 0xFFFFFFFE), "every entry below capacity belongs to an item"
    dropped = got_items == 0xFFFFFFFF
    assert dropped.any() and np.isnan(got[dropped]).all()
    assert np.all(dev_tris[cap * 9:].cpu().numpy() == -7.0)
    part = _per_item(ka, got[~dropped], got_items[~dropped])
    for k, t in part.items():  # a written item is complete
        assert np.array_equal(t.view(np.uint32), full[k].view(np.uint32)), k
    missing = set(full) - set(part)
    assert missing
    big_tris = torch.empty(len(tris) * 9, dtype=torch.float32, device="cuda:0")
    big_items = torch.empty(len(tris), dtype=torch.int32, device="cuda:0")
    ka2, da2, m2, n2 = _device_call(a, big_tris, big_items)
    again = _per_item(ka2, big_tris[: 9 * n2].cpu().numpy().reshape(-1, 3, 3), big_items[:n2].cpu().numpy().view(np.uint32))
    print(f"[fused] past capacity: {len(part)} items written, {len(missing)} dropped, {len(again)} re-listed",
          flush=True)
    assert missing <= set(again)
    for k, t in again.items():
        assert np.array_equal(t.view(np.uint32), full[k].view(np.uint32)), k
    a.close()
    b.close()


def test_fused_after_point_cloud_and_frames(gpu):
    """Blocks written by point-cloud integration carry both near-iso sign
    hints (the fused pass's item filter must never drop them), then frames
    update them through the integrator: the fused device output still equals
    the two-pass host output per item."""
    import torch
    import test_point_cloud as tpc
    w = synth.tiny(n_frames=3, seed=9, width=96, height=72)
    frames = synth.render(w.scene, w.K, w.poses, w.noise)
    a, b = _volumes(w)
    cloud = tpc.depth_cloud(frames[0].astype(np.float32) * np.float32(0.001), w.K)
    dev_tris = torch.empty(1_000_000 * 9, dtype=torch.float32, device="cuda:0")
    dev_items = torch.empty(1_000_000, dtype=torch.int32, device="cuda:0")
    for v in (a, b):
        v.integrate_point_cloud(cloud, w.poses[0])
    for step in range(2):
        if step == 1:
            for v in (a, b):
                v.integrate(frames[1], w.K, w.poses[1])
                v.integrate(frames[2], w.K, w.poses[2])
        ka, da, m, n = _device_call(a, dev_tris, dev_items)
        kb, db, tris, items = b.extract_mesh_incremental()
        print(f"[fused] point cloud step {step}: {m} blocks, {n} triangles", flush=True)
        assert (m, n) == (len(kb), len(tris)) and n > 0
        got = dev_tris[: 9 * n].cpu().numpy().reshape(-1, 3, 3)
        ga = _per_item(ka, got, dev_items[:n].cpu().numpy().view(np.uint32))
        gb = _per_item(kb, tris, items, in_order=True)
        assert ga.keys() == gb.keys()
        for k in ga:
            assert np.array_equal(ga[k].view(np.uint32), gb[k].view(np.uint32)), k
    a.close()
    b.close()
