"""Parity at the configurations' full frame sizes.

* 640x480 plain-small room, 1 cm voxels: the first 2 frames against the
  oracle (keys, every voxel layer bit for bit, counters, mesh).
* 1280x720 plain-at-scale room, 5 mm voxels: frame 0 against the oracle
  (keys, layers, counters), then 40 frames through the GPU only, checked by
  size-independent properties: incremental meshes per (block, direction)
  reassemble the full extraction after every call, a second volume fed the
  same frames and meshed into device buffers (the fused pass) yields the
  same triangles per (block, direction) as the first one's host output
  (the two-pass path), and the two volumes end bit-identical on a block
  sample.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
import voxelhashmt_amd as vhm
from voxelhashmt_amd import synth

pytestmark = pytest.mark.gpu


def _params(w):
    return vhm.default_params(w.voxel_size, w.truncation, filter_sigma_spatial=w.filter_sigma_spatial,
                              filter_sigma_range=w.filter_sigma_range, hash_log2_buckets=w.hash_log2_buckets,
                              max_blocks=w.max_blocks, max_width=w.K.width, max_height=w.K.height)


def _against_oracle(w, n_frames, mesh):
    frames = synth.render(w.scene, w.K, w.poses[:n_frames], w.noise)
    vp = _params(w)
    ov = O.OracleVolume(O.oracle_params(vp))
    with vhm.Volume(vp) as vol:
        for i in range(n_frames):
            print(f"[fullsize] {w.name} frame {i}", flush=True)
            vol.integrate(frames[i], w.K, w.poses[i])
            ov.integrate(frames[i].astype(np.float32) * np.float32(0.001), w.K, vhm.pose_array(w.poses[i]))
        vol.synchronize()
        keys = O.sorted_keys(ov.block_keys())
        assert np.array_equal(O.sorted_keys(vol.block_keys()), keys)
        assert vol.observed_voxel_count() == ov.observed_voxel_count()
        assert vol.stats()["updates"] == ov.updates()
        diff = 0
        for s in range(0, len(keys), 2048):
            g, o = vol.download_blocks(keys[s:s + 2048]), ov.dump_blocks(keys[s:s + 2048])
            diff += int((g.view(np.uint32) != o.view(np.uint32)).sum())
        print(f"[fullsize] {w.name}: {len(keys)} blocks, {vol.stats()['updates']} updates, "
              f"{diff} differing words", flush=True)
        assert diff == 0
        if mesh:
            gm, om = O.canonical_mesh(vol.extract_mesh()), O.canonical_mesh(ov.extract_mesh())
            print(f"[fullsize] {w.name}: {len(gm)} triangles", flush=True)
            assert np.array_equal(gm, om)


def test_plain_small_640x480_first_frames(gpu):
    _against_oracle(synth.plain_small(n_frames=2), 2, mesh=True)


def test_plain_scale_1280x720_first_frame(gpu):
    _against_oracle(synth.plain_scale(n_frames=1), 1, mesh=False)


def _items(keys, tris, items):
    """{(key, direction): triangles of that item in output order}."""
    m = len(keys)
    out = {}
    items = items.astype(np.int64)
    order = np.argsort(items, kind="stable")
    ids, starts = np.unique(items[order], return_index=True)
    ends = np.append(starts[1:], len(order))
    for t, s0, e0 in zip(ids, starts, ends):
        out[(tuple(keys[int(t) % m]), int(t) // m)] = tris[order[s0:e0]]
    return out


def _device_incremental(vol, dev_tris, dev_items):
    """vhm_extract_mesh_incremental into device buffers (host keys and directions)."""
    cap_b = max(vol.block_count(), 1)
    keys = np.zeros((cap_b, 3), np.int32)
    dirs = np.zeros(cap_b, np.uint8)
    nb, nt = C.c_size_t(), C.c_size_t()
    rc = vhm.lib().vhm_extract_mesh_incremental(vol._h, 0.0, keys.ctypes.data, dirs.ctypes.data, cap_b, C.byref(nb),
                                                dev_tris.data_ptr(), dev_items.data_ptr(), dev_tris.numel() // 9,
                                                vhm.VHM_DEVICE, C.byref(nt))
    assert rc == 0, vhm.lib().vhm_last_error(vol._h)
    assert nt.value <= dev_items.numel()
    tris = dev_tris[: 9 * nt.value].cpu().numpy().reshape(-1, 3, 3)
    items = dev_items[: nt.value].cpu().numpy().view(np.uint32)
    return keys[: nb.value], dirs[: nb.value], tris, items


def test_plain_scale_1280x720_properties(gpu):
    import torch
    w = synth.plain_scale(n_frames=40)
    frames = torch.empty((40, w.K.height, w.K.width), dtype=torch.int16, device="cuda:0")
    synth.render(w.scene, w.K, w.poses, w.noise, 0, frames)
    vp = _params(w)
    vp.max_blocks = 200_000
    a, b = vhm.Volume(vp), vhm.Volume(vp)
    cap = 16_000_000
    dev_tris = torch.empty(cap * 9, dtype=torch.float32, device="cuda:0")
    dev_items = torch.empty(cap, dtype=torch.int32, device="cuda:0")
    per_item = {}
    for i in range(40):
        a.integrate(frames[i], w.K, w.poses[i])
        b.integrate(frames[i], w.K, w.poses[i])
        if i % 10 == 9:
            keys, dirs, tris, items = a.extract_mesh_incremental()
            m = len(keys)
            order = np.argsort(items, kind="stable")
            ids, starts = np.unique(items[order], return_index=True)
            ends = np.append(starts[1:], len(order))
            got = {int(t): (s0, e0) for t, s0, e0 in zip(ids, starts, ends)}
            for j in range(m):
                for d in range(6):
                    if dirs[j] & (1 << d):  # re-meshed: replace (possibly by nothing)
                        s0, e0 = got.get(d * m + j, (0, 0))
                        per_item[(tuple(keys[j]), d)] = tris[order[s0:e0]]
            assert set(got) <= {d * m + j for j in range(m) for d in range(6) if dirs[j] & (1 << d)}
            # the other volume through the fused device-output pass: same items, same triangles per item
            kb_, db_, tb_, ib_ = _device_incremental(b, dev_tris, dev_items)
            assert dict(zip(map(tuple, kb_), db_)) == dict(zip(map(tuple, keys), dirs))
            ga, gb = _items(keys, tris, items), _items(kb_, tb_, ib_)
            assert ga.keys() == gb.keys() and len(tb_) == len(tris)
            assert all(np.array_equal(ga[k].view(np.uint32), gb[k].view(np.uint32)) for k in ga)
            if i in (19, 39):
                full = a.extract_mesh()
                union = np.concatenate([t for t in per_item.values() if len(t)])
                print(f"[fullsize] frame {i}: {m} blocks re-meshed, {len(tris)} new triangles, full {len(full)}",
                      flush=True)
                assert len(union) == len(full)
                assert np.array_equal(O.canonical_mesh(union), O.canonical_mesh(full))
    ka, kb = O.sorted_keys(a.block_keys()), O.sorted_keys(b.block_keys())
    assert np.array_equal(ka, kb)
    assert a.stats()["updates"] == b.stats()["updates"]
    assert a.observed_voxel_count() == b.observed_voxel_count()
    sample = ka[np.random.default_rng(0).choice(len(ka), 2000, replace=False)]
    assert np.array_equal(a.download_blocks(sample).view(np.uint32), b.download_blocks(sample).view(np.uint32))
    a.close()
    b.close()
