"""vhm_extract_mesh_incremental_noalloc (fixed device scratch, no
allocation) against vhm_extract_mesh_incremental with device output, on two
volumes fed the same frames: the same listed blocks and direction masks,
the same triangle count, and per (block, direction) item the same triangles
in the same order.  Two calls in a row, the second with a smaller input (one
more frame: fewer dirty blocks), so the second call runs on scratch holding
the first call's contents.  A list longer than the scratch's limit is
refused with VHM_ERR_CAPACITY and its blocks stay dirty: the next call of
the allocating entry point lists them, with the triangles of a volume that
was never refused.
"""
import ctypes as C

import numpy as np
import pytest

import voxelhashmt_amd as vhm
from meshitems import per_item
from voxelhashmt_amd import synth

pytestmark = pytest.mark.gpu

NOALLOC_MAX_LISTED = 131072
VHM_ERR_CAPACITY = 3


def _call(fn, vol, dev_tris, dev_items, with_kind):
    cap_b = max(vol.block_count(), 1)
    keys = np.zeros((cap_b, 3), np.int32)
    dirs = np.zeros(cap_b, np.uint8)
    nb, nt = C.c_size_t(), C.c_size_t()
    args = [vol._h, 0.0, keys.ctypes.data, dirs.ctypes.data, cap_b, C.byref(nb), dev_tris.data_ptr(),
            dev_items.data_ptr(), dev_tris.numel() // 9]
    rc = fn(*(args + ([vhm.VHM_DEVICE] if with_kind else []) + [C.byref(nt)]))
    return rc, keys[: nb.value], dirs[: nb.value], nt.value


def _lib():
    L = vhm.lib()
    f = L.vhm_extract_mesh_incremental_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_float, C.c_void_p, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t), C.c_void_p,
                  C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]
    return L


def _items(keys, dev_tris, dev_items, n):
    tris = dev_tris[: 9 * n].cpu().numpy().reshape(-1, 3, 3)
    return per_item(keys, tris, dev_items[:n].cpu().numpy().view(np.uint32))


def test_noalloc_equals_allocating_entry_point(gpu):
    import torch
    L = _lib()
    w = synth.plain_small(n_frames=4)
    frames = synth.render(w.scene, w.K, w.poses, w.noise)
    vp = vhm.default_params(w.voxel_size, w.truncation, filter_sigma_spatial=w.filter_sigma_spatial,
                            filter_sigma_range=w.filter_sigma_range, hash_log2_buckets=w.hash_log2_buckets,
                            max_blocks=w.max_blocks, max_width=w.K.width, max_height=w.K.height)
    cap = 8_000_000
    ta, ia = (torch.empty(cap * 9, dtype=torch.float32, device="cuda:0"),
              torch.empty(cap, dtype=torch.int32, device="cuda:0"))
    tb, ib = (torch.empty(cap * 9, dtype=torch.float32, device="cuda:0"),
              torch.empty(cap, dtype=torch.int32, device="cuda:0"))
    sizes = []
    with vhm.Volume(vp) as a, vhm.Volume(vp) as b:
        for call, span in enumerate([(0, 3), (3, 4)]):
            for i in range(*span):
                a.integrate(frames[i], w.K, w.poses[i])
                b.integrate(frames[i], w.K, w.poses[i])
            rca, ka, da, na = _call(L.vhm_extract_mesh_incremental, a, ta, ia, True)
            rcb, kb, db, nb = _call(L.vhm_extract_mesh_incremental_noalloc, b, tb, ib, False)
            print(f"[noalloc] call {call}: {len(kb)} blocks, {nb} triangles", flush=True)
            assert rca == 0 and rcb == 0, L.vhm_last_error(b._h)
            assert np.array_equal(ka, kb) and np.array_equal(da, db) and na == nb and 0 < nb <= cap
            ga, gb = _items(ka, ta, ia, na), _items(kb, tb, ib, nb)
            assert ga.keys() == gb.keys()
            for k in ga:
                assert np.array_equal(ga[k].view(np.uint32), gb[k].view(np.uint32)), (call, k)
            sizes.append(len(kb))
    assert sizes[1] < sizes[0]


def test_noalloc_refuses_a_list_over_its_limit(gpu):
    import torch
    L = _lib()
    w = synth.plain_small(n_frames=1)
    frames = synth.render(w.scene, w.K, w.poses, w.noise)
    vp = vhm.default_params(0.0025, 0.01, filter_sigma_spatial=w.filter_sigma_spatial,
                            filter_sigma_range=w.filter_sigma_range, hash_log2_buckets=21, max_blocks=300_000,
                            max_width=w.K.width, max_height=w.K.height)
    cap = 30_000_000
    ta, ia = (torch.empty(cap * 9, dtype=torch.float32, device="cuda:0"),
              torch.empty(cap, dtype=torch.int32, device="cuda:0"))
    with vhm.Volume(vp) as a, vhm.Volume(vp) as b:
        a.integrate(frames[0], w.K, w.poses[0])
        b.integrate(frames[0], w.K, w.poses[0])
        rc, _, _, _ = _call(L.vhm_extract_mesh_incremental_noalloc, b, ta, ia, False)
        msg = L.vhm_last_error(b._h).decode()
        print(f"[noalloc] {b.block_count()} blocks: rc {rc}: {msg}", flush=True)
        assert rc == VHM_ERR_CAPACITY and str(NOALLOC_MAX_LISTED) in msg
        rca, ka, _, na = _call(L.vhm_extract_mesh_incremental, a, ta, ia, True)
        want = _items(ka, ta, ia, na)
        rcb, kb, _, nb = _call(L.vhm_extract_mesh_incremental, b, ta, ia, True)
        got = _items(kb, ta, ia, nb)
        assert rca == 0 and rcb == 0 and len(ka) > NOALLOC_MAX_LISTED
        assert set(map(tuple, ka)) <= set(map(tuple, kb))
        assert want.keys() <= got.keys()
        for k in want:
            assert np.array_equal(want[k].view(np.uint32), got[k].view(np.uint32)), k
