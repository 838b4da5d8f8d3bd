"""Batched multi-scene fusion (vhm_batch_*) against the one-volume path.

Every batched step (a replay of the captured hipGraph, or direct launches)
must leave each scene exactly where vhm_integrate_u16 + the incremental mesh
leave a separate volume fed the same frames: the same listed (block,
direction) items with the same triangles per item after every step, then
the same block keys (in pool order) and bit-identical voxel layers.  The
one-volume path is itself bit-exact against the oracle (test_gpu_parity.py,
test_gpu_config_golden.py), so this pins the batched path to the oracle.
Scenes differ in resolution and sequence length; steps continue past the
shortest sequence.  The batched workload itself (48 rooms of 1k-60k
blocks at 320x240 to 1280x720) runs its first steps the same way, scene
volumes compared through device digests of every block.
"""
import numpy as np
import pytest

import oracle_lib as O
import voxelhashmt_amd as vhm
from meshitems import item_digest, per_item
from voxelhashmt_amd import synth

pytestmark = pytest.mark.gpu

SCENES = [dict(seed=21, width=96, height=72, n=5), dict(seed=22, width=160, height=120, n=3),
          dict(seed=23, width=128, height=96, n=4)]


def _workloads():
    out = []
    for s in SCENES:
        w = synth.tiny(n_frames=s["n"], seed=s["seed"], width=s["width"], height=s["height"])
        out.append((w, synth.render(w.scene, w.K, w.poses, w.noise)))
    return out


def _params(w):
    return vhm.default_params(w.voxel_size, w.truncation, filter_sigma_spatial=w.filter_sigma_spatial,
                              filter_sigma_range=w.filter_sigma_range, hash_log2_buckets=w.hash_log2_buckets,
                              max_blocks=w.max_blocks, max_width=w.K.width, max_height=w.K.height)


def _blocks(vol, keys):
    return vol.download_blocks(keys) if len(keys) else None


def _run(torch, use_graph, autotune, tiles=None):
    wl = _workloads()
    singles = [vhm.Volume(_params(w)) for w, _ in wl]
    with vhm.Batch([_params(w) for w, _ in wl]) as b:
        for i, (w, frames) in enumerate(wl):
            b.set_sequence(i, torch.from_numpy(frames.view(np.int16)).to("cuda:0"), w.K, w.poses,
                           tri_capacity=1 << 18)
        b.prepare(use_graph=use_graph)
        steps = max(s["n"] for s in SCENES) + 1
        step = 0
        if autotune:
            gi, gm = b.autotune((1, 4))
            print(f"[batch] autotuned grids integrate {gi} mesh {gm}", flush=True)
            for _ in range(2):  # autotune consumed two steps; feed the singles the same frames
                for i, (w, frames) in enumerate(wl):
                    if step < len(frames):
                        singles[i].integrate(frames[step], w.K, w.poses[step])
                        singles[i].extract_mesh_incremental()
                step += 1
        if tiles is not None:
            got = b.autotune_tiles(*tiles)
            print(f"[batch] tiles {tiles} -> filter {got[0]}, allocation {got[1]}, chunk {got[2]}", flush=True)
            assert got == tuple(t[-1] for t in tiles) or len(tiles[0]) > 1
            for _ in range(max(len(t) for t in tiles)):  # consumed steps
                for i, (w, frames) in enumerate(wl):
                    if step < len(frames):
                        singles[i].integrate(frames[step], w.K, w.poses[step])
                        singles[i].extract_mesh_incremental()
                step += 1
        while step < steps:
            print(f"[batch] graph={use_graph} step {step}", flush=True)
            b.step(1)
            b.synchronize()
            for i, (w, frames) in enumerate(wl):
                if step < len(frames):
                    singles[i].integrate(frames[step], w.K, w.poses[step])
                keys, dirs, tris, items = singles[i].extract_mesh_incremental()
                nb, nt, btris = b.mesh_output(i)
                assert nb == len(keys), f"scene {i} step {step}: re-meshed blocks {nb} vs {len(keys)}"
                assert nt == len(tris), f"scene {i} step {step}: triangles {nt} vs {len(tris)}"
                bkeys, bdirs, bitems = b.mesh_items(i)
                assert np.array_equal(bkeys, keys) and np.array_equal(bdirs, dirs)
                got, want = per_item(bkeys, btris, bitems), per_item(keys, tris, items, in_order=True)
                assert got.keys() == want.keys(), f"scene {i} step {step}: items differ"
                for k in want:
                    assert np.array_equal(got[k].view(np.uint32), want[k].view(np.uint32)), (i, step, k)
            step += 1
        assert b.steps_done == steps
        for i, (w, frames) in enumerate(wl):
            bv, sv = b.volumes[i], singles[i]
            assert np.array_equal(bv.block_keys(), sv.block_keys()), f"scene {i}: pool order differs"
            kb, ks = O.sorted_keys(bv.block_keys()), O.sorted_keys(sv.block_keys())
            assert np.array_equal(kb, ks), f"scene {i}: block keys differ"
            assert np.array_equal(_blocks(bv, kb).view(np.uint32), _blocks(sv, ks).view(np.uint32))
            assert bv.stats()["updates"] == sv.stats()["updates"]
            assert bv.stats()["frames"] == len(frames)
            assert bv.observed_voxel_count() == sv.observed_voxel_count()
            print(f"[batch] scene {i}: {len(kb)} blocks equal", flush=True)
            # the volume keeps working through the one-volume API after the batch
            bv.integrate(frames[0], w.K, w.poses[0])
            sv.integrate(frames[0], w.K, w.poses[0])
            assert np.array_equal(O.sorted_keys(bv.block_keys()), O.sorted_keys(sv.block_keys()))
            kk = O.sorted_keys(sv.block_keys())
            assert np.array_equal(_blocks(bv, kk).view(np.uint32), _blocks(sv, kk).view(np.uint32))
    for v in singles:
        v.close()


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_equals_single_volumes(gpu, use_graph):
    import torch
    _run(torch, use_graph, autotune=False)


def test_batch_autotune_keeps_results(gpu):
    import torch
    _run(torch, True, autotune=True)


@pytest.mark.parametrize("tiles", [((64,), (64,), (8,)), ((16,), (32,), (1,)), ((32, 16), (16, 64), (0, 32))])
def test_batch_tile_shapes_keep_results(gpu, tiles):
    """Filter / allocation tile widths and the integrator's chunk (forced
    single candidates, then a timed choice between two): every later step
    and the final volumes still equal the one-volume path."""
    import torch
    _run(torch, True, autotune=False, tiles=tiles)


def test_batched_48_rooms_first_steps(gpu):
    """The batched workload's 48 rooms (block targets 1k-60k, 320x240 to
    1280x720), first 3 steps through the captured graph: every scene's
    per-step incremental mesh (items and their triangles) and final volume
    (keys in pool order, counters, the device digest of every block) equal a
    one-volume run of the same frames, made one scene at a time."""
    import torch
    steps = 3
    ws = synth.batched_scenes()
    params = [_params(w) for w in ws]
    caps = [200 * w.target_blocks + 65536 for w in ws]
    frames = []
    for w in ws:
        t = torch.empty((steps, w.K.height, w.K.width), dtype=torch.int16, device="cuda:0")
        synth.render(w.scene, w.K, w.poses[:steps], w.noise, 0, t)
        frames.append(t)
    digests = [[] for _ in ws]
    with vhm.Batch(params) as b:
        for i, w in enumerate(ws):
            b.set_sequence(i, frames[i], w.K, w.poses[:steps], tri_capacity=caps[i])
        b.prepare(use_graph=True)
        for step in range(steps):
            print(f"[batch48] step {step}", flush=True)
            b.step(1)
            b.synchronize()
            for i in range(len(ws)):
                nb, nt, tris = b.mesh_output(i)
                assert nt <= caps[i], f"scene {i}: triangle buffer overflow"
                keys, dirs, items = b.mesh_items(i)
                digests[i].append(item_digest(keys, dirs, tris, items))
        final = []
        for i, v in enumerate(b.volumes):
            keys = v.block_keys()
            final.append((keys, v.block_digests(keys), v.stats()["updates"], v.observed_voxel_count()))
    total_blocks = 0
    for i, w in enumerate(ws):
        with vhm.Volume(params[i]) as sv:
            for step in range(steps):
                sv.integrate(frames[i][step], w.K, w.poses[step])
                keys, dirs, tris, items = sv.extract_mesh_incremental()
                assert item_digest(keys, dirs, tris, items) == digests[i][step], f"scene {i} step {step}: mesh"
            keys = sv.block_keys()
            assert np.array_equal(keys, final[i][0]), f"scene {i}: block keys / pool order"
            assert np.array_equal(sv.block_digests(keys), final[i][1]), f"scene {i}: block contents"
            assert (sv.stats()["updates"], sv.observed_voxel_count()) == final[i][2:], f"scene {i}: counters"
            total_blocks += len(keys)
        print(f"[batch48] scene {i} ({w.K.width}x{w.K.height}, {len(keys)} blocks) equal", flush=True)
    print(f"[batch48] {len(ws)} scenes, {total_blocks} blocks", flush=True)
