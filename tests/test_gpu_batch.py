"""Batched multi-scene fusion (vhm_batch_*) against the one-volume path.

Every batched step (a replay of the captured hipGraph, or direct launches)
must leave each scene exactly where vhm_integrate_u16 + the incremental mesh
leave a separate volume fed the same frames: the same re-meshed block count
and triangle multiset after every step, then the same block keys and
bit-identical voxel layers.  The one-volume path is itself bit-exact against
the oracle (test_gpu_parity.py), so this pins the batched path to the oracle.
Scenes differ in resolution and sequence length; steps continue past the
shortest sequence.
"""
import numpy as np
import pytest

import oracle_lib as O
import voxelhashmt_amd as vhm
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


def _run(torch, use_graph, autotune):
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
        while step < steps:
            print(f"[batch] graph={use_graph} step {step}", flush=True)
            b.step(1)
            b.synchronize()
            for i, (w, frames) in enumerate(wl):
                if step < len(frames):
                    singles[i].integrate(frames[step], w.K, w.poses[step])
                keys, _, tris, _ = singles[i].extract_mesh_incremental()
                nb, nt, btris = b.mesh_output(i)
                assert nb == len(keys), f"scene {i} step {step}: re-meshed blocks {nb} vs {len(keys)}"
                assert nt == len(tris), f"scene {i} step {step}: triangles {nt} vs {len(tris)}"
                assert np.array_equal(O.canonical_mesh(btris), O.canonical_mesh(tris))
            step += 1
        assert b.steps_done == steps
        for i, (w, frames) in enumerate(wl):
            bv, sv = b.volumes[i], singles[i]
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
