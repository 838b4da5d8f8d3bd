"""Whole-configuration parity on the GPU against the oracle's golden digests.

tests/golden/make_config_golden.py ran the oracle (pinned bit for bit to the
reference's own CPU path by test_golden.py) over

* all 100 frames of the plain-small room (640x480, voxel 1 cm) and its one
  extraction at the end, and
* frames 0-9 of the plain-at-scale random walk (1280x720, voxel 5 mm) with
  extractions after frames 0 and 9,

with the configurations' filter (sigma_s 4.5 px, sigma_r 30 mm), and
committed per-frame digests.  Here the HIP path integrates the same frames
(rendered on the GPU; their digests pin the input) and must match every
one: each frame's active block set, the block / observed-voxel / update
counts after it, every block's bytes at the checkpoint frames and the
canonical triangle multiset at the mesh frames.  At frame 9 of the
plain-at-scale workload the first incremental extraction into a device
buffer (the fused pass the benchmark times) must also give that multiset.
"""
import json
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
import voxelhashmt_amd as vhm

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
import make_config_golden as G  # noqa: E402

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def _render(w, n):
    import torch
    frames = torch.empty((n, w.K.height, w.K.width), dtype=torch.int16, device="cuda:0")
    from voxelhashmt_amd import synth
    synth.render(w.scene, w.K, w.poses[:n], w.noise, 0, frames)
    return frames, frames.cpu().numpy().view(np.uint16)


def _run(name):
    import torch
    w, n, checkpoints, mesh_frames = G.workload(name)
    rec = json.loads((GOLDEN / f"config_{w.name}.json").read_text())
    assert rec["frames"] == n
    frames, host = _render(w, n)
    vp = G.params(w)
    sp = G.params(w)
    sp.max_blocks = 65536
    with vhm.Volume(vp) as vol, vhm.Volume(sp) as scratch:
        for i in range(n):
            g = rec["per_frame"][i]
            print(f"[config] {w.name} frame {i}", flush=True)
            assert G.digest(host[i]) == g["frame"], f"frame {i}: the generator's frame differs"
            scratch.reset()
            filtered = scratch.filter_depth(frames[i], w.K)
            active = O.sorted_keys(scratch.allocate_active_blocks(filtered, w.K, w.poses[i]))
            assert (len(active), G.digest(active)) == (g["active"], g["active_digest"]), f"frame {i}: active set"
            vol.integrate(frames[i], w.K, w.poses[i])
            got = (vol.block_count(), vol.observed_voxel_count(), vol.stats()["updates"])
            assert got == (g["blocks"], g["observed"], g["updates"]), f"frame {i}: counts {got}"
            if i in checkpoints:
                c = rec["checkpoints"][str(i)]
                keys = O.sorted_keys(vol.block_keys())
                assert (len(keys), G.digest(keys)) == (c["blocks"], c["keys_digest"])
                whole, runs = G.block_digests(keys, vol.download_blocks)
                bad = [r for r, (a, b) in enumerate(zip(runs, c["runs"])) if a != b]
                print(f"[config] {w.name} frame {i}: {len(keys)} blocks, runs differing: {bad}", flush=True)
                assert not bad and whole == c["digest"]
            if i in mesh_frames:
                m = rec["meshes"][str(i)]
                if i == n - 1 and name == "plain_scale":
                    cap = m["triangles"] + 1024
                    buf = torch.empty(cap * 9, dtype=torch.float32, device="cuda:0")
                    _, nt = vol.extract_mesh_incremental(device_tris=buf)
                    inc = buf[: 9 * nt].cpu().numpy().reshape(-1, 3, 3)
                    assert G.mesh_record(inc) == m, f"frame {i}: incremental (device) mesh"
                full = G.mesh_record(vol.extract_mesh())
                print(f"[config] {w.name} mesh after frame {i}: {full['triangles']} triangles", flush=True)
                assert full == m, f"frame {i}: mesh"


def test_plain_small_640x480_all_frames_and_mesh(gpu):
    _run("plain_small")


def test_plain_scale_1280x720_first_ten_frames_and_meshes(gpu):
    _run("plain_scale")
