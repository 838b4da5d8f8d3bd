"""Writes whole-configuration golden digests from the oracle.

    python tests/golden/make_config_golden.py [plain_small|plain_scale]

The oracle (oracle/cpu_ref.hpp, liboracle.so) is the restated reference CPU
path; tests/test_golden.py pins it bit for bit to the reference's own CPU
path compiled from its sources.  The reference build has its filter
constants compiled in (reconstruction_backend.cpp:26-32: sigma_s 1.5 px,
sigma_r 1.2 mm), so the configurations' filter (sigma_s 4.5 px, sigma_r
30 mm) runs through the oracle.

Two workloads, with the frame loop of the reference demo (main.cpp:433-455:
integrate every frame, stride 1) and its extraction (main.cpp:467-469):

* plain_small_640x480: all 100 frames of the plain-small room (640x480,
  voxel 1 cm, truncation 4 cm) and one marching-cubes extraction at the end.
* plain_scale_1280x720: frames 0-9 of the plain-at-scale random walk
  (1280x720, voxel 5 mm, truncation 2 cm) with extractions after frames 0
  and 9 (the first incremental extraction of that workload meshes every
  block, so it equals the full one).

Frames come from the seeded generator (synth.render); their digests pin the
input.  What is recorded per frame: the frame's digest, its active block
set (count and digest of the sorted keys), and the block count, observed
voxel count and voxel update count after it.  At checkpoint frames: the
digest of every block's 147,456 bytes (AoS Voxel layout) in sorted key
order, also per run of 1024 blocks to localise a difference.  At mesh
frames: triangle count and the digest of the canonical (sorted) triangle
multiset.  Only these digests are committed (config_<name>.json).
"""
from __future__ import annotations

import hashlib
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
sys.path.insert(0, str(HERE.parent))

import oracle_lib as O  # noqa: E402
import voxelhashmt_amd as vhm  # noqa: E402
from voxelhashmt_amd import synth  # noqa: E402

BLOCK_RUN = 1024


def digest(data) -> str:
    return hashlib.blake2b(np.ascontiguousarray(data).tobytes(), digest_size=16).hexdigest()


def block_digests(keys: np.ndarray, dump) -> tuple[str, list[str]]:
    """Digest of every block's AoS bytes, in the given (sorted) key order:
    the whole sequence and each run of BLOCK_RUN blocks.  dump(keys) returns
    the blocks (n x 512 x 6 x 3 Voxel records)."""
    per_block = []
    for s in range(0, len(keys), BLOCK_RUN):
        blocks = dump(keys[s:s + BLOCK_RUN])
        raw = np.ascontiguousarray(blocks).view(np.uint8).reshape(len(blocks), -1)
        per_block.extend(hashlib.blake2b(r.tobytes(), digest_size=8).digest() for r in raw)
    runs = [hashlib.blake2b(b"".join(per_block[s:s + BLOCK_RUN]), digest_size=8).hexdigest()
            for s in range(0, len(per_block), BLOCK_RUN)]
    return hashlib.blake2b(b"".join(per_block), digest_size=16).hexdigest(), runs


def mesh_record(tris: np.ndarray) -> dict:
    canon = O.canonical_mesh(tris)
    return {"triangles": int(len(canon)), "digest": digest(canon)}


def params(w):
    return vhm.default_params(w.voxel_size, w.truncation, filter_sigma_spatial=w.filter_sigma_spatial,
                              filter_sigma_range=w.filter_sigma_range, hash_log2_buckets=w.hash_log2_buckets,
                              max_blocks=w.max_blocks, max_width=w.K.width, max_height=w.K.height)


def workload(name: str):
    """(workload, frames used, checkpoint frames, mesh frames)."""
    if name == "plain_small":
        return synth.plain_small(), 100, (49, 99), (99,)
    if name == "plain_scale":
        return synth.plain_scale(), 10, (0, 9), (0, 9)
    raise SystemExit(f"unknown workload {name}")


def make(name: str) -> dict:
    w, n, checkpoints, mesh_frames = workload(name)
    frames = synth.render(w.scene, w.K, w.poses[:n], w.noise)
    op = O.oracle_params(params(w))
    ov = O.OracleVolume(op)
    rec = {"workload": w.name, "frames": n, "voxel_size": w.voxel_size, "truncation": w.truncation,
           "filter_sigma_spatial": w.filter_sigma_spatial, "filter_sigma_range": w.filter_sigma_range,
           "per_frame": [], "checkpoints": {}, "meshes": {}}
    t0 = time.time()
    for i in range(n):
        metres = frames[i].astype(np.float32) * np.float32(0.001)
        pose = vhm.pose_array(w.poses[i])
        scratch = O.OracleVolume(op)
        active = O.sorted_keys(scratch.prepare(O.oracle_filter(op, metres, w.K), w.K, pose))
        del scratch
        ov.integrate(metres, w.K, pose)
        rec["per_frame"].append({"frame": digest(frames[i]), "active": int(len(active)),
                                 "active_digest": digest(active), "blocks": int(ov.block_count()),
                                 "observed": int(ov.observed_voxel_count()), "updates": int(ov.updates())})
        print(f"{name} frame {i}: {len(active)} active, {ov.block_count()} blocks, {ov.updates()} updates "
              f"({time.time() - t0:.0f} s)", flush=True)
        if i in checkpoints:
            keys = O.sorted_keys(ov.block_keys())
            whole, runs = block_digests(keys, ov.dump_blocks)
            rec["checkpoints"][str(i)] = {"blocks": int(len(keys)), "keys_digest": digest(keys),
                                          "digest": whole, "runs": runs}
        if i in mesh_frames:
            rec["meshes"][str(i)] = mesh_record(ov.extract_mesh())
            print(f"{name} mesh after frame {i}: {rec['meshes'][str(i)]['triangles']} triangles", flush=True)
    return rec


if __name__ == "__main__":
    names = sys.argv[1:] or ["plain_small", "plain_scale"]
    for nm in names:
        out = make(nm)
        path = HERE / f"config_{out['workload']}.json"
        path.write_text(json.dumps(out, indent=1) + "\n")
        print(f"wrote {path}")
