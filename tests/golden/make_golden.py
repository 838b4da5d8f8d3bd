"""Writes golden fixtures from the reference's own CPU path.

    python tests/golden/make_golden.py      (needs oracle/_ref/libvhm_ref.so,
                                             built by `make -C oracle ref`)

For each small seeded scene the fixture holds the inputs (uint16 depth
frames, poses, intrinsics, volume geometry) and what the reference CPU path
(TsdfIntegrationBackend(kCpu) over VoxelHashTSDF) returns for them: the
filtered first frame, each frame's active block keys, the final block key
set, a 64-bit polynomial hash of every block's 147,456 bytes (AoS Voxel
layout), two full blocks, the observed-voxel count and the mesh (triangle
count, hash of the canonical multiset, first triangles).  The reference
itself never leaves this container; only these arrays are committed.
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
sys.path.insert(0, str(HERE.parent))

import oracle_lib as O  # noqa: E402
import voxelhashmt_amd as vhm  # noqa: E402
from voxelhashmt_amd import synth  # noqa: E402

CASES = [("tiny_seed3", 3, 96, 72, 4), ("tiny_seed17", 17, 128, 96, 3)]


def block_hash(blocks: np.ndarray) -> np.ndarray:
    """Per row: sum_i w_i * P^i (mod 2^64) over the row's 32-bit words."""
    words = np.ascontiguousarray(blocks).reshape(len(blocks), -1).view(np.uint32).astype(np.uint64)
    powers = np.empty(words.shape[1], np.uint64)
    acc = np.uint64(1)
    with np.errstate(over="ignore"):
        for i in range(words.shape[1]):
            powers[i] = acc
            acc = acc * np.uint64(0x100000001B3)
        return (words * powers).sum(axis=1, dtype=np.uint64)


def mesh_hash(canon: np.ndarray) -> np.uint64:
    return block_hash(np.ascontiguousarray(canon).reshape(1, -1))[0] if len(canon) else np.uint64(0)


def make(name, seed, width, height, n_frames):
    w = synth.tiny(n_frames=n_frames, seed=seed, width=width, height=height)
    frames = synth.render(w.scene, w.K, w.poses, w.noise)
    poses = np.stack([vhm.pose_array(p) for p in w.poses])
    K = np.array([w.K.width, w.K.height, w.K.fx, w.K.fy, w.K.cx, w.K.cy], np.float64)
    metres = frames.astype(np.float32) * np.float32(0.001)
    filtered0 = O.ref_filter(metres[0], w.K)
    rv = O.RefVolume(w.voxel_size, w.truncation)
    frame_keys, counts = [], []
    for i in range(n_frames):
        # the frame's active set, on the filtered frame, from a scratch volume
        scratch = O.RefVolume(w.voxel_size, w.truncation)
        frame_keys.append(O.sorted_keys(scratch.prepare(O.ref_filter(metres[i], w.K), w.K, poses[i])))
        rv.integrate(metres[i], w.K, poses[i])
        counts.append((rv.block_count(), rv.observed_voxel_count()))
    keys = O.sorted_keys(rv.block_keys())
    blocks = rv.dump_blocks(keys)
    hashes = block_hash(blocks)
    obs_per_block = (blocks["weight"] > 0).any(axis=(2, 3)).sum(axis=1)
    full_idx = np.argsort(-obs_per_block, kind="stable")[:2]
    mesh = O.canonical_mesh(rv.extract_mesh())
    out = dict(frames=frames, poses=poses, K=K, voxel_size=np.float32(w.voxel_size),
               truncation=np.float32(w.truncation), filtered0=filtered0,
               frame_key_counts=np.array([len(k) for k in frame_keys]),
               frame_keys=np.concatenate(frame_keys), block_keys=keys, block_hashes=hashes,
               full_block_index=full_idx, full_blocks=blocks[full_idx].view(np.uint32),
               counts=np.array(counts, np.int64), mesh_count=np.int64(len(mesh)), mesh_hash=mesh_hash(mesh),
               mesh_head=mesh[:64])
    np.savez_compressed(HERE / f"{name}.npz", **out)
    print(f"{name}: {len(keys)} blocks, {len(mesh)} triangles")


if __name__ == "__main__":
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libvhm_ref.so missing: run `make -C oracle ref`")
    for c in CASES:
        make(*c)
