"""Learned block denoiser: training data from plain fusion, GPU training, and
the learned-vs-plain comparison.

Training pairs (denoiser training config): for each seeded small-room
scene (4x4x3 m, 20 primitives, 640x480, circular trajectory),
* input = plain fusion after a few noisy frames;
* target = plain fusion after many frames with the noise sigma halved.
Samples are the blocks observed in both volumes (a random subset per scene).
Training runs entirely on the GPU through vhm_denoiser_train_step (BF16
MFMA forward/backward, Adam).

Comparison (learned + end-to-end config): unseen scenes fused from a few
frames and denoised, against plain fusion of many frames, by Chamfer
distance, TSDF L1 and wall time (evaluate.py).
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from . import Denoiser, Noise, Volume, default_params
from . import synth


def _render_gpu(scene, K, poses, noise, torch):
    out = torch.empty((len(poses), K.height, K.width), dtype=torch.int16, device="cuda:0")
    synth.render(scene, K, poses, noise, 0, out)
    return out


def small_room_scene(seed: int, n_frames: int):
    w = synth.plain_small(n_frames=n_frames, seed=seed)
    return w


def halved_noise(seed: int) -> Noise:
    return Noise(a=0.0006, b=0.00095, c=0.4, dropout=0.01, max_depth=0.0, seed=seed, enabled=1)


def fuse(w, frames, poses, params, n_frames=None) -> tuple[Volume, float]:
    import torch
    vol = Volume(params)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    n = len(poses) if n_frames is None else n_frames
    for i in range(n):
        vol.integrate(frames[i], w.K, poses[i])
    vol.synchronize()
    return vol, time.perf_counter() - t0


def volume_params(w, max_blocks=None):
    return default_params(w.voxel_size, w.truncation, filter_sigma_spatial=w.filter_sigma_spatial,
                          filter_sigma_range=w.filter_sigma_range, hash_log2_buckets=w.hash_log2_buckets,
                          max_blocks=max_blocks or w.max_blocks, max_width=w.K.width, max_height=w.K.height)


def observed_keys(vol: Volume) -> np.ndarray:
    keys = vol.block_keys()
    if len(keys) == 0:
        return keys
    obs = np.zeros(len(keys), bool)
    for i in range(0, len(keys), 2048):
        b = vol.download_blocks(keys[i:i + 2048])
        obs[i:i + 2048] = (b["weight"] > 0).reshape(len(b), -1).any(1)
    return keys[obs]


def build_training_set(dn: Denoiser, seeds, frames_in=10, frames_target=300, blocks_per_scene=1000, log=print):
    """Appends training samples from each seed's scene to dn."""
    import torch
    for s in seeds:
        w = small_room_scene(s, frames_target)
        vp = volume_params(w, max_blocks=80_000)
        f_in = _render_gpu(w.scene, w.K, w.poses[:frames_in], w.noise, torch)
        vin, _ = fuse(w, f_in, w.poses[:frames_in], vp)
        f_t = _render_gpu(w.scene, w.K, w.poses, halved_noise(s), torch)
        vt, _ = fuse(w, f_t, w.poses, vp)
        k_in = {tuple(k) for k in observed_keys(vin)}
        keys = np.array([k for k in observed_keys(vt) if tuple(k) in k_in], np.int32).reshape(-1, 3)
        rng = np.random.default_rng(s)
        if len(keys) > blocks_per_scene:
            keys = keys[rng.choice(len(keys), blocks_per_scene, replace=False)]
        n = min(len(keys), dn.capacity - dn.count)
        if n > 0:
            dn.append(vin, keys[:n], vt)
        log(f"[train-data] scene {s}: {len(keys)} blocks (set {dn.count})")
        vin.close()
        vt.close()
        del f_in, f_t
        if dn.count >= dn.capacity:
            break


def train(dn: Denoiser, steps: int, batch: int = 4096, lr: float = 1e-3, seed: int = 0, log=print) -> list:
    """Adam steps on random contiguous windows of the device sample set."""
    rng = np.random.default_rng(seed)
    losses = []
    n = dn.count
    batch = min(batch, n, dn.batch_capacity)
    for it in range(steps):
        off = int(rng.integers(0, max(n - batch, 0) + 1))
        losses.append(dn.train_step(off, batch, lr))
        if it % 50 == 0 or it == steps - 1:
            log(f"[train] step {it}: loss {losses[-1]:.5f}")
    return losses


@dataclass
class Comparison:
    scene_seed: int
    plain_frames: int
    learned_frames: int
    plain: dict
    learned: dict
    few_plain: dict


def compare_on_scene(dn: Denoiser, w, seed: int, frames_plain: int, frames_learned: int, evaluate,
                     apply_kwargs=None, gt_stride=16, log=print) -> Comparison:
    """Plain fusion of frames_plain frames vs frames_learned frames + the
    denoiser (and the same few frames without it)."""
    import torch
    vp = volume_params(w)
    frames = _render_gpu(w.scene, w.K, w.poses[:frames_plain], w.noise, torch)
    gt = evaluate.gt_surface_points(w.scene, w.K, w.poses[:frames_plain:max(1, frames_plain // 30)],
                                    stride=gt_stride)

    def metrics(vol, wall):
        keys = vol.block_keys()
        tris = vol.extract_mesh()
        c = evaluate.chamfer_mm(tris, w.scene, gt)
        blocks = np.concatenate([vol.download_blocks(keys[i:i + 2048]) for i in range(0, len(keys), 2048)]) \
            if len(keys) else np.zeros((0, 512, 6, 3), dtype=vol.download_blocks(keys[:0]).dtype)
        l1 = evaluate.tsdf_l1_mm(keys, blocks, w.scene, w.voxel_size, w.truncation)
        return {**c, **l1, "wall_s": wall, "blocks": int(len(keys)), "triangles": int(len(tris))}

    vplain, t_plain = fuse(w, frames, w.poses[:frames_plain], vp)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tri = vplain.extract_mesh()
    t_plain += time.perf_counter() - t0
    plain = metrics(vplain, t_plain)
    plain["triangles"] = int(len(tri))
    vplain.close()

    vfew, t_few = fuse(w, frames, w.poses[:frames_learned], vp)
    few = metrics(vfew, t_few)
    vfew.close()

    vlr, t_lr = fuse(w, frames, w.poses[:frames_learned], vp)
    t0 = time.perf_counter()
    keys = observed_keys(vlr)
    dn.clear()
    out = {"cells_replaced": 0, "blocks_kept_plain": 0}
    for i in range(0, len(keys), dn.capacity):
        dn.clear()
        k = keys[i:i + dn.capacity]
        dn.append(vlr, k)
        dn.forward(0, len(k), fetch=False)
        r = dn.apply(vlr, **(apply_kwargs or {}))
        out = {a: out[a] + r[a] for a in out}
    vlr.synchronize()
    t_lr += time.perf_counter() - t0
    learned = metrics(vlr, t_lr)
    learned.update(out)
    vlr.close()
    log(f"[compare] seed {seed}: plain {frames_plain}f chamfer {plain['chamfer_mm']:.2f} mm "
        f"({plain['wall_s']:.2f} s); few {frames_learned}f {few['chamfer_mm']:.2f} mm; "
        f"learned {learned['chamfer_mm']:.2f} mm ({learned['wall_s']:.2f} s)")
    return Comparison(seed, frames_plain, frames_learned, plain, learned, few)
