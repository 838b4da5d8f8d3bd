"""Learned block denoiser: training data from plain fusion, GPU training, and
the learned-vs-plain comparison.

    python -m voxelhashmt_amd.learned [--train-scenes 200] [--steps 1500] [--eval-scenes 16] [--out r.json]

Training pairs (denoiser training config): for each seeded small room
(4x4x3 m, 20 primitives, 640x480, 300 frames on a circle of radius 1.5 m,
voxel 1 cm, truncation 4 cm):
* input  = plain fusion of 10 noisy frames spread evenly over the circle;
* target = plain fusion of all 300 frames rendered with the noise sigma
  halved.
Samples are blocks observed in both volumes (a seeded subset per scene).
Training runs on the GPU through vhm_denoiser_train_step: BF16 MFMA
forward/backward kernels, FP32 master weights, Adam, batches of 4096 blocks.

Comparison (learned + end-to-end config): unseen seeded rooms in the
plain-at-scale style (8x8x3 m, 150 primitives, 1280x720, random walk, voxel
5 mm, truncation 2 cm).  Plain fusion of all 300 frames against fusion of
30 of them (every 10th) followed by the denoiser on every observed block in
one launch, both finished by a full marching-cubes extraction into device
memory; and the same 30 frames without the denoiser.  Reported per scene:
Chamfer distance (mm) against the analytic surface, TSDF L1 (mm) on a
common block sample, and wall time on the GPU (fusion + denoiser +
extraction, frames resident in HBM, the mesh left in HBM; the metrics copy
it to the host afterwards).  Both volumes use the same units of trunc and voxel (4 voxels), which
the network's normalised inputs rely on.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from dataclasses import asdict, dataclass

import numpy as np

from . import Denoiser, Noise, Volume, default_params
from . import evaluate as E
from . import synth


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def halved_noise(seed: int) -> Noise:
    return Noise(a=0.0006, b=0.00095, c=0.4, dropout=0.01, max_depth=0.0, seed=seed, enabled=1)


def volume_params(w, max_blocks=None):
    return default_params(w.voxel_size, w.truncation, filter_sigma_spatial=w.filter_sigma_spatial,
                          filter_sigma_range=w.filter_sigma_range, hash_log2_buckets=w.hash_log2_buckets,
                          max_blocks=max_blocks or w.max_blocks, max_width=w.K.width, max_height=w.K.height)


def render_gpu(w, noise, poses=None):
    import torch
    poses = w.poses if poses is None else poses
    out = torch.empty((len(poses), w.K.height, w.K.width), dtype=torch.int16, device="cuda:0")
    synth.render(w.scene, w.K, poses, noise, 0, out)
    return out


def spread(n_total: int, n: int) -> np.ndarray:
    """n frame indices spread evenly over [0, n_total)."""
    return (np.arange(n) * n_total) // n


def fuse(vol: Volume, w, frames, idx) -> float:
    """Integrates frames[idx] (device tensor) into vol; returns seconds."""
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in idx:
        vol.integrate(frames[int(i)], w.K, w.poses[int(i)])
    vol.synchronize()
    return time.perf_counter() - t0


def build_training_set(dn: Denoiser, seeds, frames_in=10, frames_target=300, blocks_per_scene=1000,
                       target_noise="halved", style="small"):
    """Appends training samples from each seed's room to dn: a small room
    (style "small": 640x480, voxel 1 cm, the training configuration) or a
    room in the plain-at-scale style ("scale": 8x8x3 m, 1280x720 random walk,
    voxel 5 mm, the evaluation's style).  The target frames carry the noise
    sigma halved ("halved", the training configuration) or no noise at all
    ("none")."""
    vin = vt = None
    for s in seeds:
        if style == "scale":
            w = synth.plain_scale(n_frames=frames_target, seed=s)
        else:
            w = synth.plain_small(n_frames=frames_target, seed=s)
        if vin is None:
            vp = volume_params(w, max_blocks=80_000 if style == "small" else 400_000)
            vin, vt = Volume(vp), Volume(vp)
        else:
            vin.reset()
            vt.reset()
        frames = render_gpu(w, w.noise)
        fuse(vin, w, frames, spread(frames_target, frames_in))
        del frames
        frames = render_gpu(w, halved_noise(s) if target_noise == "halved" else Noise(enabled=0))
        fuse(vt, w, frames, range(frames_target))
        del frames
        k_in = vin.observed_block_keys()
        k_t = vt.observed_block_keys()
        common = np.array(sorted({tuple(k) for k in k_in} & {tuple(k) for k in k_t}), np.int32).reshape(-1, 3)
        rng = np.random.default_rng(s)
        if len(common) > blocks_per_scene:
            common = common[rng.choice(len(common), blocks_per_scene, replace=False)]
        n = min(len(common), dn.capacity - dn.count)
        if n > 0:
            dn.append(vin, common[:n], vt)
        log(f"[train-data] scene {s}: {len(common)} blocks (set {dn.count})")
        if dn.count >= dn.capacity:
            break
    for v in (vin, vt):
        if v is not None:
            v.close()


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform W1, W2 (bound sqrt(6 / fan_in)), W3 at a tenth of that,
    zero biases; flat float32 in the order W1, b1, W2, b2, W3, b3."""
    rng = np.random.Generator(np.random.PCG64(seed))
    parts = []
    for shape, scale in (((3, 3, 3, 12, 32), 1.0), ((32,), 0.0), ((3, 3, 3, 32, 32), 1.0), ((32,), 0.0),
                         ((3, 3, 3, 32, 6), 0.1), ((6,), 0.0)):
        if scale == 0.0:
            parts.append(np.zeros(shape))
        else:
            lim = np.sqrt(6.0 / (27 * shape[3])) * scale
            parts.append(rng.uniform(-lim, lim, size=shape))
    return np.concatenate([p.reshape(-1) for p in parts]).astype(np.float32)


def train(dn: Denoiser, steps: int, n: int, batch: int = 4096, lr: float = 1e-3, seed: int = 0) -> list:
    """Adam steps on seeded random contiguous windows of samples [0, n)."""
    rng = np.random.default_rng(seed)
    losses = []
    batch = min(batch, n, dn.batch_capacity)
    t0 = time.perf_counter()
    for it in range(steps):
        off = int(rng.integers(0, max(n - batch, 0) + 1))
        losses.append(dn.train_step(off, batch, lr))
        if it % 50 == 0 or it == steps - 1:
            log(f"[train] step {it}: loss {losses[-1]:.5f} ({time.perf_counter() - t0:.1f} s)")
    return losses


def held_out_loss(dn: Denoiser, offset: int, n: int, batch: int) -> tuple[float, float]:
    """(masked loss of the network, masked loss of the identity -- plain
    few-frame fusion) over samples [offset, offset + n)."""
    ln = li = 0.0
    cnt = 0
    saved = dn.get_params()
    for o in range(offset, offset + n, batch):
        m = min(batch, offset + n - o)
        ln += dn.train_step(o, m, 0.0) * m
        cnt += m
    # identity: zero residual head (W3, b3 = 0) leaves y = input tsdf
    ident = saved.copy()
    ident[-(27 * 32 * 6 + 6):] = 0.0
    dn.set_params(ident)
    for o in range(offset, offset + n, batch):
        m = min(batch, offset + n - o)
        li += dn.train_step(o, m, 0.0) * m
    dn.set_params(saved)
    return ln / max(cnt, 1), li / max(cnt, 1)


@dataclass
class SceneResult:
    seed: int
    plain: dict
    few: dict
    learned: dict


def _metrics(vol: Volume, w, gt, sample_keys, wall: float, tris) -> dict:
    c = E.chamfer_mm(tris, w.scene, gt)
    blocks = vol.download_blocks(sample_keys)
    l1 = E.tsdf_l1_mm(sample_keys, blocks, w.scene, w.voxel_size, w.truncation)
    return {**c, **l1, "wall_s": wall, "triangles": int(len(tris))}


def compare_on_scene(dn: Denoiser, w, frames_plain=300, frames_learned=30, apply_kwargs=None,
                     l1_blocks=2000) -> SceneResult:
    import torch
    vp = volume_params(w)
    frames = render_gpu(w, w.noise, w.poses[:frames_plain])
    gt = E.gt_surface_points(w.scene, w.K, w.poses[:frames_plain:max(1, frames_plain // 30)], stride=8,
                             max_points=200_000)
    few_idx = spread(frames_plain, frames_learned)
    out = {}
    vol = Volume(vp)
    # learned: few frames + denoiser on every observed block + extraction
    t = fuse(vol, w, frames, few_idx)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    keys = vol.observed_block_keys()
    dn.clear()
    n = min(len(keys), dn.capacity)
    dn.append(vol, keys[:n])
    dn.forward(0, n, fetch=False)
    r = dn.apply(vol, **(apply_kwargs or {}))
    tris = vol.extract_mesh(device=True)
    t += time.perf_counter() - t0
    tris = tris.cpu().numpy()  # for the metrics, after the timed region
    rng = np.random.default_rng(w.noise.seed)
    sample_keys = keys[rng.choice(len(keys), min(l1_blocks, len(keys)), replace=False)] if len(keys) else keys
    out["learned"] = {**_metrics(vol, w, gt, sample_keys, t, tris), **r, "blocks_denoised": int(n)}
    # the same few frames without the denoiser
    vol.reset()
    t = fuse(vol, w, frames, few_idx)
    t0 = time.perf_counter()
    tris = vol.extract_mesh(device=True)
    t += time.perf_counter() - t0
    out["few"] = _metrics(vol, w, gt, sample_keys, t, tris.cpu().numpy())
    # plain: every frame
    vol.reset()
    t = fuse(vol, w, frames, range(frames_plain))
    t0 = time.perf_counter()
    tris = vol.extract_mesh(device=True)
    t += time.perf_counter() - t0
    out["plain"] = _metrics(vol, w, gt, sample_keys, t, tris.cpu().numpy())
    vol.close()
    del frames
    log(f"[compare] seed {w.noise.seed}: plain {frames_plain}f chamfer {out['plain']['chamfer_mm']:.3f} mm "
        f"l1 {out['plain']['tsdf_l1_mm']:.3f} mm ({out['plain']['wall_s']:.2f} s); few {frames_learned}f "
        f"{out['few']['chamfer_mm']:.3f} mm l1 {out['few']['tsdf_l1_mm']:.3f}; learned "
        f"{out['learned']['chamfer_mm']:.3f} mm l1 {out['learned']['tsdf_l1_mm']:.3f} "
        f"({out['learned']['wall_s']:.2f} s)")
    return SceneResult(int(w.noise.seed), out["plain"], out["few"], out["learned"])


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--train-scenes", type=int, default=200)
    ap.add_argument("--train-style", choices=["small", "scale"], default="small",
                    help="training rooms: the training configuration's small rooms, or the evaluation's style")
    ap.add_argument("--train-seed", type=int, default=1000)
    ap.add_argument("--blocks-per-scene", type=int, default=1000)
    ap.add_argument("--frames-in", type=int, default=10)
    ap.add_argument("--frames-target", type=int, default=300)
    ap.add_argument("--target-noise", choices=["halved", "none"], default="halved",
                    help="noise of the many-frame target fusion: sigma halved (the training configuration) or none")
    ap.add_argument("--holdout-scenes", type=int, default=8, help="of the training scenes, kept out of training")
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval-scenes", type=int, default=16)
    ap.add_argument("--eval-seed", type=int, default=2000)
    ap.add_argument("--eval-frames-plain", type=int, default=300)
    ap.add_argument("--eval-frames-learned", type=int, default=30)
    ap.add_argument("--apply-max-delta", type=float, default=0.5, help="per-cell confidence bound (tsdf units)")
    ap.add_argument("--apply-block-max-delta", type=float, default=0.25,
                    help="per-block fallback to plain fusion above this mean |change|")
    ap.add_argument("--params-in", default=None, help="skip training: load parameters (.npy float32)")
    ap.add_argument("--params-out", default=None)
    ap.add_argument("--out", default=None, help="results JSON")
    args = ap.parse_args(argv)

    import torch  # noqa: F401  (device memory for the frames)
    res: dict = {"config": vars(args)}
    dn = Denoiser(capacity=max(args.train_scenes * args.blocks_per_scene, 400_000), batch_capacity=args.batch)
    if args.params_in:
        dn.set_params(np.load(args.params_in, allow_pickle=False).astype(np.float32))
    else:
        dn.set_params(init_params(0))
        t0 = time.perf_counter()
        seeds = list(range(args.train_seed, args.train_seed + args.train_scenes))
        build_training_set(dn, seeds, args.frames_in, args.frames_target, args.blocks_per_scene, args.target_noise,
                           args.train_style)
        res["train_set_blocks"] = dn.count
        res["train_set_seconds"] = time.perf_counter() - t0
        # the last scenes' samples are held out of training
        hold = min(args.holdout_scenes * args.blocks_per_scene, dn.count // 5)
        n_train = dn.count - hold
        t0 = time.perf_counter()
        losses = train(dn, args.steps, n_train, args.batch, args.lr)
        torch.cuda.synchronize()
        res["train_seconds"] = time.perf_counter() - t0
        res["train_loss_first"] = losses[0] if losses else None
        res["train_loss_last"] = float(np.mean(losses[-20:])) if losses else None
        if hold > 0:
            net, ident = held_out_loss(dn, n_train, hold, args.batch)
            res["holdout_loss_network"] = net
            res["holdout_loss_identity"] = ident
            log(f"[train] held-out masked MSE: network {net:.5f} vs plain few-frame {ident:.5f}")
        if args.params_out:
            np.save(args.params_out, dn.get_params())
    scenes = []
    for i in range(args.eval_scenes):
        w = synth.plain_scale(n_frames=args.eval_frames_plain, seed=args.eval_seed + i)
        scenes.append(asdict(compare_on_scene(dn, w, args.eval_frames_plain, args.eval_frames_learned,
                                              apply_kwargs={"max_delta": args.apply_max_delta,
                                                            "block_max_delta": args.apply_block_max_delta})))
        if args.out:
            res["scenes"] = scenes
            with open(args.out, "w") as f:
                json.dump(res, f, indent=1)
    res["scenes"] = scenes
    if scenes:
        def mean(k, m):
            return float(np.mean([s[k][m] for s in scenes]))
        res["summary"] = {k: {m: mean(k, m) for m in ("chamfer_mm", "tsdf_l1_mm", "wall_s")}
                          for k in ("plain", "few", "learned")}
        res["summary"]["speedup_plain_over_learned"] = res["summary"]["plain"]["wall_s"] / \
            max(res["summary"]["learned"]["wall_s"], 1e-9)
    print(json.dumps(res.get("summary", {})))
    if args.out:
        with open(args.out, "w") as f:
            json.dump(res, f, indent=1)
    dn.close()


if __name__ == "__main__":
    main()
