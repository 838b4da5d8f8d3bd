"""The learned configurations on the GPU.

* Denoiser training: samples from small rooms of the plain-small style
  (input: 10 noisy frames; target: 300 frames with the noise sigma halved),
  one training step at batch 4096.  The batch's masked-mean loss and
  gradient are the mask-weighted combination of its chunks' (the loss is
  linear in them), one chunk's gradient matches the FP64 oracle, and an
  Adam step at that batch lowers the batch loss.
* Learned + end-to-end: a scene of the plain-at-scale style (1280x720,
  voxel 5 mm) fused from 30 frames, the denoiser run on every observed
  block in one launch, then marching cubes.  With a block threshold below
  zero every block keeps plain fusion and the mesh equals, byte for byte,
  the mesh of the same 30 frames without the denoiser; with the default
  thresholds the write-back changes the mesh, and Chamfer distance and TSDF
  L1 against the analytic scene are reported for both.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

import voxelhashmt_amd as vhm
import voxelhashmt_amd.learned as L
from voxelhashmt_amd import evaluate as E
from voxelhashmt_amd import synth

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import denoiser_ref as D  # noqa: E402

pytestmark = pytest.mark.gpu

BATCH = 4096


def test_training_step_at_batch_4096(gpu):
    with vhm.Denoiser(capacity=BATCH + 1024, batch_capacity=BATCH) as dn:
        dn.set_params(L.init_params(0))
        L.build_training_set(dn, range(1000, 1010), blocks_per_scene=1000)
        print(f"[c3] {dn.count} samples", flush=True)
        assert dn.count >= BATCH
        _, _, m = dn.samples(0, BATCH)
        mcount = m.reshape(BATCH, -1).sum(1)
        loss, g = dn.train_step(0, BATCH, 0.0, want_grad=True)
        chunk = 256
        acc = np.zeros_like(g, dtype=np.float64)
        lacc = 0.0
        for o in range(0, BATCH, chunk):
            lc, gc = dn.train_step(o, chunk, 0.0, want_grad=True)
            nc = max(float(mcount[o:o + chunk].sum()), 1.0)
            acc += gc.astype(np.float64) * nc
            lacc += lc * nc
        n_all = max(float(mcount.sum()), 1.0)
        acc /= n_all
        lacc /= n_all
        rel = np.linalg.norm(g - acc) / np.linalg.norm(acc)
        print(f"[c3] batch loss {loss:.6f} vs chunks {lacc:.6f}; gradient rel err {rel:.2e}", flush=True)
        assert abs(loss - lacc) <= 1e-4 * abs(lacc) and rel < 1e-4
        # one chunk of 16 samples against the FP64 oracle
        x, t, mm = dn.samples(0, 16)
        X = D.bf16_to_float(x)[..., :12].astype(np.float64)
        pr = D.flat_to_params(dn.get_params().astype(np.float64))
        l16, g16 = dn.train_step(0, 16, 0.0, want_grad=True)
        Lr, gr = D.loss_and_grads(pr, X, t.reshape(16, 8, 8, 8, 6).astype(np.float64),
                                  mm.reshape(16, 8, 8, 8, 6).astype(np.float64))
        gflat = D.params_to_flat(gr)
        cos = float(g16 @ gflat / (np.linalg.norm(g16) * np.linalg.norm(gflat)))
        rel16 = np.linalg.norm(g16 - gflat) / np.linalg.norm(gflat)
        print(f"[c3] 16-sample chunk: loss {l16:.6f} vs fp64 {Lr:.6f}; gradient rel err {rel16:.2e}, cos {cos:.6f}",
              flush=True)
        assert abs(l16 - Lr) <= 2e-2 * abs(Lr) and rel16 < 5e-2 and cos > 0.998
        before = dn.train_step(0, BATCH, 0.0)
        for _ in range(3):
            dn.train_step(0, BATCH, 1e-3)
        after = dn.train_step(0, BATCH, 0.0)
        print(f"[c3] batch loss {before:.6f} -> {after:.6f} after 3 Adam steps", flush=True)
        assert after < before


def test_learned_pipeline_fallback_and_write_back(gpu):
    import torch
    w = synth.plain_scale(n_frames=300, seed=2000)
    idx = L.spread(300, 30)
    frames = L.render_gpu(w, w.noise, [w.poses[int(i)] for i in idx])
    vp = L.volume_params(w, max_blocks=400_000)
    with vhm.Volume(vp) as plain, vhm.Volume(vp) as vol:
        for j, i in enumerate(idx):
            plain.integrate(frames[j], w.K, w.poses[int(i)])
            vol.integrate(frames[j], w.K, w.poses[int(i)])
        ref_mesh = plain.extract_mesh()
        keys = vol.observed_block_keys()
        print(f"[c4] {len(keys)} observed blocks, plain mesh {len(ref_mesh)} triangles", flush=True)
        with vhm.Denoiser(capacity=len(keys), batch_capacity=BATCH) as dn:
            dn.set_params(L.init_params(0))
            dn.append(vol, keys)
            dn.forward(0, len(keys), fetch=False)
            r = dn.apply(vol, block_max_delta=-1.0)
            assert (r["cells_replaced"], r["blocks_kept_plain"]) == (0, len(keys))
            mesh = vol.extract_mesh()
            assert mesh.tobytes() == ref_mesh.tobytes()
            r = dn.apply(vol, max_delta=0.5, block_max_delta=0.25)
            print(f"[c4] default thresholds: {r}", flush=True)
            assert r["cells_replaced"] > 0
            learned = vol.extract_mesh()
            assert learned.tobytes() != ref_mesh.tobytes()
        gt = E.gt_surface_points(w.scene, w.K, w.poses[::10], stride=8, max_points=100_000)
        sample = keys[np.random.default_rng(0).choice(len(keys), min(1000, len(keys)), replace=False)]
        cp, cl = E.chamfer_mm(ref_mesh, w.scene, gt), E.chamfer_mm(learned, w.scene, gt)
        lp = E.tsdf_l1_mm(sample, plain.download_blocks(sample), w.scene, w.voxel_size, w.truncation)
        ll = E.tsdf_l1_mm(sample, vol.download_blocks(sample), w.scene, w.voxel_size, w.truncation)
        print(f"[c4] chamfer plain {cp['chamfer_mm']:.3f} mm, learned (untrained) {cl['chamfer_mm']:.3f} mm; "
              f"tsdf L1 plain {lp['tsdf_l1_mm']:.3f} mm, learned {ll['tsdf_l1_mm']:.3f} mm", flush=True)
        assert np.isfinite(cp["chamfer_mm"]) and np.isfinite(cl["chamfer_mm"])
    del frames
    torch.cuda.synchronize()
