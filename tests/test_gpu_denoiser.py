"""The BF16 MFMA denoiser kernels against the FP64 oracle
(oracle/denoiser_ref.py): operand layout, feature gather, forward,
gradients, training progress and the write-back fallback."""
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
import voxelhashmt_amd as vhm
from voxelhashmt_amd import synth

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import denoiser_ref as D  # noqa: E402

pytestmark = pytest.mark.gpu


def test_mfma_operand_layout(gpu):
    out = vhm.mfma_layout_probe()
    for lane in range(64):
        for r in range(4):
            i, n = 4 * (lane // 16) + r, lane % 16
            assert out[lane, r] == i + n / 16.0, (lane, r, out[lane, r])


def test_mfma_k32_operand_layout(gpu):
    """v_mfma_f32_16x16x32_bf16 with operands placed as the kernels place
    them equals A B on small integers (exact in bf16 and fp32)."""
    rng = np.random.default_rng(1)
    a = rng.integers(-3, 4, size=(16, 32)).astype(np.float32)
    b = rng.integers(-3, 4, size=(32, 16)).astype(np.float32)
    out = np.zeros(256, np.float32)
    assert vhm.lib().vhm_mfma_layout_probe_k32(a.ctypes.data, b.ctypes.data, out.ctypes.data) == 0
    d = a @ b
    for lane in range(64):
        for r in range(4):
            assert out[4 * lane + r] == d[4 * (lane // 16) + r, lane % 16], (lane, r)


def _scene_volume(seed=5, frames=3):
    w = synth.tiny(n_frames=frames, seed=seed)
    f = synth.render(w.scene, w.K, w.poses, w.noise)
    vp = vhm.default_params(w.voxel_size, w.truncation, hash_log2_buckets=16, max_blocks=16384,
                            max_width=w.K.width, max_height=w.K.height)
    vol = vhm.Volume(vp)
    for i in range(frames):
        vol.integrate(f[i], w.K, w.poses[i])
    vol.synchronize()
    return w, vol


def _features_x(x_bits):
    return (x_bits.astype(np.uint32) << 16).view(np.float32)


def test_gather_equals_numpy_features(gpu):
    w, vol = _scene_volume()
    keys = O.sorted_keys(vol.block_keys())
    blocks = vol.download_blocks(keys)
    lut = {tuple(k): blocks[i] for i, k in enumerate(keys)}
    # blocks with observations
    obs = (blocks["weight"] > 0).reshape(len(keys), -1).sum(1)
    pick = keys[np.argsort(-obs)[:6]]
    with vhm.Denoiser(64, 16) as dn:
        dn.append(vol, pick, vol)
        x, t, m = dn.samples(0, len(pick))
    for i, k in enumerate(pick):
        ref = D.bf16_round(D.block_features(lut, tuple(k), w.voxel_size, w.truncation))
        assert np.array_equal(_features_x(x[i]).view(np.uint32), ref.view(np.uint32))
        # targets: the unrounded features of the same volume at the interior
        full = D.block_features(lut, tuple(k), w.voxel_size, w.truncation)[1:9, 1:9, 1:9]
        assert np.array_equal(t[i].reshape(8, 8, 8, 6), full[..., :6])
        assert np.array_equal(m[i].reshape(8, 8, 8, 6) > 0, full[..., 6:12] > 0)
    vol.close()


def _random_batch(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1, 1, size=(n, 10, 10, 10, 12))
    X[..., 6:] = np.abs(X[..., 6:])
    return X


def _load_synthetic_samples(dn, vol, keys):
    dn.append(vol, keys, vol)


def test_forward_and_gradients_match_fp64_oracle(gpu):
    w, vol = _scene_volume(seed=9, frames=4)
    keys = O.sorted_keys(vol.block_keys())
    obs = (vol.download_blocks(keys)["weight"] > 0).reshape(len(keys), -1).sum(1)
    pick = keys[np.argsort(-obs)[:8]]
    p = D.init_params(7)
    p["W3"] *= 10.0
    rng = np.random.default_rng(2)
    for k in ("b1", "b2", "b3"):
        p[k] = rng.normal(scale=0.05, size=p[k].shape)
    flat = D.params_to_flat(p).astype(np.float32)
    with vhm.Denoiser(64, 16) as dn:
        dn.set_params(flat)
        dn.append(vol, pick, vol)
        x, t, m = dn.samples(0, len(pick))
        # the targets here are the inputs' own features; perturb them so the loss is not ~0
        y = dn.forward(0, len(pick))
        X = _features_x(x)[..., :12].astype(np.float64)
        pr = D.flat_to_params(flat.astype(np.float64))
        yr, _ = D.forward(pr, X)
        err = np.sqrt(((y.reshape(yr.shape) - yr) ** 2).mean()) / np.sqrt((yr ** 2).mean())
        print(f"[denoiser] forward rms rel err vs fp64 = {err:.2e}", flush=True)
        assert err < 1e-2
        loss, g = dn.train_step(0, len(pick), 0.0, want_grad=True)
        T = t.reshape(yr.shape).astype(np.float64)
        M = m.reshape(yr.shape).astype(np.float64)
        Lr, gr = D.loss_and_grads(pr, X, T, M)
        print(f"[denoiser] loss gpu={loss:.6e} fp64={Lr:.6e}", flush=True)
        assert abs(loss - Lr) <= 2e-2 * abs(Lr) + 1e-9
        gflat = D.params_to_flat(gr)
        o = 0
        for name in D.ORDER:
            nparam = int(np.prod(D.SHAPES[name]))
            a, b = g[o:o + nparam].astype(np.float64), gflat[o:o + nparam]
            o += nparam
            rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            cos = float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30))
            print(f"[denoiser] grad {name}: rel L2 err {rel:.2e}, cosine {cos:.6f}", flush=True)
            assert rel < 5e-2 and cos > 0.998, name
    vol.close()


def test_training_reduces_loss_and_apply_fallback(gpu):
    """Inputs after 2 noisy frames, targets after 8 frames of the same scene:
    a few Adam steps reduce the masked loss; apply() with a block threshold
    below zero keeps every block plain and leaves the volume unchanged."""
    w = synth.tiny(n_frames=8, seed=4)
    f = synth.render(w.scene, w.K, w.poses, w.noise)
    vp = vhm.default_params(w.voxel_size, w.truncation, hash_log2_buckets=16, max_blocks=16384,
                            max_width=w.K.width, max_height=w.K.height)
    few, many = vhm.Volume(vp), vhm.Volume(vp)
    for i in range(8):
        if i in (0, 4):
            few.integrate(f[i], w.K, w.poses[i])
        many.integrate(f[i], w.K, w.poses[i])
    kf = {tuple(k) for k in few.block_keys()}
    keys = np.array([k for k in O.sorted_keys(many.block_keys()) if tuple(k) in kf], np.int32)[:512]
    with vhm.Denoiser(1024, 512) as dn:
        dn.set_params(D.params_to_flat(D.init_params(1)).astype(np.float32))
        dn.append(few, keys, many)
        losses = [dn.train_step(0, len(keys), 1e-3) for _ in range(30)]
        print(f"[denoiser] loss {losses[0]:.4e} -> {losses[-1]:.4e}", flush=True)
        assert losses[-1] < 0.8 * losses[0]
        dn.forward(0, len(keys), fetch=False)
        before = few.download_blocks(keys)
        r = dn.apply(few, block_max_delta=-1.0)
        assert r["cells_replaced"] == 0 and r["blocks_kept_plain"] == len(keys)
        assert np.array_equal(before.view(np.uint8), few.download_blocks(keys).view(np.uint8))
        r = dn.apply(few, max_delta=10.0, block_max_delta=10.0)
        assert r["cells_replaced"] > 0
    few.close()
    many.close()


def _pair(seed, frames_in=2, frames_target=12, sigma=0.004):
    """(few-frame volume, many-frame half-noise volume, their common observed
    keys) for one small seeded room rendered at 160x120."""
    w = synth.tiny(n_frames=frames_target, seed=seed, width=160, height=120)
    noisy = vhm.Noise(a=sigma, b=0.0, c=0.4, dropout=0.01, max_depth=0.0, seed=seed, enabled=1)
    clean = vhm.Noise(a=sigma / 2, b=0.0, c=0.4, dropout=0.01, max_depth=0.0, seed=seed + 7919, enabled=1)
    f_in = synth.render(w.scene, w.K, w.poses, noisy)
    f_t = synth.render(w.scene, w.K, w.poses, clean)
    vp = vhm.default_params(w.voxel_size, w.truncation, hash_log2_buckets=16, max_blocks=16384,
                            max_width=w.K.width, max_height=w.K.height)
    few, many = vhm.Volume(vp), vhm.Volume(vp)
    for i in range(frames_target):
        if i % (frames_target // frames_in) == 0:
            few.integrate(f_in[i], w.K, w.poses[i])
        many.integrate(f_t[i], w.K, w.poses[i])
    kf = {tuple(k) for k in few.observed_block_keys()}
    keys = np.array([k for k in O.sorted_keys(many.observed_block_keys()) if tuple(k) in kf], np.int32)
    return few, many, keys


def test_trained_network_beats_plain_few_frame_fusion(gpu):
    """The gain test: trained on pairs from 6 rooms, the network's masked
    error against the many-frame target on 2 unseen rooms is below that of
    plain few-frame fusion (the identity: zero residual head) -- what the
    learned part adds over the method without it."""
    import voxelhashmt_amd.learned as L
    with vhm.Denoiser(8192, 1024) as dn:
        dn.set_params(L.init_params(3))
        for seed in range(40, 48):
            few, many, keys = _pair(seed)
            dn.append(few, keys[:600], many)
            few.close()
            many.close()
            if seed == 45:
                n_train = dn.count
        print(f"[gain] {n_train} training blocks, {dn.count - n_train} held out", flush=True)
        net0, ident = L.held_out_loss(dn, n_train, dn.count - n_train, 1024)
        losses = L.train(dn, 300, n_train, batch=1024, lr=2e-3, seed=1)
        net, ident2 = L.held_out_loss(dn, n_train, dn.count - n_train, 1024)
        print(f"[gain] train loss {losses[0]:.5f} -> {np.mean(losses[-10:]):.5f}; held-out: identity {ident:.5f}, "
              f"untrained {net0:.5f}, trained {net:.5f}", flush=True)
        assert abs(ident - ident2) <= 1e-4 * max(ident, 1e-12) + 1e-9  # the identity ignores W1, W2
        assert net < 0.9 * ident


def test_adam_kernel_matches_oracle(gpu):
    """Adam's step in dnReduceKernel (FP32 master weights and moments)
    against the oracle's adam_step (FP64; itself pinned to
    torch.optim.Adam) over two steps fed the GPU's own gradients; the
    gradient reduction is deterministic."""
    w, vol = _scene_volume(seed=9, frames=4)
    keys = O.sorted_keys(vol.block_keys())
    obs = (vol.download_blocks(keys)["weight"] > 0).reshape(len(keys), -1).sum(1)
    pick = keys[np.argsort(-obs)[:8]]
    p = D.init_params(5)
    p["W3"] *= 10.0
    p0 = D.params_to_flat(p).astype(np.float32)
    with vhm.Denoiser(64, 16) as dn:
        dn.set_params(p0)
        dn.append(vol, pick, vol)
        _, g1 = dn.train_step(0, len(pick), 0.0, want_grad=True)
        _, g1b = dn.train_step(0, len(pick), 1e-3, want_grad=True)
        assert np.array_equal(g1, g1b)
        p1 = dn.get_params()
        _, g2 = dn.train_step(0, len(pick), 1e-3, want_grad=True)
        p2 = dn.get_params()
    ref = D.flat_to_params(p0.astype(np.float64))
    m = {k: np.zeros(D.SHAPES[k]) for k in D.ORDER}
    v = {k: np.zeros(D.SHAPES[k]) for k in D.ORDER}
    D.adam_step(ref, D.flat_to_params(g1.astype(np.float64)), m, v, 1)
    r1 = D.params_to_flat(ref)
    err1 = np.abs(p1 - r1).max()
    ref = D.flat_to_params(p1.astype(np.float64))
    D.adam_step(ref, D.flat_to_params(g2.astype(np.float64)), m, v, 2)
    r2 = D.params_to_flat(ref)
    err2 = np.abs(p2 - r2).max()
    print(f"[denoiser] adam max abs err vs fp64: step 1 {err1:.2e}, step 2 {err2:.2e}", flush=True)
    assert np.allclose(p1, r1, rtol=1e-6, atol=1e-8) and np.allclose(p2, r2, rtol=1e-6, atol=1e-8)
    assert np.abs(p1 - p0).max() > 5e-4  # the step moved the weights (lr 1e-3)
    vol.close()


def test_update_rewrites_packed_weights(gpu):
    """The reduction's Adam step rewrites the packed bf16 weights of every
    parameter it updates: after two training steps the network predicts
    exactly what a fresh denoiser given the updated parameters (packed by
    dnPackKernel) predicts, bit for bit."""
    w, vol = _scene_volume(seed=7, frames=3)
    keys = O.sorted_keys(vol.block_keys())
    obs = (vol.download_blocks(keys)["weight"] > 0).reshape(len(keys), -1).sum(1)
    pick = keys[np.argsort(-obs)[:12]]
    p0 = D.params_to_flat(D.init_params(3)).astype(np.float32)
    with vhm.Denoiser(64, 16) as a, vhm.Denoiser(64, 16) as b:
        a.set_params(p0)
        a.append(vol, pick, vol)
        for _ in range(2):
            a.train_step(0, len(pick), 1e-2)
        p = a.get_params()
        assert np.abs(p - p0).max() > 1e-3
        b.set_params(p)
        b.append(vol, pick, vol)
        ya, yb = a.forward(0, len(pick)), b.forward(0, len(pick))
    assert np.array_equal(ya.view(np.uint32), yb.view(np.uint32))
    vol.close()


def test_apply_kernel_matches_numpy_statement(gpu):
    """dnApplyKernel against the oracle's apply_blocks, bit for bit: the
    blocks after write-back, cells replaced and blocks kept plain, for a
    threshold that keeps about half the blocks plain and then for a
    weight / delta filter on the already written volume."""
    w, vol = _scene_volume(seed=5, frames=3)
    keys = O.sorted_keys(vol.block_keys())
    obs = (vol.download_blocks(keys)["weight"] > 0).reshape(len(keys), -1).sum(1)
    pick = keys[np.argsort(-obs)[:16]]
    p = D.init_params(2)
    p["W3"] *= 30.0
    with vhm.Denoiser(64, 32) as dn:
        dn.set_params(D.params_to_flat(p).astype(np.float32))
        dn.append(vol, pick, vol)
        y = dn.forward(0, len(pick))
        x, _, _ = dn.samples(0, len(pick))
        xin = D.bf16_to_float(x)[:, 1:9, 1:9, 1:9, :].reshape(len(pick), 512, 16)
        means = [np.abs(y[b] - xin[b, :, :6])[xin[b, :, 6:12] > 0].mean() for b in range(len(pick))]
        valid = xin[..., 6:12] > 0
        med_delta = float(np.median(np.abs(y - xin[..., :6])[valid]))
        med_w = float(np.median(xin[..., 6:12][valid]))  # weight / max_weight of observed cells
        settings = [(0.0, 10.0, float(np.median(means))), (med_w, med_delta, 10.0)]
        for min_w, max_delta, bmd in settings:
            before = vol.download_blocks(pick)
            want, rep, kept = D.apply_blocks(before, pick, y, x, w.voxel_size, w.truncation, 100.0, min_w, max_delta,
                                             bmd)
            r = dn.apply(vol, min_weight=min_w, max_delta=max_delta, block_max_delta=bmd)
            got = vol.download_blocks(pick)
            print(f"[denoiser] apply min_w={min_w} max_delta={max_delta} block_max={bmd:.4f}: {r}", flush=True)
            assert (r["cells_replaced"], r["blocks_kept_plain"]) == (rep, kept)
            assert rep > 0
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    vol.close()
