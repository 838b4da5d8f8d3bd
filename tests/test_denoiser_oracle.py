"""Pins of the FP64 denoiser oracle (oracle/denoiser_ref.py): parameter
count, the convolution against a direct loop, and every gradient against
central finite differences."""
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import denoiser_ref as D  # noqa: E402


def test_parameter_count():
    assert D.num_params() == 43270


def test_conv_matches_direct_loop():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(1, 5, 5, 5, 3))
    W = rng.normal(size=(3, 3, 3, 3, 2))
    b = rng.normal(size=2)
    y = D._conv(x, W, b)
    for (z, yy, xx) in [(0, 0, 0), (2, 1, 0), (1, 2, 2)]:
        ref = b.copy()
        for kz in range(3):
            for ky in range(3):
                for kx in range(3):
                    ref += x[0, z + kz, yy + ky, xx + kx] @ W[kz, ky, kx]
        assert np.allclose(y[0, z, yy, xx], ref)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    p = D.init_params(3)
    p["W3"] *= 10.0  # make the head matter
    for k in ("b1", "b2", "b3"):
        p[k] = rng.normal(scale=0.05, size=p[k].shape)
    X = rng.uniform(-1, 1, size=(2, 10, 10, 10, 12))
    T = rng.uniform(-1, 1, size=(2, 8, 8, 8, 6))
    M = (rng.uniform(size=T.shape) > 0.3).astype(np.float64)
    L, g = D.loss_and_grads(p, X, T, M)
    eps = 1e-6
    for name in D.ORDER:
        flat = p[name].reshape(-1)
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + eps
            Lp, _ = D.loss_and_grads(p, X, T, M)
            flat[idx] = old - eps
            Lm, _ = D.loss_and_grads(p, X, T, M)
            flat[idx] = old
            fd = (Lp - Lm) / (2 * eps)
            an = g[name].reshape(-1)[idx]
            assert abs(fd - an) <= 1e-6 + 1e-4 * abs(fd), (name, idx, fd, an)


def test_adam_step_matches_torch_adam_fp64():
    """adam_step equals torch.optim.Adam (same hyper-parameters, bias
    correction, eps outside the square root) in float64 over several steps."""
    import torch
    rng = np.random.default_rng(4)
    p = {k: rng.normal(size=D.SHAPES[k]) for k in D.ORDER}
    tp = [torch.tensor(p[k].copy(), dtype=torch.float64, requires_grad=True) for k in D.ORDER]
    opt = torch.optim.Adam(tp, lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    m = {k: np.zeros(D.SHAPES[k]) for k in D.ORDER}
    v = {k: np.zeros(D.SHAPES[k]) for k in D.ORDER}
    for t in range(1, 6):
        g = {k: rng.normal(scale=10.0 ** rng.uniform(-4, 1), size=D.SHAPES[k]) for k in D.ORDER}
        D.adam_step(p, g, m, v, t)
        for k, x in zip(D.ORDER, tp):
            x.grad = torch.tensor(g[k], dtype=torch.float64)
        opt.step()
        for k, x in zip(D.ORDER, tp):
            assert np.allclose(p[k], x.detach().numpy(), rtol=1e-13, atol=1e-15), (t, k)


def _synthetic_blocks(n, seed):
    """n blocks of random Voxel records (any 32-bit octahedral code is a valid
    one), their keys, inputs X (bf16 bits) and predictions Y."""
    from voxelhashmt_amd import VOXEL_DTYPE
    rng = np.random.default_rng(seed)
    blocks = np.zeros((n, 512, 6, 3), VOXEL_DTYPE)
    w = rng.uniform(0, 100, size=blocks.shape).astype(np.float32)
    w[rng.uniform(size=w.shape) < 0.4] = 0
    blocks["weight"] = w
    blocks["plane_offset"] = rng.uniform(-3, 3, size=blocks.shape).astype(np.float32)
    blocks["packed_normal"] = rng.integers(0, 2 ** 32, size=blocks.shape, dtype=np.uint64).astype(np.uint32)
    keys = rng.integers(-50, 50, size=(n, 3)).astype(np.int32)
    xf = np.zeros((n, 10, 10, 10, 16), np.float32)
    xf[..., :6] = rng.uniform(-1, 1, size=(n, 10, 10, 10, 6))
    xf[..., 6:12] = np.where(rng.uniform(size=(n, 10, 10, 10, 6)) < 0.5, rng.uniform(0, 1, (n, 10, 10, 10, 6)), 0)
    X = (D.bf16_round(xf).view(np.uint32) >> 16).astype(np.uint16)
    interior = D.bf16_to_float(X)[:, 1:9, 1:9, 1:9, :6].reshape(n, 512, 6)
    Y = (interior + rng.normal(scale=0.05, size=interior.shape)).astype(np.float32)
    return blocks, keys, X, Y


def test_apply_blocks_fallback_and_write_back():
    """The write-back oracle against what its rules fix: a negative block
    threshold keeps every block; otherwise each replaced cell is exactly the
    heaviest slot of a cell passing the weight and delta tests, its plane's
    signed distance at the voxel centre is clamp(y) * trunc (to rounding),
    and nothing else changes; the block mean decides the fallback."""
    voxel, trunc, max_w = 0.01, 0.04, 100.0
    blocks, keys, X, Y = _synthetic_blocks(3, 11)
    out, rep, kept = D.apply_blocks(blocks, keys, Y, X, voxel, trunc, max_w, 0.0, 0.5, -1.0)
    assert (rep, kept) == (0, 3) and out.tobytes() == blocks.tobytes()
    min_w, max_delta = 0.2, 0.08
    out, rep, kept = D.apply_blocks(blocks, keys, Y, X, voxel, trunc, max_w, min_w, max_delta, 10.0)
    assert kept == 0 and rep > 0
    changed = out["plane_offset"] != blocks["plane_offset"]
    for f in ("weight", "surface_coord", "packed_normal"):
        assert np.array_equal(out[f].view(np.uint32), blocks[f].view(np.uint32))
    xin = D.bf16_to_float(X)[:, 1:9, 1:9, 1:9, :6].reshape(3, 512, 6)
    w = blocks["weight"]
    heavy = np.argmax(w, axis=-1)  # first maximum
    eligible = (w.max(-1) > 0) & (w.max(-1) / max_w >= min_w) & (np.abs(Y - xin) <= max_delta)
    want = np.zeros_like(changed)
    b, v, d = np.nonzero(eligible)
    want[b, v, d, heavy[b, v, d]] = True
    # a replaced offset may equal the old one by chance; every change is a wanted cell
    assert not (changed & ~want).any() and changed.sum() >= want.sum() - 2
    assert rep == int(want.sum())
    vi = np.arange(512)
    for b_, v_, d_ in zip(b, v, d):
        s = heavy[b_, v_, d_]
        c = (np.array([keys[b_][0] * 8 + (v_ & 7), keys[b_][1] * 8 + ((v_ >> 3) & 7),
                       keys[b_][2] * 8 + (v_ >> 6)], np.float64) + 0.5) * voxel
        n = D._unpack(np.array([blocks["packed_normal"][b_, v_, d_, s]]))[0].astype(np.float64)
        sd = n @ c - float(out["plane_offset"][b_, v_, d_, s])
        assert abs(sd - np.clip(Y[b_, v_, d_], -1, 1) * trunc) < 1e-5, (b_, v_, d_)
    del vi
    # the block mean: just above it the block stays plain, at it the block is written
    valid = D.bf16_to_float(X)[0, 1:9, 1:9, 1:9, 6:12].reshape(512, 6) > 0
    mean = float(np.abs(Y[0] - xin[0])[valid].mean())
    _, _, kept_hi = D.apply_blocks(blocks[:1], keys[:1], Y[:1], X[:1], voxel, trunc, max_w, 0, 1, mean * 0.999)
    _, _, kept_lo = D.apply_blocks(blocks[:1], keys[:1], Y[:1], X[:1], voxel, trunc, max_w, 0, 1, mean * 1.001)
    assert (kept_hi, kept_lo) == (1, 0)
