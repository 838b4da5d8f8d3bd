"""CPU ORACLE for the learned block denoiser -- test infrastructure only.

Float64 numpy statement of the per-block 3D-conv denoiser (no reference
counterpart: the reference has no learned component).  Only tests/ and the
training checks import it; the GPU kernels never call it.

Network (channels-last, one 8^3 block with a 1-voxel halo):
  X   (B, 10, 10, 10, 12)  channels 0..5 tsdf_d in [-1, 1], 6..11 w_d in [0, 1]
  h1 = relu(conv3(X, W1) + b1)          valid 3^3:   (B, 8, 8, 8, 32)
  h2 = relu(conv3(pad1(h1), W2) + b2)   same 3^3:    (B, 8, 8, 8, 32)
  r  = conv3(pad1(h2), W3) + b3         same 3^3:    (B, 8, 8, 8, 6)
  y  = X[:, 1:9, 1:9, 1:9, 0:6] + r     residual on the input tsdf
Weights W (kz, ky, kx, Cin, Cout); 43,270 parameters.
Loss: masked mean squared error, L = sum(M (y - T)^2) / max(sum(M), 1).
Adam (Kingma & Ba, 2015) with bias correction.
"""
from __future__ import annotations

import numpy as np

SHAPES = {"W1": (3, 3, 3, 12, 32), "b1": (32,), "W2": (3, 3, 3, 32, 32), "b2": (32,),
          "W3": (3, 3, 3, 32, 6), "b3": (6,)}
ORDER = ["W1", "b1", "W2", "b2", "W3", "b3"]


def init_params(seed: int = 0) -> dict:
    """He-uniform init for W1/W2, small init for the residual head."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = {}
    for name in ORDER:
        shp = SHAPES[name]
        if name.startswith("b"):
            p[name] = np.zeros(shp)
        else:
            fan_in = 27 * shp[3]
            lim = np.sqrt(6.0 / fan_in) * (0.1 if name == "W3" else 1.0)
            p[name] = rng.uniform(-lim, lim, size=shp)
    return p


def num_params() -> int:
    return int(sum(np.prod(SHAPES[n]) for n in ORDER))


def _patches(x: np.ndarray) -> np.ndarray:
    """(B, D+2, H+2, W+2, C) -> (B, D, H, W, 27, C) in (kz, ky, kx) order."""
    B, D2, H2, W2, C = x.shape
    D, H, W = D2 - 2, H2 - 2, W2 - 2
    out = np.empty((B, D, H, W, 27, C), x.dtype)
    i = 0
    for kz in range(3):
        for ky in range(3):
            for kx in range(3):
                out[:, :, :, :, i, :] = x[:, kz:kz + D, ky:ky + H, kx:kx + W, :]
                i += 1
    return out


def _conv(x, W, b):
    P = _patches(x)
    return np.einsum("bdhwkc,kco->bdhwo", P, W.reshape(27, W.shape[3], W.shape[4])) + b


def _pad1(x):
    return np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))


def forward(p: dict, X: np.ndarray):
    X = np.asarray(X, np.float64)
    a1 = _conv(X, p["W1"], p["b1"])
    h1 = np.maximum(a1, 0)
    a2 = _conv(_pad1(h1), p["W2"], p["b2"])
    h2 = np.maximum(a2, 0)
    r = _conv(_pad1(h2), p["W3"], p["b3"])
    y = X[:, 1:9, 1:9, 1:9, 0:6] + r
    return y, (X, a1, h1, a2, h2)


def _conv_backward(x, W, dy):
    """Gradients of y = conv(x, W) + b (x already padded): dx (padded shape), dW, db."""
    P = _patches(x)
    Cin, Cout = W.shape[3], W.shape[4]
    dW = np.einsum("bdhwkc,bdhwo->kco", P, dy).reshape(3, 3, 3, Cin, Cout)
    db = dy.sum(axis=(0, 1, 2, 3))
    dP = np.einsum("bdhwo,kco->bdhwkc", dy, W.reshape(27, Cin, Cout))
    dx = np.zeros_like(x)
    D, H, Wd = dy.shape[1:4]
    i = 0
    for kz in range(3):
        for ky in range(3):
            for kx in range(3):
                dx[:, kz:kz + D, ky:ky + H, kx:kx + Wd, :] += dP[:, :, :, :, i, :]
                i += 1
    return dx, dW, db


def loss_and_grads(p: dict, X, T, M):
    y, (X, a1, h1, a2, h2) = forward(p, X)
    n = max(float(M.sum()), 1.0)
    diff = (y - T) * M
    L = float((diff * (y - T)).sum() / n)
    dy = 2.0 * diff / n
    g = {}
    dh2p, g["W3"], g["b3"] = _conv_backward(_pad1(h2), p["W3"], dy)
    da2 = dh2p[:, 1:9, 1:9, 1:9, :] * (a2 > 0)
    dh1p, g["W2"], g["b2"] = _conv_backward(_pad1(h1), p["W2"], da2)
    da1 = dh1p[:, 1:9, 1:9, 1:9, :] * (a1 > 0)
    _, g["W1"], g["b1"] = _conv_backward(X, p["W1"], da1)
    return L, g


def adam_step(p: dict, g: dict, m: dict, v: dict, t: int, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    for k in ORDER:
        m[k] = b1 * m[k] + (1 - b1) * g[k]
        v[k] = b2 * v[k] + (1 - b2) * g[k] * g[k]
        mh = m[k] / (1 - b1 ** t)
        vh = v[k] / (1 - b2 ** t)
        p[k] = p[k] - lr * mh / (np.sqrt(vh) + eps)


# ----------------------------------------------------------- input features --

def _unpack(packed):
    """unpackNormalOctahedral (voxel_hash_tsdf.h:291-309) in float32, vectorised."""
    f = np.float32
    packed = np.asarray(packed, np.uint32)
    x = ((packed & 0xFFFF).astype(f) / f(65535.0)) * f(2.0) - f(1.0)
    y = (((packed >> 16) & 0xFFFF).astype(f) / f(65535.0)) * f(2.0) - f(1.0)
    z = f(1.0) - np.abs(x) - np.abs(y)
    neg = z < 0
    px = np.where(neg, (f(1) - np.abs(y)) * np.where(x >= 0, f(1), f(-1)), x).astype(f)
    py = np.where(neg, (f(1) - np.abs(x)) * np.where(y >= 0, f(1), f(-1)), y).astype(f)
    z = z.astype(f)
    ln = np.sqrt((px * px + py * py) + z * z).astype(f)
    ok = ln > f(1e-8)
    safe = np.where(ok, ln, f(1))
    return np.stack([np.where(ok, px / safe, f(0)), np.where(ok, py / safe, f(0)), np.where(ok, z / safe, f(0))], -1)


def bf16_round(x: np.ndarray) -> np.ndarray:
    """float32 -> bf16 (round to nearest even) -> float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def block_features(blocks: dict, key, voxel: float, trunc: float, max_w: float = 100.0) -> np.ndarray:
    """Denoiser input of one block (10, 10, 10, 16) as float32 (before bf16),
    from blocks {key tuple: (512, 6, 3) structured voxel array with fields
    weight, plane_offset, packed_normal}.  Per direction: the heaviest slot
    (first on ties), tsdf = clamp((n . centre - offset) / trunc, -1, 1) with
    float32 arithmetic in the kernel's order, weight / max_w."""
    f = np.float32
    out = np.zeros((10, 10, 10, 16), f)
    inv_t, inv_w = f(1.0) / f(trunc), f(1.0) / f(max_w)
    kx, ky, kz = key
    for zz in range(10):
        for yy in range(10):
            for xx in range(10):
                x, y, z = xx - 1, yy - 1, zz - 1
                nb = (kx + (x // 8), ky + (y // 8), kz + (z // 8))
                blk = blocks.get(nb)
                if blk is None:
                    continue
                vi = (x % 8) + 8 * ((y % 8) + 8 * (z % 8))
                centre = np.array([(f(kx * 8 + x) + f(0.5)) * f(voxel), (f(ky * 8 + y) + f(0.5)) * f(voxel),
                                   (f(kz * 8 + z) + f(0.5)) * f(voxel)], f)
                for d in range(6):
                    w = blk[vi, d]["weight"]
                    k = -1
                    best = f(0)
                    for s in range(3):
                        if w[s] > best:
                            best, k = w[s], s
                    if k < 0:
                        continue
                    n = _unpack(blk[vi, d]["packed_normal"][k])
                    sd = f(f(f(n[0] * centre[0]) + f(n[1] * centre[1])) + f(n[2] * centre[2])) - \
                        f(blk[vi, d]["plane_offset"][k])
                    out[zz, yy, xx, d] = min(max(f(sd * inv_t), f(-1)), f(1))
                    out[zz, yy, xx, 6 + d] = f(best * inv_w)
    return out


def bf16_to_float(bits: np.ndarray) -> np.ndarray:
    """bf16 bit patterns (uint16) -> float32."""
    return (np.asarray(bits, np.uint32) << 16).view(np.float32)


def _dot3(n, c):
    """(n.x c.x + n.y c.y) + n.z c.z, each operation rounded to float32."""
    f = np.float32
    return f(f(f(n[..., 0] * c[..., 0]) + f(n[..., 1] * c[..., 1])) + f(n[..., 2] * c[..., 2]))


def apply_blocks(blocks: np.ndarray, keys: np.ndarray, Y: np.ndarray, X: np.ndarray, voxel: float, trunc: float,
                 max_w: float, min_w: float, max_delta: float, block_max_delta: float):
    """Write-back of denoised tsdf with the per-block confidence fallback, in
    float32 in the kernel's operation order (dnApplyKernel).

    blocks: (n, 512, 6, 3) Voxel records of the blocks (AoS, as
    vhm_download_blocks returns them); Y: (n, 512, 6) predictions; X:
    (n, 10, 10, 10, 16) the inputs as bf16 bits.  Per block, the mean of
    |y - x| over the cells (voxel, direction) whose input weight channel is
    > 0 is formed as 256 partial sums (partial t takes voxels t and t + 256,
    directions in order) added pairwise in a tree; a block with no such cell
    or a mean above block_max_delta keeps plain fusion.  Otherwise every cell
    whose heaviest slot (first on ties) has weight / max_w >= min_w and
    |y - x| <= max_delta gets that slot's plane_offset = n . centre -
    clamp(y, -1, 1) * trunc, so that the plane's signed distance at the voxel
    centre is the prediction in truncation units (planeSignedDistance,
    voxel_hash_tsdf.h:336-339).  Returns (new blocks, cells replaced,
    blocks kept plain)."""
    f = np.float32
    out = blocks.copy()
    xin = bf16_to_float(X).reshape(len(keys), 10, 10, 10, 16)[:, 1:9, 1:9, 1:9].reshape(len(keys), 512, 16)
    inv_w = f(1.0) / f(max_w)
    vi = np.arange(512)
    lx, ly, lz = vi & 7, (vi >> 3) & 7, vi >> 6
    replaced = kept = 0
    for b, key in enumerate(np.asarray(keys, np.int64)):
        y = np.asarray(Y[b], f)
        x = xin[b, :, :6]
        delta = np.abs(f(1) * (y - x)).astype(f)
        valid = xin[b, :, 6:12] > 0
        part = np.zeros(256, f)
        cnt = np.zeros(256, np.int64)
        for half in (0, 256):
            for d in range(6):
                m = valid[half:half + 256, d]
                part = np.where(m, (part + delta[half:half + 256, d]).astype(f), part)
                cnt += m
        off = 128
        while off > 0:
            part[:off] = (part[:off] + part[off:2 * off]).astype(f)
            cnt[:off] += cnt[off:2 * off]
            off >>= 1
        if cnt[0] == 0 or f(part[0] / f(cnt[0])) > f(block_max_delta):
            kept += 1
            continue
        blk = out[b]
        centre = np.stack([(f(1) * (key[0] * 8 + lx).astype(f) + f(0.5)) * f(voxel),
                           (f(1) * (key[1] * 8 + ly).astype(f) + f(0.5)) * f(voxel),
                           (f(1) * (key[2] * 8 + lz).astype(f) + f(0.5)) * f(voxel)], -1).astype(f)
        for d in range(6):
            w = blk["weight"][:, d, :]
            bw = np.zeros(512, f)
            bk = np.full(512, -1)
            for k in range(3):  # heaviest slot, the first on ties
                better = w[:, k] > bw
                bw = np.where(better, w[:, k], bw)
                bk = np.where(better, k, bk)
            take = (bk >= 0) & ~((bw * inv_w).astype(f) < f(min_w)) & ~(delta[:, d] > f(max_delta))
            if not take.any():
                continue
            idx = np.flatnonzero(take)
            slot = bk[idx]
            n = _unpack(blk["packed_normal"][idx, d, slot])
            yc = np.clip(y[idx, d], f(-1), f(1)).astype(f)
            po = (_dot3(n, centre[idx]) - (yc * f(trunc)).astype(f)).astype(f)
            blk["plane_offset"][idx, d, slot] = po
            replaced += len(idx)
    return out, replaced, kept


def params_to_flat(p: dict) -> np.ndarray:
    return np.concatenate([p[k].reshape(-1) for k in ORDER]).astype(np.float64)


def flat_to_params(v: np.ndarray) -> dict:
    out, o = {}, 0
    for k in ORDER:
        n = int(np.prod(SHAPES[k]))
        out[k] = np.asarray(v[o:o + n], np.float64).reshape(SHAPES[k])
        o += n
    return out
