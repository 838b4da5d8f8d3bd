"""Reconstruction quality against the analytic scene.

* Chamfer distance (mm) = (accuracy + completeness) / 2.
  * accuracy: mean over mesh vertices of |SDF(v)|. SDF is the scene's exact
    signed distance, whose magnitude is the Euclidean distance to the
    surface outside the primitives.
  * completeness: mean over ground-truth surface samples of the distance to
    the nearest mesh vertex. The samples are noise-free depth pixels
    back-projected from the trajectory's frames, so only surfaces the camera
    saw are counted.
* TSDF L1 (mm). For every observed voxel, the plane signed distance of its
  heaviest slot is evaluated at the voxel centre (planeSignedDistance,
  voxel_hash_tsdf.h:336-339). It is compared with the scene SDF at the same
  centre, both clamped to +-truncation.

This is analysis code, not a hot path. It runs on the host with numpy and
scipy's k-d tree.
"""
from __future__ import annotations

import numpy as np

from . import VOXEL_DTYPE


def unpack_octahedral(packed: np.ndarray) -> np.ndarray:
    """unpackNormalOctahedral (voxel_hash_tsdf.h:291-309), vectorised in float32."""
    packed = np.asarray(packed, np.uint32)
    f = np.float32
    x = (((packed & 0xFFFF).astype(f) / f(65535.0)) * f(2.0) - f(1.0)).astype(f)
    y = ((((packed >> 16) & 0xFFFF).astype(f) / f(65535.0)) * f(2.0) - f(1.0)).astype(f)
    z = (f(1.0) - np.abs(x) - np.abs(y)).astype(f)
    neg = z < 0
    sx = np.where(x >= 0, f(1), f(-1))
    sy = np.where(y >= 0, f(1), f(-1))
    px = np.where(neg, (f(1) - np.abs(y)) * sx, x)
    py = np.where(neg, (f(1) - np.abs(x)) * sy, y)
    n = np.stack([px, py, z], -1).astype(f)
    ln = np.sqrt((n * n).sum(-1, keepdims=True)).astype(f)
    return np.where(ln > 1e-8, n / np.maximum(ln, f(1e-30)), f(0)).astype(f)


def gt_surface_points(scene, K, poses, stride: int = 8, max_points: int = 400_000, seed: int = 0) -> np.ndarray:
    """Noise-free depth of every `stride`-th pixel of the given frames, back-projected to the world."""
    from . import synth
    try:  # the GPU ray caster when a device is present (same pixels as the host one)
        import torch
        if not torch.cuda.is_available():
            raise ImportError
        out = torch.empty((len(poses), K.height, K.width), dtype=torch.int16, device="cuda:0")
        synth.render(scene, K, poses, None, 0, out)
        depth = out.cpu().numpy().view(np.uint16).astype(np.float64) * 1e-3
    except ImportError:
        depth = synth.render(scene, K, poses, None).astype(np.float64) * 1e-3
    pts = []
    vv, uu = np.mgrid[0:K.height:stride, 0:K.width:stride]
    for i, p in enumerate(poses):
        d = depth[i][vv, uu]
        ok = d > 0
        x = (uu[ok] - K.cx) / K.fx * d[ok]
        y = (vv[ok] - K.cy) / K.fy * d[ok]
        pc = np.stack([x, y, d[ok]], -1)
        R = np.array(p.R, np.float64).reshape(3, 3)
        pts.append(pc @ R.T + np.array(p.t, np.float64))
    pts = np.concatenate(pts) if pts else np.zeros((0, 3))
    if len(pts) > max_points:
        pts = pts[np.random.default_rng(seed).choice(len(pts), max_points, replace=False)]
    return pts


def _unique_vertices(tris32: np.ndarray) -> np.ndarray:
    """Distinct vertices (exact float32 bit patterns) of a triangle soup, as float64."""
    v = np.ascontiguousarray(tris32, np.float32).reshape(-1, 3)
    b = v.view(np.uint32)
    k1 = (b[:, 0].astype(np.uint64) << np.uint64(32)) | b[:, 1].astype(np.uint64)
    k2 = b[:, 2]
    order = np.lexsort((k2, k1))
    s1, s2 = k1[order], k2[order]
    keep = np.ones(len(order), bool)
    keep[1:] = (s1[1:] != s1[:-1]) | (s2[1:] != s2[:-1])
    return v[order[keep]].astype(np.float64)


def _nearest_sq_gpu(verts, pts, cell=0.02, chunk=8192):
    """Exact squared distance from each point to its nearest vertex, on the
    GPU: vertices bucketed in `cell` cubes, each point searches its 27
    neighbouring cells; points whose best candidate is farther than `cell`
    (where a vertex outside the 27 cells could be nearer) are resolved by
    brute force over all vertices."""
    import torch
    dev = torch.device("cuda:0")
    v = torch.as_tensor(verts, dtype=torch.float32, device=dev)
    g = torch.as_tensor(pts, dtype=torch.float32, device=dev)
    lo = torch.minimum(v.min(0).values, g.min(0).values) - 2 * cell
    hi = torch.maximum(v.max(0).values, g.max(0).values) + 2 * cell
    dims = (torch.floor((hi - lo) / cell).long() + 1)
    def code(c):
        return (c[:, 0] * dims[1] + c[:, 1]) * dims[2] + c[:, 2]
    vcode = code(torch.floor((v - lo) / cell).long())
    vcode, order = torch.sort(vcode)
    vs = v[order]
    best = torch.full((len(g),), float("inf"), device=dev)
    offs = torch.tensor([(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)], device=dev)
    gc = torch.floor((g - lo) / cell).long()
    for s in range(0, len(g), chunk):
        e = min(s + chunk, len(g))
        nc = (gc[s:e, None, :] + offs[None]).reshape(-1, 3)
        q = code(nc)
        left = torch.searchsorted(vcode, q)
        cnt = torch.searchsorted(vcode, q, right=True) - left
        tot = int(cnt.sum())
        if tot == 0:
            continue
        pair_q = torch.repeat_interleave(torch.arange(len(q), device=dev), cnt)
        start = torch.cumsum(cnt, 0) - cnt
        vidx = left[pair_q] + (torch.arange(tot, device=dev) - start[pair_q])
        gidx = s + pair_q // 27
        d = ((g[gidx] - vs[vidx]) ** 2).sum(1)
        best.scatter_reduce_(0, gidx, d, reduce="amin")
    far = torch.nonzero(best > cell * cell).flatten()
    for i in far.tolist():
        best[i] = ((vs - g[i]) ** 2).sum(1).min()
    return best.cpu().numpy().astype(np.float64)


def chamfer_mm(triangles: np.ndarray, scene, gt_points: np.ndarray, max_accuracy_samples: int = 500_000,
               seed: int = 0) -> dict:
    """accuracy over (a seeded random subset of at most max_accuracy_samples)
    unique mesh vertices; completeness over all gt_points against all unique
    vertices (exact nearest neighbours: on the GPU when one is present,
    else a k-d tree)."""
    verts = np.asarray(triangles).reshape(-1, 3)
    if len(verts) == 0 or len(gt_points) == 0:
        return {"chamfer_mm": float("inf"), "accuracy_mm": float("inf"), "completeness_mm": float("inf"),
                "vertices": 0}
    verts = _unique_vertices(verts)
    sample = verts
    if len(verts) > max_accuracy_samples:
        sample = verts[np.random.default_rng(seed).choice(len(verts), max_accuracy_samples, replace=False)]
    acc = np.abs(scene.sdf(sample)).mean() * 1e3
    try:
        import torch
        if not torch.cuda.is_available():
            raise ImportError
        comp = np.sqrt(_nearest_sq_gpu(verts, gt_points)).mean() * 1e3
    except ImportError:
        from scipy.spatial import cKDTree
        comp_d, _ = cKDTree(verts, balanced_tree=False, compact_nodes=False).query(gt_points, k=1, workers=-1)
        comp = comp_d.mean() * 1e3
    return {"chamfer_mm": float((acc + comp) / 2), "accuracy_mm": float(acc), "completeness_mm": float(comp),
            "vertices": int(len(verts))}


def tsdf_l1_mm(keys: np.ndarray, blocks: np.ndarray, scene, voxel_size: float, truncation: float) -> dict:
    """blocks: (n, 512, 6, 3) VOXEL_DTYPE as Volume.download_blocks returns."""
    assert blocks.dtype == VOXEL_DTYPE
    w = blocks["weight"].reshape(len(blocks), 512, 18)
    best = w.argmax(-1)
    obs = np.take_along_axis(w, best[..., None], -1)[..., 0] > 0
    if not obs.any():
        return {"tsdf_l1_mm": float("inf"), "observed": 0}
    pn = np.take_along_axis(blocks["packed_normal"].reshape(len(blocks), 512, 18), best[..., None], -1)[..., 0]
    po = np.take_along_axis(blocks["plane_offset"].reshape(len(blocks), 512, 18), best[..., None], -1)[..., 0]
    li = np.arange(512)
    local = np.stack([li & 7, (li >> 3) & 7, li >> 6], -1)
    vox = keys[:, None, :].astype(np.int64) * 8 + local[None]
    centre = ((vox.astype(np.float32) + np.float32(0.5)) * np.float32(voxel_size)).astype(np.float32)
    n = unpack_octahedral(pn)
    sd = (n * centre).sum(-1) - po
    gt = scene.sdf(centre[obs].astype(np.float64))
    t = truncation
    err = np.abs(np.clip(sd[obs], -t, t) - np.clip(gt, -t, t))
    return {"tsdf_l1_mm": float(err.mean() * 1e3), "observed": int(obs.sum())}
