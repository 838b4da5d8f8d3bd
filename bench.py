"""Benchmark: plain fusion at scale on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload plain_scale]

A step is one depth frame through the fusion path (bilateral filter ->
block allocation -> directional TSDF integration) of the 1280x720 plain-
at-scale workload: 8x8x3 m room with 150 seeded primitives, 500 frames on a
random walk (0.02 m / +-1 deg per frame), voxel 5 mm, truncation 2 cm, six
directional bins x 3 fp32 slots per voxel, 5x5 bilateral (sigma_s 4.5 px,
sigma_r 30 mm), incremental marching cubes over the dirty blocks every 10
frames (triangles written into a device buffer).  Frames are rendered on the
GPU and resident in HBM before the timed region; W warm-up frames are
integrated untimed, then K frames are timed (W + K <= 500, in sequence).
After the timed region one full marching-cubes extraction is timed
separately.

value = voxel updates / s over the timed frames (an update is a voxel that
reaches the layer update with weight > 0 -- voxel_hash_tsdf.h:1031-1055).
roofline: the integration pass (integrateKernel + integrateDeferredKernel,
the latter taking the voxels that project within 2 px of the top/left image
border), algorithmic bytes = 96 B per update (3 slots x 16 B read +
written), HIP events on the library's stream around the pair; avg_launch_ms
is per frame.
cpu_baseline: frame 0 of the same workload on one host core: allocation and
integration by the reference's own CPU code (oracle/_ref, compiled from its
sources) after the restated filter, or, if that build is absent, the
restated CPU path throughout.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
BYTES_PER_UPDATE = 96
# Marching cubes' algorithmic bytes: a staged (block, direction) item reads
# its 9x9x9 corner region's 3 slots (48 B per voxel); a triangle writes 36 B.
MC_BYTES_PER_ITEM = 48 * 729
MC_BYTES_PER_TRIANGLE = 36
# VALU issue peak: one wave64 instruction per CU per cycle (4 SIMDs of 16
# lanes), 256 CUs at 2.4 GHz.
VALU_PEAK_GINST = 256 * 2.4


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(w, frame0_u16, gpu_updates_frame0: int) -> dict | None:
    """Times the CPU path on frame 0 (filter + allocation + integration), one
    core.  Where the reference's own CPU path is built (oracle/_ref, from its
    sources), allocation and integration are its compiled code
    (VoxelHashTSDF::integrate, voxel_hash_tsdf.h:125-131) on the frame
    filtered by the restated filter: the reference build has its filter
    constants compiled in (reconstruction_backend.cpp:26-32) and cannot run
    the workload's (sigma_s 4.5 px, sigma_r 30 mm).  Otherwise the restated
    path (the oracle) does all three."""
    import numpy as np
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as O
    import voxelhashmt_amd as vhm
    depth = frame0_u16.astype(np.float32) * np.float32(0.001)
    pose = vhm.pose_array(w.poses[0])
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except (AttributeError, OSError):
        pass
    vp = vhm.default_params(w.voxel_size, w.truncation, filter_sigma_spatial=w.filter_sigma_spatial,
                            filter_sigma_range=w.filter_sigma_range)
    op = O.oracle_params(vp)
    t0 = time.perf_counter()
    if O.ref_available():
        filtered = O.oracle_filter(op, depth, w.K)
        t1 = time.perf_counter()
        rv = O.RefVolume(w.voxel_size, w.truncation)
        rv.integrate(filtered, w.K, pose, apply_filter=False)
        t2 = time.perf_counter()
        updates, kind = gpu_updates_frame0, "reference"
        how = (f"filter (restated, config sigmas) {t1 - t0:.1f} s + allocation and integration by the reference's "
               f"compiled CPU code {t2 - t1:.1f} s")
    else:
        ov = O.OracleVolume(op)
        ov.integrate(depth, w.K, pose)
        t2 = time.perf_counter()
        updates, kind = ov.updates(), "port"
        how = "filter + allocation + integration, restated CPU path"
    dt = t2 - t0
    return {"value": updates / dt, "unit": "voxel updates/s", "cores": 1, "kind": kind,
            "sample": f"frame 0 of {w.name} ({w.K.width}x{w.K.height}): {updates} updates in {dt:.1f} s on 1 host "
                      f"core (of {os.cpu_count()}); {how}",
            "seconds": dt}


def calibrate_copy(nbytes: int = 1 << 30, reps: int = 5):
    """A plain device-to-device copy of known size: read nbytes, write nbytes
    per launch.  Its FETCH_SIZE / WRITE_SIZE under rocprofv3 --pmc give the
    counters' scale (KiB per byte moved)."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read_per_launch": nbytes, "copy_bytes_written_per_launch": nbytes,
                      "launches": reps}))


def bench_point_cloud(args):
    """integratePointCloud at scale: the valid pixels of the 1280x720
    plain-at-scale frames, back-projected to camera space on the GPU (about
    0.9 M points a frame), integrated one cloud per step into one volume
    (voxel 5 mm, truncation 2 cm: 19 samples per ray).  The clouds are
    resident in HBM before the timed region."""
    import numpy as np
    import torch
    import voxelhashmt_amd as vhm
    from voxelhashmt_amd import synth

    W = min(args.warmup, 10)
    K = min(args.steps, 40)
    w = synth.plain_scale(n_frames=W + K)
    ok, msg = vhm.device_available()
    if not ok:
        raise SystemExit(f"no GPU: {msg}")
    dev = torch.device("cuda:0")
    frames = torch.empty((W + K, w.K.height, w.K.width), dtype=torch.int16, device=dev)
    synth.render(w.scene, w.K, w.poses, w.noise, 0, frames)
    v, u = torch.meshgrid(torch.arange(w.K.height, device=dev, dtype=torch.float32),
                          torch.arange(w.K.width, device=dev, dtype=torch.float32), indexing="ij")
    clouds = []
    for i in range(W + K):
        d = frames[i].to(torch.int32).to(torch.float32) * 0.001
        m = d > 0
        z = d[m]
        clouds.append(torch.stack([(u[m] - w.K.cx) * z / w.K.fx, (v[m] - w.K.cy) * z / w.K.fy, z], 1).contiguous())
    torch.cuda.synchronize()
    vp = vhm.default_params(w.voxel_size, w.truncation, hash_log2_buckets=w.hash_log2_buckets,
                            max_blocks=args.max_blocks or w.max_blocks, max_width=w.K.width, max_height=w.K.height)
    vol = vhm.Volume(vp)
    for i in range(W):
        vol.integrate_point_cloud(clouds[i], w.poses[i])
    s0 = vol.stats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(W, W + K):
        vol.integrate_point_cloud(clouds[i], w.poses[i])
    vol.synchronize()
    elapsed = time.perf_counter() - t0
    points = sum(int(c.shape[0]) for c in clouds[W:])
    updates = vol.stats()["updates"] - s0["updates"]
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle_lib as O
        sample = clouds[W][:20000].cpu().numpy()
        ov = O.OracleVolume(O.oracle_params(vp))
        t = time.perf_counter()
        ov.integrate_point_cloud(sample, vhm.pose_array(w.poses[W]))
        cpu_s = time.perf_counter() - t
        cpu = {"value": len(sample) / cpu_s, "unit": "points/s", "cores": 1, "kind": "port",
               "sample": f"first {len(sample)} points of the first timed cloud, empty volume"}
    out = {
        "metric": "point-cloud points integrated/s", "value": points / elapsed, "unit": "points/s", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": elapsed * 1e3 / K, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (plain-at-scale frames back-projected on the GPU, resident in HBM)",
        "config": {"workload": "point_cloud_1280x720", "points_per_step": points / K, "voxel_size_m": w.voxel_size,
                   "truncation_m": w.truncation, "samples_per_ray": int(2 * w.truncation / max(w.voxel_size / 2, 0.0025)) + 1,
                   "blocks_allocated": vol.block_count()},
        "voxel_updates_per_s": updates / elapsed,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    vol.close()


def bench_batched(args):
    """Batched plain method: 48 seeded rooms (target block counts log-uniform
    in 1k-60k, resolutions from {320x240, 640x480, 1280x720}), 60 frames
    each, voxel 1 cm; one step = one frame of every scene through filter,
    allocation, integration and incremental marching cubes, replayed as one
    captured hipGraph.  Grid sizes come from the autotuner; an untimed full
    pass precedes the timed pass (the sequences are rewound in between)."""
    import numpy as np
    import torch
    import voxelhashmt_amd as vhm
    from voxelhashmt_amd import synth

    ws = synth.batched_scenes()
    n_frames = len(ws[0].poses)
    K = min(args.steps, n_frames)
    ok, msg = vhm.device_available()
    if not ok:
        raise SystemExit(f"no GPU: {msg}")
    log(f"[bench-batched] {msg}; {len(ws)} scenes x {n_frames} frames, timed steps {K}")
    dev = torch.device("cuda:0")
    frames = []
    for w in ws:
        t = torch.empty((n_frames, w.K.height, w.K.width), dtype=torch.int16, device=dev)
        synth.render(w.scene, w.K, w.poses, w.noise, 0, t)
        frames.append(t)
    torch.cuda.synchronize()
    log("[bench-batched] rendered all frames into HBM")
    params = [vhm.default_params(w.voxel_size, w.truncation, filter_sigma_spatial=w.filter_sigma_spatial,
                                 filter_sigma_range=w.filter_sigma_range, hash_log2_buckets=w.hash_log2_buckets,
                                 max_blocks=w.max_blocks, max_width=w.K.width, max_height=w.K.height)
              for w in ws]
    caps = [200 * w.target_blocks + 65536 for w in ws]
    b = vhm.Batch(params)
    for i, w in enumerate(ws):
        b.set_sequence(i, frames[i], w.K, w.poses, tri_capacity=caps[i])
    b.prepare(use_graph=True)
    gi, gm = b.autotune((2, 4, 8, 16))
    log(f"[bench-batched] autotuned grids: integrate {gi}, mesh {gm}")
    ftw, atw, chunk = b.autotune_tiles((32, 16, 64, 32, 16, 64), (16, 32, 64, 16, 32, 64), (0, 16, 64, 0, 16, 64))
    log(f"[bench-batched] autotuned tiles: filter {ftw}x{256 // ftw}, allocation {atw}x{256 // atw}, "
        f"integrator chunk {chunk}")
    b.rewind()
    if args.warmup:
        b.step(n_frames)
        b.synchronize()
        b.rewind()
        log("[bench-batched] warm-up pass done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.step(K)
    b.synchronize()
    elapsed = time.perf_counter() - t0
    updates = sum(v.stats()["updates"] for v in b.volumes)
    blocks = np.array([v.block_count() for v in b.volumes])
    log(f"[bench-batched] timed {K} steps in {elapsed * 1e3:.1f} ms; updates={updates}; "
        f"blocks min/median/max {blocks.min()}/{int(np.median(blocks))}/{blocks.max()}")

    # stage breakdown: the same steps again with direct launches and events
    b.rewind()
    tot = {"front_ms": 0.0, "integrate_ms": 0.0, "mesh_ms": 0.0, "updates": 0, "blocks_remeshed": 0}
    overflow = 0
    for _ in range(K):
        r = b.timed_step()
        for k in tot:
            tot[k] += r[k]
        for i in range(len(ws)):
            if b.mesh_output(i, fetch=False)[1] > caps[i]:
                overflow += 1
    achieved = BYTES_PER_UPDATE * tot["updates"] / max(tot["integrate_ms"] * 1e-3, 1e-12) / 1e9

    # the same scenes and steps without batching: one scene after another
    # through the one-volume path (vhm_integrate_u16 + the incremental mesh
    # into a device buffer after every frame), frames resident in HBM
    b.rewind()
    b.synchronize()
    seq_buf = torch.empty(max(caps) * 9, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i, w in enumerate(ws):
        v = b.volumes[i]
        for step in range(K):
            v.integrate(frames[i][step], w.K, w.poses[step])
            v.extract_mesh_incremental(device_tris=seq_buf)
        v.synchronize()
    torch.cuda.synchronize()
    seq_elapsed = time.perf_counter() - t0
    seq_updates = sum(v.stats()["updates"] for v in b.volumes)
    log(f"[bench-batched] sequential: {K} steps of {len(ws)} scenes one after another in {seq_elapsed * 1e3:.1f} ms")
    assert seq_updates == updates, "the sequential run integrates the same voxel updates"
    del seq_buf

    cpu = None
    if not args.no_cpu_baseline:
        i0 = min(range(len(ws)), key=lambda i: ws[i].K.width * ws[i].K.height)
        f0 = frames[i0][0].cpu().numpy().view(np.uint16)
        with vhm.Volume(params[i0]) as v0:
            v0.integrate(frames[i0][0], ws[i0].K, ws[i0].poses[0])
            u0 = v0.stats()["updates"]
        log("[bench-batched] timing the CPU baseline ...")
        cpu = cpu_baseline(ws[i0], f0, u0)
        cpu.pop("seconds", None)
    out = {
        "metric": "voxel updates/s (% HBM peak)",
        "value": updates / elapsed,
        "unit": "voxel updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": n_frames if args.warmup else 0,
        "ms_per_step": elapsed * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (48 seeded analytic rooms, GPU-rendered uint16 depth resident in HBM)",
        "config": {"workload": "batched_48_rooms", "scenes": len(ws), "frames_per_scene": n_frames,
                   "resolutions": sorted({f"{w.K.width}x{w.K.height}" for w in ws}), "voxel_size_m": 0.01,
                   "truncation_m": 0.04, "blocks_min": int(blocks.min()), "blocks_median": int(np.median(blocks)),
                   "blocks_max": int(blocks.max()), "blocks_total": int(blocks.sum()),
                   "target_blocks_min": min(w.target_blocks for w in ws),
                   "target_blocks_max": max(w.target_blocks for w in ws),
                   "filter_tile": f"{ftw}x{256 // ftw}", "alloc_tile": f"{atw}x{256 // atw}",
                   "integrate_chunk": chunk,
                   "integrate_grid": gi, "mesh_grid": gm, "hipgraph": True},
        "frames_per_s": len(ws) * K / elapsed,
        "stage_ms_direct_launches": {k: round(v, 3) for k, v in tot.items() if k.endswith("_ms")},
        "blocks_remeshed": tot["blocks_remeshed"],
        "mesh_capacity_overflows": overflow,
        "sequential_one_volume": {"ms_per_step": round(seq_elapsed * 1e3 / K, 3),
                                  "speedup_of_batching": round(seq_elapsed / elapsed, 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "integrateBatchKernel+integrateDeferredBatchKernel",
                     "avg_launch_ms": round(tot["integrate_ms"] / K, 4),
                     "updates_per_launch": tot["updates"] / K},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    b.close()


# Algorithmic FLOP of the denoiser per block (three 3x3x3 convolutions on the
# 8^3 interior, 12->32->32->6): forward = 2 * 512 * 27 * (12*32 + 32*32 +
# 32*6); a training step adds the three weight gradients (the forward's
# FLOP again) and the data gradients of layers 3 and 2 (32*6 + 32*32 per
# voxel and tap).
DN_FWD_FLOP = 2 * 512 * 27 * (12 * 32 + 32 * 32 + 32 * 6)
DN_TRAIN_FLOP = 2 * DN_FWD_FLOP + 2 * 512 * 27 * (32 * 6 + 32 * 32)
BF16_DENSE_PEAK_TFLOPS = 2500.0


def bench_denoiser(args):
    """Denoiser training (the learned configuration's step): samples from
    seeded small rooms in the plain-small style (input: 10 noisy frames,
    target: 300 frames with the noise sigma halved), then K training steps
    at batch 4096 (forward, backward, gradient reduction, Adam) over a
    resident sample set, timed with the device idle before and after."""
    import numpy as np
    import torch
    import voxelhashmt_amd as vhm
    import voxelhashmt_amd.learned as L

    ok, msg = vhm.device_available()
    if not ok:
        raise SystemExit(f"no GPU: {msg}")
    batch = 4096
    K = max(1, min(args.steps, 200))
    W = max(1, min(args.warmup, 20))
    dn = vhm.Denoiser(capacity=3 * batch, batch_capacity=batch)
    dn.set_params(L.init_params(0))
    L.build_training_set(dn, range(1000, 1020), blocks_per_scene=1000)
    n = dn.count
    log(f"[bench-denoiser] {msg}; {n} samples, batch {batch}, warmup {W}, steps {K}")
    offs = np.random.default_rng(0).integers(0, n - batch + 1, size=W + K)
    for i in range(W):
        dn.train_step(int(offs[i]), batch, 1e-3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses = [dn.train_step(int(offs[W + i]), batch, 1e-3) for i in range(K)]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    flop = DN_TRAIN_FLOP * batch
    tflops = flop * K / dt / 1e12
    t1 = time.perf_counter()
    dn.forward(0, batch, fetch=False)
    torch.cuda.synchronize()
    fwd_dt = time.perf_counter() - t1
    out = {
        "metric": "denoiser training blocks/s (% MFMA peak)",
        "value": batch * K / dt,
        "unit": "blocks/s",
        "n_gpus": 1, "steps": K, "warmup": W, "ms_per_step": dt * 1e3 / K,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (fp32 accumulate)",
        "data": "synthetic (denoiser samples from 20 seeded plain-small rooms fused on the GPU)",
        "config": {"workload": "denoiser_train_batch4096", "batch": batch, "samples": n,
                   "network": "3x3x3 conv 12->32->32->6, ReLU", "optimizer": "Adam lr 1e-3"},
        "loss_first": losses[0], "loss_last": losses[-1],
        "forward_ms_per_batch": fwd_dt * 1e3,
        "roofline": {"bound": "mfma", "achieved": round(tflops, 2), "peak": BF16_DENSE_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(tflops / BF16_DENSE_PEAK_TFLOPS, 4), "traffic": None,
                     "kernel": "training step (forward, backward, reduce, Adam)",
                     "flop_per_step": flop},
        "cpu_baseline": None,
    }
    print(json.dumps(out))
    dn.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=480)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="plain_scale",
                    choices=["plain_scale", "plain_small", "batched", "point_cloud", "denoiser_train"])
    ap.add_argument("--max-blocks", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mesh", action="store_true")
    ap.add_argument("--time-all-stages", action="store_true",
                    help="also time the filter and allocation launches with HIP events (costs about 0.5 %%)")
    ap.add_argument("--mesh-every", type=int, default=10, help="incremental MC period in frames (0 = off)")
    ap.add_argument("--mesh-tri-capacity", type=int, default=64_000_000,
                    help="device triangle buffer of the incremental calls (a call that overflows it is counted)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy 5 times (HBM counter calibration under rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU only (--gpus must be 1); fusion would shard "
                                   "by block ownership, one volume shard per GPU"}))
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.workload == "batched":
        bench_batched(args)
        return
    if args.workload == "point_cloud":
        bench_point_cloud(args)
        return
    if args.workload == "denoiser_train":
        bench_denoiser(args)
        return

    import numpy as np
    import torch
    import voxelhashmt_amd as vhm
    from voxelhashmt_amd import synth

    w = synth.plain_scale() if args.workload == "plain_scale" else synth.plain_small()
    n_total = len(w.poses)
    W = min(args.warmup, n_total - 1)
    K = min(args.steps, n_total - W)
    ok, msg = vhm.device_available()
    if not ok:
        raise SystemExit(f"no GPU: {msg}")
    log(f"[bench] {msg}; workload {w.name}, warmup {W}, steps {K}")

    dev = torch.device("cuda:0")
    frames = torch.empty((W + K, w.K.height, w.K.width), dtype=torch.int16, device=dev)
    synth.render(w.scene, w.K, w.poses[: W + K], w.noise, 0, frames)
    torch.cuda.synchronize()
    log(f"[bench] rendered {W + K} frames into HBM")

    max_blocks = args.max_blocks or w.max_blocks
    vp = vhm.default_params(w.voxel_size, w.truncation, filter_sigma_spatial=w.filter_sigma_spatial,
                            filter_sigma_range=w.filter_sigma_range, hash_log2_buckets=w.hash_log2_buckets,
                            max_blocks=max_blocks, max_width=w.K.width, max_height=w.K.height)
    vol = vhm.Volume(vp)
    updates_frame0 = None
    for i in range(W):
        vol.integrate(frames[i], w.K, w.poses[i])
        if i == 0:
            updates_frame0 = vol.stats()["updates"]
    vol.synchronize()
    s0 = vol.stats()
    log(f"[bench] warmup done: blocks={vol.block_count()}")

    mesh_every = args.mesh_every
    tri_buf = torch.empty(args.mesh_tri_capacity * 9, dtype=torch.float32, device=dev) if mesh_every else None
    if mesh_every:
        vol.extract_mesh_incremental(device_tris=tri_buf)  # consume the warm-up's dirty blocks
    inc_blocks = inc_tris = inc_calls = inc_overflows = inc_max = 0
    items0 = vol.mesh_items_staged()
    # HIP events around the integration and meshing launches (the two
    # rooflines); each timed stage leaves a few microseconds between launches,
    # so filter and allocation are timed only on request
    timed = vol.TIMING_STAGES if args.time_all_stages else ("integrate", "mesh")
    vol.set_timing(True, stages=timed)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for j, i in enumerate(range(W, W + K)):
        vol.integrate(frames[i], w.K, w.poses[i])
        if mesh_every and (j + 1) % mesh_every == 0:
            nb, nt = vol.extract_mesh_incremental(device_tris=tri_buf)
            inc_blocks += nb
            inc_tris += nt
            inc_calls += 1
            inc_overflows += int(nt > args.mesh_tri_capacity)
            inc_max = max(inc_max, nt)
    vol.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    s1 = vol.stats()
    inc_items = vol.mesh_items_staged() - items0
    kt = vol.kernel_times()
    vol.set_timing(False)
    updates = s1["updates"] - s0["updates"]
    visited = s1["visited"] - s0["visited"]
    blocks = vol.block_count()
    log(f"[bench] timed {K} frames in {elapsed * 1e3:.1f} ms; updates={updates} blocks={blocks}")

    mesh = {}
    if not args.no_mesh:
        vol.set_timing(True)
        t = time.perf_counter()
        tris = vol.extract_mesh()
        mesh_s = time.perf_counter() - t
        mk = vol.kernel_times()["mesh"]
        vol.set_timing(False)
        mesh = {"mesh_ms": mesh_s * 1e3, "mesh_gpu_ms": mk["ms"], "triangles": int(len(tris))}
        log(f"[bench] mesh: {len(tris)} triangles in {mesh_s * 1e3:.1f} ms")

    integ = kt["integrate"]
    avg_ms = integ["ms"] / max(integ["launches"], 1)
    achieved = BYTES_PER_UPDATE * updates / max(integ["ms"] * 1e-3, 1e-12) / 1e9
    # HBM traffic per launch: bytes per voxel update measured with
    # rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (each scaled by a plain copy
    # kernel's reading), times this run's updates per launch
    traffic = None
    tfile = ROOT / "profiles" / "integrateKernel_plain_scale_1280x720_traffic.json"
    if tfile.exists():
        try:
            bpu = json.loads(tfile.read_text())["bytes_per_update"]
            traffic = round(bpu * updates / max(integ["launches"], 1))
        except (KeyError, ValueError):
            traffic = None
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "integrateKernel+integrateDeferredKernel", "avg_launch_ms": round(avg_ms, 4),
                "updates_per_launch": updates / max(integ["launches"], 1)}
    # VALU issue of the same launches: VALU wavefront instructions per voxel
    # update measured with rocprofv3 --pmc SQ_INSTS_VALU SQ_WAVES (the
    # profile's figure), times this run's updates per launch, over the
    # launch time, against one wave64 instruction per CU per cycle
    vfile = ROOT / "profiles" / "integrateKernel_plain_scale_1280x720_valu.json"
    if vfile.exists():
        try:
            ipu = json.loads(vfile.read_text())["valu_insts_per_update"]
            ginst = ipu * updates / max(integ["ms"] * 1e-3, 1e-12) / 1e9
            roofline["valu"] = {"achieved": round(ginst, 1), "peak": VALU_PEAK_GINST, "unit": "G wave-instr/s",
                                "frac": round(ginst / VALU_PEAK_GINST, 4), "insts_per_update": ipu}
        except (KeyError, ValueError):
            pass
    mesh_ms = kt["mesh"]["ms"]
    mc_bytes = MC_BYTES_PER_ITEM * inc_items + MC_BYTES_PER_TRIANGLE * inc_tris
    mc_achieved = mc_bytes / max(mesh_ms * 1e-3, 1e-12) / 1e9
    mc_roofline = {"bound": "hbm", "achieved": round(mc_achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                   "frac": round(mc_achieved / HBM_PEAK_GBS, 4), "traffic": None,
                   "kernel": "incremental marching cubes (meshNeighboursKernel+meshFusedKernel and the listing)",
                   "avg_call_ms": round(mesh_ms / max(inc_calls, 1), 4),
                   "items_staged_per_call": inc_items / max(inc_calls, 1),
                   "triangles_per_call": inc_tris / max(inc_calls, 1)}

    cpu = None
    if not args.no_cpu_baseline:
        f0 = synth.render(w.scene, w.K, w.poses[:1], w.noise)[0]
        if updates_frame0 is None:
            with vhm.Volume(vp) as v2:
                v2.integrate(f0, w.K, w.poses[0])
                updates_frame0 = v2.stats()["updates"]
        log("[bench] timing the CPU baseline on frame 0 ...")
        cpu = cpu_baseline(w, f0, updates_frame0)
        log(f"[bench] cpu baseline: {cpu['value']:.0f} updates/s ({cpu['seconds']:.1f} s)")

    out = {
        "metric": "voxel updates/s (% HBM peak)",
        "value": updates / elapsed,
        "unit": "voxel updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": elapsed * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded analytic scene, GPU-rendered uint16 depth resident in HBM)",
        "config": {"workload": w.name, "frames": K, "resolution": f"{w.K.width}x{w.K.height}",
                   "voxel_size_m": w.voxel_size, "truncation_m": w.truncation, "directional_bins": 6,
                   "slots_per_bin": 3, "bilateral": "5x5 sigma_s=4.5px sigma_r=30mm",
                   "blocks_allocated": blocks, "voxels_visited": visited},
        "frames_per_s": K / elapsed,
        "kernel_ms": {k: round(v["ms"], 3) for k, v in kt.items() if k in timed},
        "incremental_mesh": {"every_frames": args.mesh_every, "calls": inc_calls, "blocks_remeshed": inc_blocks,
                             "triangles": inc_tris, "ms": round(kt["mesh"]["ms"], 3),
                             "items_staged": inc_items,
                             "tri_capacity": args.mesh_tri_capacity, "max_triangles_per_call": inc_max,
                             "capacity_overflows": inc_overflows},
        "roofline": roofline,
        "roofline_mesh": mc_roofline,
        "cpu_baseline": cpu,
        **mesh,
    }
    if cpu:
        out["cpu_baseline"] = {k: v for k, v in cpu.items() if k != "seconds"}
    print(json.dumps(out))
    vol.close()


if __name__ == "__main__":
    main()
