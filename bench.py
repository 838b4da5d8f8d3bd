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
cpu_baseline: the reference's CPU path (oracle/_ref, compiled from its
sources) or, if absent, the restated oracle, on frame 0 of the same workload,
one core.
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(w, frame0_u16, gpu_updates_frame0: int) -> dict | None:
    """Times the CPU path on frame 0 (filter + allocation + integration)."""
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
    if O.ref_available() and w.filter_sigma_spatial == 1.5 and w.filter_sigma_range == 0.0012:
        kind = "reference"
    else:
        kind = "port"
    vp = vhm.default_params(w.voxel_size, w.truncation, filter_sigma_spatial=w.filter_sigma_spatial,
                            filter_sigma_range=w.filter_sigma_range)
    ov = O.OracleVolume(O.oracle_params(vp))
    t0 = time.perf_counter()
    if kind == "reference":
        rv = O.RefVolume(w.voxel_size, w.truncation)
        rv.integrate(depth, w.K, pose)
        updates = gpu_updates_frame0
    else:
        ov.integrate(depth, w.K, pose)
        updates = ov.updates()
    dt = time.perf_counter() - t0
    return {"value": updates / dt, "unit": "voxel updates/s", "cores": 1, "kind": kind,
            "sample": f"frame 0 of {w.name} ({w.K.width}x{w.K.height}): filter + allocation + integration, "
                      f"{updates} updates in {dt:.1f} s on 1 host core (of {os.cpu_count()})",
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
                   "integrate_grid": gi, "mesh_grid": gm, "hipgraph": True},
        "frames_per_s": len(ws) * K / elapsed,
        "stage_ms_direct_launches": {k: round(v, 3) for k, v in tot.items() if k.endswith("_ms")},
        "blocks_remeshed": tot["blocks_remeshed"],
        "mesh_capacity_overflows": overflow,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "integrateBatchKernel+integrateDeferredBatchKernel",
                     "avg_launch_ms": round(tot["integrate_ms"] / K, 4),
                     "updates_per_launch": tot["updates"] / K},
        "cpu_baseline": None,
    }
    print(json.dumps(out))
    b.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=480)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="plain_scale",
                    choices=["plain_scale", "plain_small", "batched", "point_cloud"])
    ap.add_argument("--max-blocks", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mesh", action="store_true")
    ap.add_argument("--mesh-every", type=int, default=10, help="incremental MC period in frames (0 = off)")
    ap.add_argument("--mesh-tri-capacity", type=int, default=64_000_000,
                    help="device triangle buffer of the incremental calls (a call that overflows it is counted)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy 5 times (HBM counter calibration under rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only (--gpus must be 1); fusion would shard "
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
    vol.set_timing(True)
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
        "kernel_ms": {k: round(v["ms"], 3) for k, v in kt.items()},
        "incremental_mesh": {"every_frames": args.mesh_every, "calls": inc_calls, "blocks_remeshed": inc_blocks,
                             "triangles": inc_tris, "ms": round(kt["mesh"]["ms"], 3),
                             "tri_capacity": args.mesh_tri_capacity, "max_triangles_per_call": inc_max,
                             "capacity_overflows": inc_overflows},
        "roofline": roofline,
        "cpu_baseline": cpu,
        **mesh,
    }
    if cpu:
        out["cpu_baseline"] = {k: v for k, v in cpu.items() if k != "seconds"}
    print(json.dumps(out))
    vol.close()


if __name__ == "__main__":
    main()
