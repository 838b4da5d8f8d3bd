"""Seeded synthetic scenes, trajectories and depth sequences.

The fusion path has no dataset in the environment, so every workload is an
analytic scene rendered by libvhm's ray caster (vhm_synth_render): an
axis-aligned room with spheres and boxes, a camera trajectory and a depth
noise model sigma(z) = 0.0012 + 0.0019 (z - 0.4)^2 m with 1 % dropouts,
stored as uint16 millimetres.

Poses follow the reference demo's look-at convention (makeLookAtPose,
main.cpp:177-187): camera axes (right, down, forward) as the columns of R,
world up = +y.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import (Intrinsics, Noise, Pose, Primitive, Scene, VHM_DEVICE, VHM_HOST, lib, make_intrinsics,
               make_pose)


def _normalized(v):
    v = np.asarray(v, dtype=np.float32)
    n = np.float32(math.sqrt(float(np.dot(v, v))))
    return v / n if n > 1e-8 else np.zeros(3, np.float32)


def look_at_pose(eye, target) -> Pose:
    """makeLookAtPose (main.cpp:177-187) in float32."""
    eye = np.asarray(eye, np.float32)
    fwd = _normalized(np.asarray(target, np.float32) - eye)
    up = np.array([0, 1, 0], np.float32)
    right = _normalized(np.cross(up, fwd).astype(np.float32))
    down = np.cross(fwd, right).astype(np.float32)
    R = np.stack([right, down, fwd], axis=1)  # columns
    return make_pose(R, eye)


def circular_poses(n: int, radius: float, eye_height: float, target) -> list[Pose]:
    out = []
    for i in range(n):
        th = 2.0 * math.pi * i / n
        eye = (radius * math.cos(th), eye_height, radius * math.sin(th))
        out.append(look_at_pose(eye, target))
    return out


def random_walk_poses(n: int, seed: int, room_min, room_max, step=0.02, yaw_deg=1.0, margin=0.5,
                      eye_height=1.5, pitch_deg=-12.0, start=None) -> tuple[list[Pose], np.ndarray]:
    """Random walk: `step` metres forward per frame, heading change U(-yaw, yaw)
    degrees, heading reflected off walls `margin` metres inside the room."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lo = np.array([room_min[0] + margin, room_min[2] + margin])
    hi = np.array([room_max[0] - margin, room_max[2] - margin])
    p = np.array(start if start is not None else rng.uniform(lo * 0.5, hi * 0.5))
    th = rng.uniform(0, 2 * math.pi)
    eyes, poses = [], []
    pitch = math.radians(pitch_deg)
    for _ in range(n):
        d = np.array([math.cos(th), math.sin(th)])
        q = p + step * d
        if q[0] < lo[0] or q[0] > hi[0]:
            th = math.pi - th
        if q[1] < lo[1] or q[1] > hi[1]:
            th = -th
        d = np.array([math.cos(th), math.sin(th)])
        p = np.clip(p + step * d, lo, hi)
        eye = np.array([p[0], eye_height, p[1]])
        look = np.array([math.cos(th) * math.cos(pitch), math.sin(pitch), math.sin(th) * math.cos(pitch)])
        poses.append(look_at_pose(eye, eye + look))
        eyes.append(eye)
        th += math.radians(rng.uniform(-yaw_deg, yaw_deg))
    return poses, np.array(eyes)


def _point_box_distance(p, c, h):
    d = np.maximum(np.abs(p - c) - h, 0.0)
    return np.sqrt((d * d).sum(-1))


@dataclass
class SceneSpec:
    room_min: tuple
    room_max: tuple
    prims: list = field(default_factory=list)  # (kind, centre(3), radius, half(3))

    def to_c(self):
        arr = (Primitive * max(len(self.prims), 1))()
        for i, (kind, c, r, h) in enumerate(self.prims):
            arr[i].kind = kind
            for j in range(3):
                arr[i].c[j] = float(c[j])
                arr[i].h[j] = float(h[j])
            arr[i].r = float(r)
        sc = Scene()
        for j in range(3):
            sc.room_min[j] = float(self.room_min[j])
            sc.room_max[j] = float(self.room_max[j])
        sc.num_primitives = len(self.prims)
        sc.primitives = C.cast(arr, C.POINTER(Primitive))
        sc._keep = arr  # keep the array alive with the struct
        return sc

    def sdf(self, p: np.ndarray) -> np.ndarray:
        """Signed distance of points (N,3) to the scene surface (positive in free space)."""
        lo, hi = np.array(self.room_min), np.array(self.room_max)
        d = np.minimum(p - lo, hi - p).min(-1)  # inside the room: distance to the nearest wall
        for kind, c, r, h in self.prims:
            c = np.asarray(c)
            if kind == 0:
                d = np.minimum(d, np.linalg.norm(p - c, axis=-1) - r)
            else:
                q = np.abs(p - c) - np.asarray(h)
                outside = np.linalg.norm(np.maximum(q, 0.0), axis=-1)
                inside = np.minimum(q.max(-1), 0.0)
                d = np.minimum(d, outside + inside)
        return d


def random_scene(seed: int, room_min, room_max, n_prims: int, path: np.ndarray, clearance=0.3,
                 size_scale=1.0) -> SceneSpec:
    """n_prims spheres (radius U[0.1,0.5] x size_scale) / boxes (sides
    U[0.2,0.8] x size_scale) with p = 1/2 each, inside the room, at least
    `clearance` from every camera position in `path`."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lo, hi = np.array(room_min, float), np.array(room_max, float)
    prims, tries = [], 0
    while len(prims) < n_prims and tries < 100000:
        tries += 1
        if rng.uniform() < 0.5:
            r = rng.uniform(0.1, 0.5) * size_scale
            c = rng.uniform(lo + r, hi - r)
            if (np.linalg.norm(path - c, axis=-1) - r).min() < clearance:
                continue
            prims.append((0, c, r, (0.0, 0.0, 0.0)))
        else:
            h = rng.uniform(0.2, 0.8, size=3) / 2 * size_scale
            c = rng.uniform(lo + h, hi - h)
            if _point_box_distance(path, c, h).min() < clearance:
                continue
            prims.append((1, c, 0.0, h))
    return SceneSpec(tuple(room_min), tuple(room_max), prims)


def render(scene: SceneSpec, K: Intrinsics, poses: list[Pose], noise: Noise | None, first_frame: int = 0,
           out=None):
    """Renders len(poses) uint16 frames.  out: None -> host numpy array;
    a torch uint16/int16 CUDA tensor (n, H, W) -> rendered on the GPU."""
    n = len(poses)
    parr = (Pose * max(n, 1))(*poses)
    sc = scene.to_c()
    nz = noise if noise is not None else Noise(enabled=0)
    if out is None:
        out = np.zeros((n, K.height, K.width), dtype=np.uint16)
        rc = lib().vhm_synth_render(C.byref(sc), C.byref(K), parr, n, first_frame, C.byref(nz), out.ctypes.data,
                                    VHM_HOST)
    else:
        rc = lib().vhm_synth_render(C.byref(sc), C.byref(K), parr, n, first_frame, C.byref(nz), out.data_ptr(),
                                    VHM_DEVICE)
    if rc != 0:
        raise RuntimeError(f"vhm_synth_render failed ({rc})")
    return out


def sensor_noise(seed: int, enabled=True, max_depth=0.0) -> Noise:
    return Noise(a=0.0012, b=0.0019, c=0.4, dropout=0.01, max_depth=max_depth, seed=seed, enabled=int(enabled))


@dataclass
class Workload:
    name: str
    K: Intrinsics
    poses: list
    scene: SceneSpec
    noise: Noise
    voxel_size: float
    truncation: float
    filter_sigma_spatial: float
    filter_sigma_range: float
    hash_log2_buckets: int
    max_blocks: int
    target_blocks: int = 0


def plain_small(n_frames: int = 100, seed: int = 0) -> Workload:
    """Plain method, small scene: 4x4x3 m room, 20 primitives, 100 frames
    640x480 on a circle of radius 1.5 m, voxel 1 cm, truncation 4 cm."""
    room_min, room_max = (-2.0, 0.0, -2.0), (2.0, 3.0, 2.0)
    poses = circular_poses(n_frames, 1.5, 1.5, (0.0, 1.0, 0.0))
    path = np.array([list(p.t) for p in circular_poses(max(n_frames, 100), 1.5, 1.5, (0.0, 1.0, 0.0))])
    scene = random_scene(seed, room_min, room_max, 20, path)
    return Workload("plain_small_640x480", make_intrinsics(640, 480, 525.0, 525.0, 319.5, 239.5), poses, scene,
                    sensor_noise(seed), 0.01, 0.04, 4.5, 0.03, 20, 120_000)


def plain_scale(n_frames: int = 500, seed: int = 1) -> Workload:
    """Plain method at scale: 8x8x3 m room, 150 primitives, 500 frames
    1280x720 on a random walk (0.02 m, +-1 deg per frame), voxel 5 mm,
    truncation 2 cm."""
    room_min, room_max = (-4.0, 0.0, -4.0), (4.0, 3.0, 4.0)
    poses, eyes = random_walk_poses(n_frames, seed, room_min, room_max)
    scene = random_scene(seed, room_min, room_max, 150, eyes)
    return Workload("plain_scale_1280x720", make_intrinsics(1280, 720, 1050.0, 1050.0, 639.5, 359.5), poses, scene,
                    sensor_noise(seed), 0.005, 0.02, 4.5, 0.03, 22, 700_000)


BATCH_RESOLUTIONS = ((320, 240), (640, 480), (1280, 720))


def room_for_blocks(seed: int, target_blocks: int, width: int, height: int, n_frames: int = 60) -> Workload:
    """A plain-small-style room (walls + 20 s^2 primitives scaled by s, circular
    trajectory) whose size s is chosen so that fusing n_frames at voxel 1 cm
    allocates about target_blocks blocks (measured on these scenes:
    blocks ~= 42,130 s^2.15 for s in [0.2, 1.3])."""
    s = (target_blocks / 42130.0) ** (1.0 / 2.15)
    L, H = 4.0 * s, 3.0 * s
    room_min, room_max = (-L / 2, 0.0, -L / 2), (L / 2, H, L / 2)
    poses = circular_poses(n_frames, 0.375 * L, 0.5 * H, (0.0, H / 3, 0.0))
    path = np.array([list(p.t) for p in poses])
    scene = random_scene(seed, room_min, room_max, max(2, int(round(20 * s * s))), path, clearance=0.3 * s,
                         size_scale=s)
    f = 525.0 * width / 640.0
    log2 = max(14, int(math.ceil(math.log2(max(target_blocks, 1) * 4))))
    return Workload(f"room{target_blocks}_{width}x{height}",
                    make_intrinsics(width, height, f, f, (width - 1) / 2, (height - 1) / 2), poses, scene,
                    sensor_noise(seed), 0.01, 0.04, 4.5, 0.03, log2, int(target_blocks * 1.6 + 2048), target_blocks)


def batched_scenes(n_scenes: int = 48, seed: int = 2, n_frames: int = 60) -> list[Workload]:
    """Batched plain method: n_scenes independent rooms with target block
    counts log-uniform in [1k, 60k] and resolutions drawn from
    {320x240, 640x480, 1280x720}; n_frames each, voxel 1 cm."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(n_scenes):
        target = int(round(math.exp(rng.uniform(math.log(1000.0), math.log(60000.0)))))
        w, h = BATCH_RESOLUTIONS[int(rng.integers(0, len(BATCH_RESOLUTIONS)))]
        out.append(room_for_blocks(seed * 1000 + i, target, w, h, n_frames))
    return out


def tiny(n_frames: int = 4, seed: int = 7, width=96, height=72) -> Workload:
    """Small test workload (runs through the CPU oracle in seconds)."""
    room_min, room_max = (-1.0, 0.0, -1.0), (1.0, 2.0, 1.0)
    poses = circular_poses(n_frames, 0.6, 1.0, (0.0, 0.8, 0.0))
    path = np.array([list(p.t) for p in poses])
    scene = random_scene(seed, room_min, room_max, 6, path, clearance=0.2)
    f = 0.82 * width
    return Workload(f"tiny_{width}x{height}", make_intrinsics(width, height, f, f, (width - 1) / 2, (height - 1) / 2),
                    poses, scene, sensor_noise(seed), 0.02, 0.06, 1.5, 0.0012, 16, 8192)
