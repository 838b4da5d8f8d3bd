"""The demo harness (voxelhashmt_amd/vhm_demo, the reference's
tsdf_voxel_hash_demo command line over libvhm) on the GPU.

* --synthetic at three settings against the reference's own demo run
  (tests/golden/demo_synthetic.json, written by make_demo_golden.py from
  oracle/_ref/tsdf_voxel_hash_demo): every frame's active blocks and
  observed voxels, the final counts, and the PLY's faces bit for bit (sha256
  of the sorted face records).
* --synthetic at the reference's defaults (voxel 0.5 mm), which the
  reference CPU path cannot hold in 64 GiB: runs, and its counts agree with
  each other.
* dataset mode on a dataset written here (16-bit PNG depth, quaternion and
  matrix pose files): every frame's counts and the mesh against the
  reference's own CPU path (oracle/_ref/libvhm_ref.so) fed the frames and
  poses its own loader functions read (those are pinned bit-exact against
  the reference's in tests/test_hostio.py).
"""
import json
import os
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
import pngenc
import voxelhashmt_amd as vhm
from voxelhashmt_amd import hostio as H

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "tests" / "golden"))
import make_demo_golden as G  # noqa: E402

pytestmark = pytest.mark.gpu
DEMO = ROOT / "voxelhashmt_amd" / "vhm_demo"
GOLDEN = json.loads((ROOT / "tests" / "golden" / "demo_synthetic.json").read_text())


def _run(args, timeout=300):
    print(f"[demo] {' '.join(args)}", flush=True)
    r = subprocess.run([str(DEMO), *args], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


@pytest.mark.parametrize("case", sorted(GOLDEN))
def test_synthetic_demo_matches_reference_run(gpu, tmp_path, case):
    g = GOLDEN[case]
    out = _run(["--synthetic", *g["args"], "--output", str(tmp_path / "m.ply"), "--max-blocks", "200000"])
    got = G.parse_stdout(out)
    print(f"[demo] {case}: {got['final']}", flush=True)
    assert "Integration backend: GPU (HIP, auto-selected)" in out
    assert got["frames"] == g["frames"]
    assert got["final"] == g["final"]
    faces, digest = G.ply_digest(tmp_path / "m.ply")
    assert (faces, digest) == (g["ply_faces"], g["ply_digest"])


def test_synthetic_demo_reference_defaults(gpu):
    out = _run(["--synthetic", "--output", "/dev/null"], timeout=600)
    got = G.parse_stdout(out)
    print(f"[demo] defaults: {got['final']}", flush=True)
    print("\n".join(l for l in out.splitlines() if l.startswith(("Frame 28/", "Timing"))), flush=True)
    assert len(got["frames"]) == 28
    blocks = [b for b, _ in got["frames"]]
    observed = [v for _, v in got["frames"]]
    assert blocks == sorted(blocks) and observed == sorted(observed)
    assert got["final"]["blocks"] == blocks[-1] > 300_000
    assert got["final"]["raw_triangles"] > 10_000_000


def _quaternion(R):
    """A unit quaternion (w, x, y, z) of a rotation matrix (Shepperd's method)."""
    R = np.asarray(R, np.float64).reshape(3, 3)
    tr = np.trace(R)
    if tr > 0:
        s = 2 * np.sqrt(1 + tr)
        return [s / 4, (R[2, 1] - R[1, 2]) / s, (R[0, 2] - R[2, 0]) / s, (R[1, 0] - R[0, 1]) / s]
    i = int(np.argmax(np.diag(R)))
    j, k = (i + 1) % 3, (i + 2) % 3
    s = 2 * np.sqrt(1 + R[i, i] - R[j, j] - R[k, k])
    q = [0.0] * 4
    q[0] = (R[k, j] - R[j, k]) / s
    q[1 + i] = s / 4
    q[1 + j] = (R[j, i] + R[i, j]) / s
    q[1 + k] = (R[k, i] + R[i, k]) / s
    return q


@pytest.mark.skipif(not O.ref_available(), reason="reference CPU path not built here")
def test_dataset_mode_matches_reference_cpu_path(gpu, tmp_path):
    w, h, n, scale, voxel, trunc = 128, 96, 8, 0.001, 0.006, 0.018
    K = H.demo_intrinsics(w, h)
    (tmp_path / "intrinsic.txt").write_text(f"{K.fx!r} {K.fy!r} {K.cx!r} {K.cy!r} {w} {h} {scale}\n")
    rnd = H.DemoRenderer(7)
    for i in range(n):
        T = H.demo_pose(i, n)
        depth = rnd.render(K, T, i)
        raw = np.clip(np.rint(depth / scale), 0, 65535).astype(np.uint16)
        d = tmp_path / str(10 + 3 * i)
        d.mkdir()
        pngenc.write_depth_png(d / "depth.png", raw)
        R = np.array(T.R, np.float64).reshape(3, 3)
        if i % 2 == 0:
            vals = _quaternion(R) + list(T.t)
        else:
            vals = np.concatenate([np.c_[R, np.array(T.t)], [[0, 0, 0, 1]]]).ravel().tolist()
        (d / "pose.txt").write_text(" ".join(f"{v:.9g}" for v in vals) + "\n")

    args = ["--dataset-root", str(tmp_path), "--frame-start", "13", "--frame-step", "6", "--voxel-size",
            str(voxel), "--truncation", str(trunc), "--output", str(tmp_path / "m.ply"), "--max-blocks", "100000"]
    out = _run(args)
    chosen = [13, 19, 25, 31]
    assert f"Frame selection: start=13, end=31, step=6, count={len(chosen)}" in out
    loads = re.findall(r"^Loading frame (\d+) \((\d+)/(\d+)\), valid_depth=(\d+)/(\d+)$", out, re.M)
    integ = re.findall(r"^Integrated frame (\d+), .*active blocks=(\d+), observed voxels=(\d+)$", out, re.M)
    final = re.search(r"^Final stats: blocks=(\d+), observed voxels=(\d+), raw triangles=(\d+)$", out, re.M)
    assert [int(x[0]) for x in loads] == chosen and [int(x[0]) for x in integ] == chosen

    Kd, sc = H.dataset_intrinsics(tmp_path / "intrinsic.txt")
    ref = O.RefVolume(voxel, trunc)
    for j, fid in enumerate(chosen):
        T, depth, valid = H.load_dataset_frame(tmp_path, fid, "pose.txt", Kd, sc)
        assert int(loads[j][3]) == valid and int(loads[j][4]) == w * h
        ref.integrate(depth, Kd, vhm.pose_array(T))
        assert (int(integ[j][1]), int(integ[j][2])) == (ref.block_count(), ref.observed_voxel_count()), fid
    mesh = ref.extract_mesh()
    assert (int(final.group(1)), int(final.group(2)), int(final.group(3))) == \
        (ref.block_count(), ref.observed_voxel_count(), len(mesh))
    H.write_ply(tmp_path / "ref.ply", mesh)
    assert G.ply_digest(tmp_path / "m.ply") == G.ply_digest(tmp_path / "ref.ply")
    print(f"[demo] dataset: {final.group(0)}", flush=True)


REF_VHM_DEMO = ROOT / "oracle" / "_ref" / "tsdf_voxel_hash_demo_vhm"


@pytest.mark.skipif(not REF_VHM_DEMO.exists(), reason="oracle/_ref/tsdf_voxel_hash_demo_vhm not built")
@pytest.mark.parametrize("case", sorted(GOLDEN))
def test_reference_demo_on_libvhm_matches_reference_cpu_run(gpu, tmp_path, case):
    """The drop-in: the reference's own main.cpp with its backend patched to
    call libvhm (oracle/vhm_backend.sed), --backend gpu, gives the reference
    CPU demo's per-frame counts, final counts and PLY faces."""
    g = GOLDEN[case]
    args = [str(REF_VHM_DEMO), "--synthetic", *g["args"], "--backend", "gpu", "--output", str(tmp_path / "m.ply")]
    print(f"[dropin] {' '.join(args[1:])}", flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=300, env={**os.environ, "VHM_MAX_BLOCKS": "200000"})
    assert r.returncode == 0, r.stderr[-2000:]
    assert "Integration backend: GPU (HIP, gfx950)" in r.stdout
    got = G.parse_stdout(r.stdout)
    assert got["frames"] == g["frames"]
    assert got["final"] == g["final"]
    assert G.ply_digest(tmp_path / "m.ply") == (g["ply_faces"], g["ply_digest"])
