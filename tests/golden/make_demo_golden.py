"""Writes the demo fixture from the reference's own CPU-only demo executable.

    python tests/golden/make_demo_golden.py     (needs oracle/_ref/tsdf_voxel_hash_demo,
                                                 built by `make -C oracle ref`)

Runs `tsdf_voxel_hash_demo --synthetic` (src/main.cpp:400-480) at three
settings and records what it prints -- every frame's active blocks and
observed voxels, the final block / observed-voxel / raw-triangle counts --
and, from the PLY it writes, the face count and the sha256 of the sorted
face records (9 vertex floats + the face normal).  Timings are dropped.

The demo's own defaults (voxel 0.5 mm, truncation 1.5 mm) are not among the
settings: the reference CPU path needs more than this container's 64 GiB
there (it was killed by the kernel's OOM handler after 15 minutes), so no
fixture of it can be made here.
"""
import hashlib
import json
import re
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
sys.path.insert(0, str(HERE.parent))

import oracle_lib as O  # noqa: E402
from voxelhashmt_amd import hostio as H  # noqa: E402

CASES = {
    "w32_h24_f3_v10mm": ["--width", "32", "--height", "24", "--frames", "3", "--voxel-size", "0.01",
                         "--truncation", "0.03"],
    "w128_h96_f28_v4mm": ["--frames", "28", "--voxel-size", "0.004", "--truncation", "0.012"],
    "w128_h96_f28_v2mm": ["--frames", "28", "--voxel-size", "0.002", "--truncation", "0.006"],
}

FRAME = re.compile(r"^Frame (\d+)/(\d+) rendered in .*active blocks=(\d+), observed voxels=(\d+)$")
FINAL = re.compile(r"^Final stats: blocks=(\d+), observed voxels=(\d+), raw triangles=(\d+)$")


def parse_stdout(text: str) -> dict:
    frames, final = [], None
    for line in text.splitlines():
        m = FRAME.match(line)
        if m:
            frames.append([int(m.group(3)), int(m.group(4))])
        m = FINAL.match(line)
        if m:
            final = {"blocks": int(m.group(1)), "observed_voxels": int(m.group(2)),
                     "raw_triangles": int(m.group(3))}
    return {"frames": frames, "final": final}


def ply_digest(path) -> tuple[int, str]:
    tris, normals = H.read_ply(path)
    rec = np.concatenate([tris, normals], axis=1).astype(np.float32).view(np.uint32)
    rec = rec[np.lexsort(rec.T[::-1])] if len(rec) else rec
    return len(rec), hashlib.sha256(rec.tobytes()).hexdigest()


def main():
    exe = O.REF_DEMO_EXE
    if not exe.exists():
        raise SystemExit(f"{exe} missing: run `make -C oracle ref`")
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, args in CASES.items():
            ply = Path(tmp) / f"{name}.ply"
            t0 = time.time()
            r = subprocess.run([str(exe), "--synthetic", *args, "--output", str(ply)], capture_output=True,
                               text=True, check=True)
            rec = parse_stdout(r.stdout)
            rec["args"] = args
            rec["ply_faces"], rec["ply_digest"] = ply_digest(ply)
            rec["reference_seconds"] = round(time.time() - t0, 1)
            out[name] = rec
            print(name, rec["final"], rec["ply_faces"], rec["reference_seconds"], "s", flush=True)
    (HERE / "demo_synthetic.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
