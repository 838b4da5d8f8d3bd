"""The drop-in, built: the reference's own demo (src/main.cpp) with its
reconstruction_backend.cpp patched as INTEGRATION.md describes
(oracle/vhm_backend.sed: the CUDA dispatch branches call the libvhm host
mirror), compiled against libvhm.so by `make -C oracle ref` into
oracle/_ref/tsdf_voxel_hash_demo_vhm.  Without a device it refuses the GPU
backend with the reference's error shape and falls back to the CPU path in
auto mode, whose output equals the reference demo's.  (On the GPU, the same
executable's GPU backend is compared with the reference CPU run in
tests/test_gpu_demo.py.)"""
import json
import subprocess
import sys
from pathlib import Path

import pytest

import voxelhashmt_amd as vhm

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "oracle" / "_ref" / "tsdf_voxel_hash_demo_vhm"
sys.path.insert(0, str(ROOT / "tests" / "golden"))
import make_demo_golden as G  # noqa: E402

pytestmark = pytest.mark.skipif(not EXE.exists(), reason="oracle/_ref/tsdf_voxel_hash_demo_vhm not built "
                                                         "(needs the reference tree: make -C oracle ref)")
GOLDEN = json.loads((ROOT / "tests" / "golden" / "demo_synthetic.json").read_text())


def test_patched_reference_demo_without_device(tmp_path):
    if vhm.device_available()[0]:
        pytest.skip("a GPU is present")
    g = GOLDEN["w32_h24_f3_v10mm"]
    r = subprocess.run([str(EXE), "--synthetic", *g["args"], "--backend", "gpu", "--output", str(tmp_path / "m.ply")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "GPU backend requested but HIP is unavailable" in r.stderr
    r = subprocess.run([str(EXE), "--synthetic", *g["args"], "--backend", "auto", "--output", str(tmp_path / "m.ply")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "Integration backend: CPU (HIP backend built but unavailable" in r.stdout
    got = G.parse_stdout(r.stdout)
    assert got["frames"] == g["frames"] and got["final"] == g["final"]
    assert G.ply_digest(tmp_path / "m.ply") == (g["ply_faces"], g["ply_digest"])
