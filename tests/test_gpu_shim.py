"""The C++ host mirror of the reference's GPU backend interface
(tsdfmc_vhm_backend.hpp), driven by tests/cpp/test_shim.cpp with
reference-shaped types, gives what the C-ABI path gives."""
import json
import subprocess
import tempfile
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
import voxelhashmt_amd as vhm
from voxelhashmt_amd import synth

BIN = Path(__file__).resolve().parent / "cpp" / "test_shim"


@pytest.mark.gpu
def test_cpp_shim_matches_capi(gpu):
    if not BIN.exists():
        pytest.fail(f"{BIN} not built (build() compiles it)")
    w = synth.tiny(n_frames=3, seed=6)
    frames = synth.render(w.scene, w.K, w.poses, w.noise)
    poses = np.stack([vhm.pose_array(p) for p in w.poses]).astype(np.float32)
    with tempfile.NamedTemporaryFile(suffix=".bin", delete=False) as f:
        f.write(frames.tobytes())
        f.write(poses.tobytes())
        path = f.name
    r = subprocess.run([str(BIN), str(w.K.width), str(w.K.height), repr(w.K.fx), repr(w.K.cx), repr(w.K.cy),
                        repr(w.voxel_size), repr(w.truncation), "3", path], capture_output=True, text=True,
                       timeout=120)
    Path(path).unlink()
    assert r.returncode == 0, r.stdout + r.stderr
    got = json.loads(r.stdout.strip().splitlines()[-1])
    vp = vhm.default_params(w.voxel_size, w.truncation, hash_log2_buckets=16, max_blocks=16384,
                            max_width=w.K.width, max_height=w.K.height)
    with vhm.Volume(vp) as vol:
        for i in range(3):
            vol.integrate(frames[i], w.K, w.poses[i])
        keys = vol.block_keys()
        assert got["blocks"] == vol.block_count() == got["keys"]
        assert got["observed"] == vol.observed_voxel_count()
        assert got["triangles"] == len(vol.extract_mesh())
        assert got["key_sum"] == int((keys.astype(np.int64) @ np.array([3, 5, 7])).sum())
        assert got["has_cache"] == 1
    # allocateActiveBlocksGpu on the unfiltered last frame == oracle rule on the same frame
    ov = O.OracleVolume(O.oracle_params(vp))
    n_last = len(ov.prepare(frames[2].astype(np.float32) * np.float32(0.001), w.K, poses[2]))
    assert got["active_last_unfiltered"] == n_last
