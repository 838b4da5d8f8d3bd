"""Golden fixtures written from the reference's own CPU path
(tests/golden/make_golden.py): the oracle must reproduce them on the CPU and
the HIP path must reproduce them on the GPU, bit for bit."""
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
import voxelhashmt_amd as vhm

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
from make_golden import block_hash, mesh_hash  # noqa: E402

FIXTURES = sorted((Path(__file__).resolve().parent / "golden").glob("*.npz"))


def _load(path):
    z = np.load(path, allow_pickle=False)
    K = z["K"]
    Ki = vhm.make_intrinsics(int(K[0]), int(K[1]), K[2], K[3], K[4], K[5])
    return z, Ki


def _check_volume(z, keys, blocks, observed, block_count, mesh):
    assert np.array_equal(keys, z["block_keys"])
    assert np.array_equal(block_hash(blocks), z["block_hashes"])
    assert np.array_equal(blocks[z["full_block_index"]].view(np.uint32), z["full_blocks"])
    assert (block_count, observed) == tuple(z["counts"][-1])
    canon = O.canonical_mesh(mesh)
    assert len(canon) == int(z["mesh_count"])
    assert mesh_hash(canon) == z["mesh_hash"]
    assert np.array_equal(canon[:64], z["mesh_head"])


@pytest.mark.parametrize("path", FIXTURES, ids=[p.stem for p in FIXTURES])
def test_oracle_reproduces_reference_fixture(path):
    z, K = _load(path)
    metres = z["frames"].astype(np.float32) * np.float32(0.001)
    vp = vhm.default_params(float(z["voxel_size"]), float(z["truncation"]))
    op = O.oracle_params(vp)
    assert np.array_equal(O.oracle_filter(op, metres[0], K).view(np.uint32), z["filtered0"].view(np.uint32))
    off = 0
    ov = O.OracleVolume(op)
    for i in range(len(metres)):
        n = int(z["frame_key_counts"][i])
        scratch = O.OracleVolume(op)
        got = O.sorted_keys(scratch.prepare(O.oracle_filter(op, metres[i], K), K, z["poses"][i]))
        assert np.array_equal(got, z["frame_keys"][off:off + n])
        off += n
        ov.integrate(metres[i], K, z["poses"][i])
        assert (ov.block_count(), ov.observed_voxel_count()) == tuple(z["counts"][i])
    keys = O.sorted_keys(ov.block_keys())
    _check_volume(z, keys, ov.dump_blocks(keys), ov.observed_voxel_count(), ov.block_count(), ov.extract_mesh())


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=[p.stem for p in FIXTURES])
def test_gpu_reproduces_reference_fixture(gpu, path):
    z, K = _load(path)
    vp = vhm.default_params(float(z["voxel_size"]), float(z["truncation"]), hash_log2_buckets=16,
                            max_blocks=16384, max_width=K.width, max_height=K.height)
    frames = np.ascontiguousarray(z["frames"])
    print(f"[golden] {path.stem}", flush=True)
    with vhm.Volume(vp) as vol:
        assert np.array_equal(vol.filter_depth(frames[0], K).view(np.uint32), z["filtered0"].view(np.uint32))
        off = 0
        for i in range(len(frames)):
            n = int(z["frame_key_counts"][i])
            with vhm.Volume(vp) as scratch:
                f = scratch.filter_depth(frames[i], K)
                got = O.sorted_keys(scratch.allocate_active_blocks(f, K, vhm.make_pose(z["poses"][i][:9],
                                                                                        z["poses"][i][9:])))
            assert np.array_equal(got, z["frame_keys"][off:off + n])
            off += n
            vol.integrate(frames[i], K, vhm.make_pose(z["poses"][i][:9], z["poses"][i][9:]))
            assert (vol.block_count(), vol.observed_voxel_count()) == tuple(z["counts"][i])
        keys = O.sorted_keys(vol.block_keys())
        _check_volume(z, keys, vol.download_blocks(keys), vol.observed_voxel_count(), vol.block_count(),
                      vol.extract_mesh())
