"""The C-ABI library loads without a GPU and exports every function that
include/*.h declares; host-only entry points work without a device."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np

import voxelhashmt_amd as vhm

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    text = "".join(h.read_text() for h in sorted((ROOT / "include").glob("*.h")))
    return sorted(set(re.findall(r"\b(vhm_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    out = subprocess.run(["nm", "-D", "--defined-only", str(vhm.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (vhm_[a-z0-9_]+)$", out, re.M))
    missing = [s for s in _declared() if s not in exported]
    assert not missing, missing
    assert set(_declared()) == set(vhm.exported_symbols())


def test_default_params_presets():
    p = vhm.default_params(0.005, 0.02)
    assert p.voxel_size == np.float32(0.005) and p.layer_normal_cos == np.float32(0.75)
    assert p.filter_disc_base == np.float32(0.0012) and p.filter_min_cos == np.float32(0.6)
    q = vhm.default_params(0.005, 0.02, preset=vhm.PRESET_REFERENCE_CUDA)
    assert q.disc_base == np.float32(0.012) and q.layer_normal_cos == np.float32(0.9)
    assert q.filter_min_cos == np.float32(0.9)


def test_host_synth_renders_without_device():
    from voxelhashmt_amd import synth
    w = synth.tiny(n_frames=1)
    f = synth.render(w.scene, w.K, w.poses, w.noise)
    assert f.dtype == np.uint16 and f.shape == (1, w.K.height, w.K.width)
    assert 0.9 < (f > 0).mean() < 1.0  # 1 % dropouts


def test_create_without_device_reports_no_device():
    ok, msg = vhm.device_available()
    if ok:
        return  # covered by the GPU tests on a device
    h = C.c_void_p()
    rc = vhm.lib().vhm_create(C.byref(vhm.default_params(0.01, 0.04)), C.byref(h))
    assert rc == 5  # VHM_ERR_NO_DEVICE
