"""Times the restated CPU path (the oracle, liboracle.so) against the
reference's own CPU path compiled from its sources (oracle/_ref/libvhm_ref.so)
on identical inputs, one core each, and writes tests/golden/cpu_calibration.json.

    python tests/golden/time_cpu_paths.py

Inputs: frames of the plain-small room (640x480, voxel 1 cm) and of the
plain-at-scale walk (1280x720, voxel 5 mm), integrated with the reference's
own constants (its filter constants are compiled in: sigma_s 1.5 px,
sigma_r 1.2 mm), frame by frame through filter + allocation + integration
(TsdfIntegrationBackend::integrate, reconstruction_backend.cpp:259-267), as
the reference demo's timer around integrate does (main.cpp:433-451).  The
two paths' block and observed-voxel counts after every frame must agree.
"""
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
sys.path.insert(0, str(HERE.parent))

import oracle_lib as O  # noqa: E402
import voxelhashmt_amd as vhm  # noqa: E402
from voxelhashmt_amd import synth  # noqa: E402


def run(w, n):
    frames = synth.render(w.scene, w.K, w.poses[:n], w.noise).astype(np.float32) * np.float32(0.001)
    vp = vhm.default_params(w.voxel_size, w.truncation)  # reference constants
    ov = O.OracleVolume(O.oracle_params(vp))
    rv = O.RefVolume(w.voxel_size, w.truncation)
    t_ref = t_port = 0.0
    for i in range(n):
        pose = vhm.pose_array(w.poses[i])
        t0 = time.perf_counter()
        rv.integrate(frames[i], w.K, pose)
        t1 = time.perf_counter()
        ov.integrate(frames[i], w.K, pose)
        t2 = time.perf_counter()
        t_ref += t1 - t0
        t_port += t2 - t1
        assert (rv.block_count(), rv.observed_voxel_count()) == (ov.block_count(), ov.observed_voxel_count()), i
        print(f"{w.name} frame {i}: reference {t1 - t0:.2f} s, port {t2 - t1:.2f} s", flush=True)
    return {"workload": w.name, "frames": n, "blocks": int(ov.block_count()), "updates": int(ov.updates()),
            "reference_s": round(t_ref, 3), "port_s": round(t_port, 3), "port_over_reference": round(t_port / t_ref, 4),
            "reference_updates_per_s": round(ov.updates() / t_ref), "port_updates_per_s": round(ov.updates() / t_port)}


if __name__ == "__main__":
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libvhm_ref.so missing: run `make -C oracle ref`")
    os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[-1]})
    cpu = next((ln.split(":", 1)[1].strip() for ln in open("/proc/cpuinfo") if ln.startswith("model name")), "?")
    out = {"host": f"{cpu}, {os.cpu_count()} logical CPUs, 1 core used", "compiler": "g++ -O2 (both)",
           "runs": [run(synth.plain_small(), 5), run(synth.plain_scale(), 2)]}
    (HERE / "cpu_calibration.json").write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps(out, indent=1))
