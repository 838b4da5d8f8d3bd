"""Builds libvhm.so (gfx950) in-tree with hipcc.

    python -m voxelhashmt_amd.build        # or voxelhashmt_amd.build.build()

Every HIP translation unit is compiled with FP contraction off (so the
reference CPU path's float rounding sequence is reproduced exactly), HIP's
correctly-rounded fp32 divide/sqrt, fp32 denormals kept, and packed-fp32
instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_obj"
LIB = PKG / "libvhm.so"
SOURCES = ["vhm_filter.hip", "vhm_alloc.hip", "vhm_integrate.hip", "vhm_mesh.hip", "vhm_synth.hip",
           "vhm_capi.hip", "vhm_denoise.hip", "vhm_denoise_api.hip", "vhm_batch.hip", "vhm_pointcloud.hip",
           "vhm_blocks.hip"]
# Host-only C++ (no GPU code): built with g++ and the reference CPU build's
# float flags (x86-64 baseline, -O2, no contraction).
HOST_SOURCES = ["vhm_io.cpp"]
HOST_FLAGS = ["-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-Wall", "-Wextra"]
DEMO = PKG / "vhm_demo"
DEMO_SOURCE = PKG / "host" / "vhm_demo.cpp"
ARCH = os.environ.get("VHM_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-gline-tables-only", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument", "-Wno-unused-value",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm required to build libvhm.so)")


def _compile(src: str) -> Path:
    out = OBJ / (Path(src).stem + ".o")
    if src.endswith(".cpp"):
        cmd = [_gxx(), *HOST_FLAGS, "-c", str(CSRC / src), "-o", str(out)]
    else:
        cmd = [hipcc(), *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    noise = ("packed-fp32-ops", "is not a recognized feature")
    msgs = [ln for ln in r.stderr.splitlines() if ln.strip() and not any(n in ln for n in noise)]
    if msgs:
        sys.stderr.write("\n".join(msgs) + "\n")
    return out


def _gxx() -> str:
    for cand in (os.environ.get("CXX"), "g++", "/usr/bin/g++"):
        if cand and shutil.which(cand):
            return cand
    raise RuntimeError("g++ not found (needed for the host-only sources)")


def build(verbose: bool = True) -> Path:
    OBJ.mkdir(exist_ok=True)
    sources = SOURCES + HOST_SOURCES
    with cf.ThreadPoolExecutor(max_workers=min(len(sources), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-lz", "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_demo(verbose)
    return LIB


def build_demo(verbose: bool = True) -> Path:
    """The reference demo's CLI (src/main.cpp) over libvhm.so: vhm_demo."""
    tmp = DEMO.with_name(DEMO.name + ".tmp")
    cmd = [_gxx(), *HOST_FLAGS, str(DEMO_SOURCE), f"-L{PKG}", "-lvhm", "-Wl,-rpath,$ORIGIN",
           "-Wl,--allow-shlib-undefined", "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"demo build failed:\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip():
        sys.stderr.write(r.stderr)
    os.replace(tmp, DEMO)
    if verbose:
        print(f"built {DEMO}")
    return DEMO


if __name__ == "__main__":
    build()
