"""Build libcochlea_amd.so in-tree with hipcc for gfx950.

Every HIP translation unit is compiled with -ffp-contract=off (the solver's
arithmetic contract forbids implicit FMA fusion) and with packed FP32 ops
disabled.  Objects go to cochlearmodel_amd/_build/, the library next to this
file.  Run as ``python -m cochlearmodel_amd.build``.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcochlea_amd.so")
DRIVER_SRC = os.path.join(HERE, "host", "cochlea_run.cpp")
DRIVER = os.path.join(HERE, "cochlea_run")
ROOT = os.path.dirname(HERE)

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = os.environ.get("CM_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", CSRC, "-I", os.path.join(ROOT, "include"), "-Wall",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *HIP_FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed for {src}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    sources = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libcochlea_amd.so failed")
    os.replace(tmp, LIB)
    # command-line host of the VOICE path, linked against the library
    cmd = [HIPCC, "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), DRIVER_SRC, "-o",
           DRIVER + ".tmp", "-L", HERE, "-lcochlea_amd", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build of cochlea_run failed")
    os.replace(DRIVER + ".tmp", DRIVER)
    if verbose:
        print(f"built {LIB} from {len(sources)} sources and {DRIVER}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
