"""Builds the HIP C-ABI library for gfx950 with hipcc (in-tree).

Output: acoustics_amd/lib/libfdtd_acoustics.so.  Every source is compiled
with -Xclang -target-feature -Xclang -packed-fp32-ops so the compiler does
not form packed FP32 instructions.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libfdtd_acoustics.so")
ARCH = os.environ.get("FDTD_OFFLOAD_ARCH", "gfx950")

# (source, compiled as HIP device code?)
SOURCES = [
    ("kernels.hip", True),
    ("device_scene.cpp", True),
    ("solver.cpp", True),
    ("batch.cpp", True),
    ("server.cpp", True),
    ("noalloc.cpp", True),
    ("setup.cpp", False),
    ("geometry.cpp", False),
    ("xml.cpp", False),
    ("io.cpp", False),
]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
          "-Wno-unused-result"]
DEVICE = ["-x", "hip", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def hipcc():
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(p):
        raise RuntimeError("hipcc not found")
    return p


def _compile(args):
    src, dev, obj = args
    cmd = [hipcc()] + COMMON + (DEVICE if dev else []) + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile of {src} failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(LIB_DIR, exist_ok=True)
    with tempfile.TemporaryDirectory(prefix="fdtd_build_") as tmp:
        jobs = [(s, d, os.path.join(tmp, os.path.splitext(s)[0] + ".o")) for s, d in SOURCES]
        with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
            objs = list(ex.map(_compile, jobs))
        out_tmp = LIB + ".tmp"
        cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", out_tmp] + objs + ["-lpthread"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(out_tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
