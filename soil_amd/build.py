"""Builds libsoil_hip.so (HIP kernels for gfx950 + the C ABI) in-tree.

    python -m soil_amd.build

Device code is compiled with -ffp-contract=off (bit-exact f64 arithmetic: the
only fused multiply-adds are explicit fma() calls) and without packed-fp32
instructions.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "libsoil_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HOSTCXX = os.environ.get("HOSTCXX", "g++")
ROCM_INCLUDE = "/opt/rocm/include"
ARCH = os.environ.get("SOIL_HIP_ARCH", "gfx950")

DEVICE_SOURCES = ["kernels.hip"]
HOST_SOURCES = ["box.cpp", "batch.cpp", "geometry.cpp", "gen.cpp", "serve.cpp", "restart.cpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall",
          "-Wno-unused-function", "-I", os.path.join(ROOT, "include"), "-I", CSRC]
DEVICE_FLAGS = ["-x", "hip", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
                "-packed-fp32-ops", "-munsafe-fp-atomics"]
HOST_FLAGS = ["-D__HIP_PLATFORM_AMD__", "-I", ROCM_INCLUDE, "-pthread"]


def _compile(src, obj, device):
    # device sources go through hipcc for gfx950; host-only sources through the host compiler
    cc = [HIPCC] if device else [HOSTCXX]
    cmd = cc + COMMON + (DEVICE_FLAGS if device else HOST_FLAGS) + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    jobs = []
    for s in DEVICE_SOURCES:
        if os.path.exists(os.path.join(CSRC, s)):
            jobs.append((os.path.join(CSRC, s), os.path.join(OUT_DIR, s + ".o"), True))
    for s in HOST_SOURCES:
        if os.path.exists(os.path.join(CSRC, s)):
            jobs.append((os.path.join(CSRC, s), os.path.join(OUT_DIR, s + ".o"), False))
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        objs = list(ex.map(lambda j: _compile(*j), jobs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
