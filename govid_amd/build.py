"""Builds the package's native libraries in-tree (govid_amd/lib).

libebivgpu.so: hipcc for gfx950 (kernels + host runtime + C ABI).
libebivenc.so: g++ (host stream encoder).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions stay off (no FP32 on this path, and the
    # feature is disabled for every HIP build here)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
    "-I", os.path.join(ROOT, "include"),
    "-Wl,-rpath,/opt/rocm/lib",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError(f"build failed: {' '.join(cmd[:3])} ...")
    return r.stdout


def build_gpu() -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libebivgpu.so")
    tmp = out + ".tmp"
    srcs = [os.path.join(CSRC, "ebiv_kernels.hip"), os.path.join(CSRC, "ebiv_runtime.cpp"),
            os.path.join(CSRC, "ebiv_host.cpp"), os.path.join(CSRC, "ebiv_container.cpp")]
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, *srcs])
    os.replace(tmp, out)
    return out


def build_encoder() -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libebivenc.so")
    tmp = out + ".tmp"
    cxx = shutil.which("g++") or "g++"
    _run([cxx, "-O3", "-std=c++17", "-fPIC", "-shared", "-pthread", "-ffp-contract=off",
          "-o", tmp, os.path.join(CSRC, "ebiv_encoder.cpp")])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_encoder()
    build_gpu()


if __name__ == "__main__":
    build_all()
    print("built", os.listdir(LIB))
