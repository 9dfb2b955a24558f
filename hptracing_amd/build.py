"""Compile the native libraries in-tree.

  hptracing_amd/lib/libhpt_host.so   g++  (CPU side: OBJ/MTL, textures, KD build)
  hptracing_amd/lib/libhptracing.so  hipcc --offload-arch=gfx950 (HIP kernels + C ABI)

Device code is compiled with contraction off and correctly rounded division
and square root, so its float32 results follow the same IEEE rounding as the
host; packed-FP32 instructions are disabled.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INCLUDE = os.path.join(ROOT, "include")
LIB_DIR = os.path.join(PKG, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HPT_OFFLOAD_ARCH", "gfx950")

HOST_FLAGS = ["-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-Wall", "-Wextra"]
DEVICE_FLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, cwd=ROOT):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_host() -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    out = os.path.join(LIB_DIR, "libhpt_host.so")
    srcs = sorted(glob.glob(os.path.join(PKG, "csrc", "host", "*.cpp")))
    _run(["g++", *HOST_FLAGS, "-shared", f"-I{INCLUDE}", *srcs, "-o", out + ".tmp", "-lpthread"])
    os.replace(out + ".tmp", out)
    return out


def build_device() -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    out = os.path.join(LIB_DIR, "libhptracing.so")
    srcs = sorted(glob.glob(os.path.join(PKG, "csrc", "device", "*.hip")))
    _run([HIPCC, *DEVICE_FLAGS, "-shared", f"-I{INCLUDE}",
          f"-I{os.path.join(PKG, 'csrc', 'device')}", *srcs, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_cli() -> str:
    """The hptracing command line, linked against both libraries."""
    out = os.path.join(LIB_DIR, "hptracing")
    src = os.path.join(PKG, "csrc", "cli", "hptracing.cpp")
    _run(["g++", *HOST_FLAGS, f"-I{INCLUDE}", src, "-o", out + ".tmp", f"-L{LIB_DIR}",
          "-lhpt_host", "-lhptracing", "-Wl,-rpath,$ORIGIN"])
    os.replace(out + ".tmp", out)
    return out


def build_boundary_client() -> str:
    """tests/cpp/boundary_client: a test program written against the C++
    boundary (include/hptracing.hpp) the way the reference's CLI uses its
    classes."""
    src = os.path.join(ROOT, "tests", "cpp", "boundary_client.cpp")
    out = os.path.join(ROOT, "tests", "cpp", "boundary_client")
    _run(["g++", *HOST_FLAGS, f"-I{INCLUDE}", src, "-o", out + ".tmp", f"-L{LIB_DIR}",
          "-lhpt_host", "-lhptracing", "-Wl,-rpath,$ORIGIN/../../hptracing_amd/lib"])
    os.replace(out + ".tmp", out)
    return out


def build_rcp_check() -> str:
    """tests/cpp/rcp_exhaustive: the reciprocal of the intersection test
    against the correctly rounded division over every float (GPU test)."""
    src = os.path.join(ROOT, "tests", "cpp", "rcp_exhaustive.hip")
    out = os.path.join(ROOT, "tests", "cpp", "rcp_exhaustive")
    _run([HIPCC, *DEVICE_FLAGS, f"-I{os.path.join(PKG, 'csrc', 'device')}", f"-I{INCLUDE}", src,
          "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_all() -> None:
    build_host()
    build_device()
    build_cli()
    build_boundary_client()
    build_rcp_check()


if __name__ == "__main__":
    what = sys.argv[1:] or ["host", "device"]
    if "host" in what:
        build_host()
    if "device" in what:
        build_device()
