"""Build the SHOCCS MI355X library: HIP device code for gfx950 plus the C++20
host library, linked into shoccs_amd/lib/libshoccs.so (in-tree).

Every translation unit is compiled with -ffp-contract=off so that FP64 sums and
products round exactly as the reference's, and the HIP units with packed FP32
instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib", "libshoccs.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

DEVICE_SOURCES = ["device/shx_core.hip", "device/shx_systems.hip", "device/shx_march.hip",
                  "device/shx_mixed.hip", "device/shx_adjoint.hip", "device/shx_euler.hip"]
HOST_SOURCES = ["host/mesh.cpp", "host/stencils.cpp", "host/operators.cpp", "host/problem.cpp",
                "host/lua_config.cpp", "host/field_io.cpp"]

COMMON = ["-O3", "-std=c++20", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
          "-Wall", "-Wno-unused-result", "-Wno-unused-value",
          f"-I{os.path.join(ROOT, 'include')}", f"-I{os.path.join(CSRC, 'device')}",
          f"-I{os.path.join(CSRC, 'host')}"]
DEVICE_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
                "-packed-fp32-ops", "-munsafe-fp-atomics"]


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src.replace("/", "_") + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + COMMON + DEVICE_FLAGS + ["-c", path, "-o", obj]
    else:
        # host-only C++20: no HIP runtime, no device code
        cmd = [CXX] + COMMON + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    sources = DEVICE_SOURCES + HOST_SOURCES
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
