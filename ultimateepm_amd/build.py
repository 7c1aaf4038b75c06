"""Build recipe for libuepm.so (HIP kernels for gfx950 + host C ABI), the
drop-in C++ EPP classes over it (libuepm_epp.so) and their test driver.

Compiles in-tree with hipcc / g++; the shared libraries land next to this file
so they travel with the repository snapshot.  Packed FP32 instructions are
disabled for all device code.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
OUT = os.path.join(HERE, "libuepm.so")
BUILD_DIR = os.path.join(HERE, "_build")
EPP_DIR = os.path.join(HERE, "epp")
EPP_OUT = os.path.join(HERE, "libuepm_epp.so")
EPP_SOURCES = ["Material.cpp", "yaml_reader.cpp", "device_context.cpp", "SymmetryPoints.cpp", "Hamiltonian.cpp",
               "BandStructure.cpp", "bz_mesh.cpp", "DielectricFunction.cpp", "bz_states.cpp"]
DROPIN_TEST_SRC = os.path.join(os.path.dirname(HERE), "tests", "cpp", "test_dropin.cpp")
DROPIN_TEST_BIN = os.path.join(BUILD_DIR, "test_dropin")
CXX = os.environ.get("CXX", "g++")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
DEVICE_SOURCES = ["solver_f64.hip", "tridiag_blocked.hip", "refine.hip", "copy.hip", "tridiag_reg.hip", "backtransform.hip",
                  "tetra_dos.hip", "dielectric.hip", "noalloc.hip"]
HOST_SOURCES = ["uepm_abi.cpp", "runtime.cpp", "msh_io.cpp", "bzmesh.cpp", "dielectric.cpp"]
COMMON = ["-O3", "-fPIC", "-std=c++17", f"-I{INCLUDE}", f"-I{CSRC}"]
DEVICE_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                "-munsafe-fp-atomics"]


# Per-file device flags.  refine.hip: MFMA accumulators in VGPRs (the default
# heuristic shuttles them through AGPRs on every pass of the H X loop).
FILE_FLAGS = {"refine.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]}


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD_DIR, exist_ok=True)
    objs = []
    for src in DEVICE_SOURCES:
        obj = os.path.join(BUILD_DIR, src + ".o")
        _run([HIPCC, "-x", "hip", *COMMON, *DEVICE_FLAGS, *FILE_FLAGS.get(src, []), "-c", os.path.join(CSRC, src),
              "-o", obj])
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD_DIR, src + ".o")
        _run([HIPCC, "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-pthread", *COMMON, "-c",
              os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = OUT + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-pthread", "-o", tmp, *objs])
    os.replace(tmp, OUT)
    build_epp()
    return OUT


def build_epp() -> str:
    """The drop-in EPP classes (host C++ only, no device code) and their test driver."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    flags = ["-std=c++17", "-O2", "-fPIC", "-Wall", f"-I{INCLUDE}", f"-I{EPP_DIR}"]
    tmp = EPP_OUT + ".tmp"
    _run([CXX, *flags, "-shared", "-o", tmp, *[os.path.join(EPP_DIR, f) for f in EPP_SOURCES], f"-L{HERE}", "-luepm",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, EPP_OUT)
    _run([CXX, *flags, "-o", DROPIN_TEST_BIN, DROPIN_TEST_SRC, f"-L{HERE}", "-luepm_epp", "-luepm",
          "-Wl,-rpath,$ORIGIN/.."])
    return EPP_OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
