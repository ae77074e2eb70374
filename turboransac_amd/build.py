"""Builds the HIP engine library for gfx950 and the host-side helpers.

    python -m turboransac_amd.build            # engine + oracle + C++ example
    python -m turboransac_amd.build engine     # engine only

The engine is one shared library, lib/libturboransac_amd.so, compiled by
hipcc from csrc/ (HIP kernels and the C ABI).  Packed FP32 instructions are
disabled for all device code (-packed-fp32-ops).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
ENGINE_LIB = os.path.join(LIBDIR, "libturboransac_amd.so")
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "_build", "liboracle.so")
EXAMPLE_BIN = os.path.join(LIBDIR, "tr_example")
EVAL_BIN = os.path.join(LIBDIR, "tr_eval7scenes")

ARCH = os.environ.get("TR_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

ENGINE_SOURCES = ["kernels.hip", "engine.hip", "fundamental.hip", "server.hip", "capi.hip"]

DEVICE_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=off",
    # MFMA results land in VGPRs, read by the VALU tail without
    # v_accvgpr_read copies.
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wall",
    "-Wno-unused-result",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_oracle() -> str:
    os.makedirs(os.path.dirname(ORACLE_LIB), exist_ok=True)
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
          "-Wall", "-o", ORACLE_LIB, os.path.join(ORACLE_DIR, "oracle.cpp"), "-lpthread"])
    return ORACLE_LIB


def build_engine(jobs: int = 4) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    objs = []
    procs = []
    for src in ENGINE_SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(objdir, src + ".o")
        objs.append(obj)
        if src.endswith(".hip"):
            cmd = [HIPCC, *DEVICE_FLAGS, *inc, "-c", path, "-o", obj]
        else:
            cmd = [HIPCC, "-O2", "-std=c++17", "-fPIC", "-Wall", *inc,
                   "-D__HIP_PLATFORM_AMD__", "-c", path, "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            if procs.pop(0).wait() != 0:
                raise RuntimeError(f"compile failed: {src}")
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("compile failed")
    tmp = ENGINE_LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
          "-lpthread"])
    os.replace(tmp, ENGINE_LIB)
    return ENGINE_LIB


def build_example() -> str:
    """C++ caller of the header-only wrapper include/turboransac.hpp."""
    src = os.path.join(CSRC, "example_main.cpp")
    _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), src,
          "-o", EXAMPLE_BIN, "-L", LIBDIR, "-lturboransac_amd", f"-Wl,-rpath,{LIBDIR}",
          "-Wl,-rpath,$ORIGIN"])
    return EXAMPLE_BIN


def build_eval() -> str:
    """Evaluator of 7-Scenes correspondence files (csrc/eval_7scenes.cpp)."""
    src = os.path.join(CSRC, "eval_7scenes.cpp")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), src,
          "-o", EVAL_BIN, "-L", LIBDIR, "-lturboransac_amd", f"-Wl,-rpath,{LIBDIR}",
          "-Wl,-rpath,$ORIGIN"])
    return EVAL_BIN


def build_all() -> None:
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    build_oracle()
    build_engine()
    build_example()
    build_eval()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"all": build_all, "engine": build_engine, "oracle": build_oracle,
     "example": build_example, "eval": build_eval}[what]()
