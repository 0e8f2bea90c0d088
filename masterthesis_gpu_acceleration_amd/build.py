"""Builds libgpu_rules.so (host C++17 + HIP kernels for gfx950) in-tree.

    python -m masterthesis_gpu_acceleration_amd.build
    python -m masterthesis_gpu_acceleration_amd.build --profile

--profile builds libgpu_rules_prof.so instead (objects in _build_prof/): the
same library with the wave-mode kernels summing cycles per search section
into their counters (GR_STEP_PROFILE); load it with GPU_RULES_LIB=<path>.

Objects go to masterthesis_gpu_acceleration_amd/_build/, the library to
masterthesis_gpu_acceleration_amd/libgpu_rules.so (both git-ignored; the
library travels with the tree).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libgpu_rules.so")
CLI = os.path.join(PKG, "gpu_rules_counter")   # command-line counter over the library (CPU_01's output)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GPU_RULES_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-pthread", "-Wall", "-Wno-unused-result", "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"),
          "-I" + CSRC]
# Device code: gfx950 only; packed-FP32 instruction selection off.
HIPFLAGS = ["-x", "hip", "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-fno-gpu-rdc"]

SOURCES = ["graph_host.cpp", "rules_host.cpp", "planner.cpp", "traffic.cpp", "snapshot.cpp", "engine.cpp", "serve.cpp",
           "refine.cpp",
           "kernels.hip", "search_s4.hip", "search_s8.hip", "search_s16.hip", "sort_unique.hip", "graph_device.hip",
           "noalloc.hip"]


def _compile(src: str, obj_dir: str = OBJ, extra=()) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(obj_dir, os.path.splitext(src)[0] + ".o")
    flags = COMMON + (HIPFLAGS + list(extra) if src.endswith(".hip") else ["-x", "c++", "-D__HIP_PLATFORM_AMD__"])
    cmd = [HIPCC] + flags + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("compile failed: " + src)
    if r.stderr.strip():
        noise = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if noise:
            sys.stderr.write("\n".join(noise) + "\n")
    return obj


def build(verbose: bool = True, profile: bool = False) -> str:
    obj_dir = OBJ + "_prof" if profile else OBJ
    lib = os.path.join(PKG, "libgpu_rules_prof.so") if profile else LIB
    extra = ["-DGR_STEP_PROFILE"] if profile else []
    os.makedirs(obj_dir, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda s: _compile(s, obj_dir, extra), SOURCES))
    tmp = lib + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "-pthread", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, lib)
    if verbose:
        print("built", os.path.relpath(lib, ROOT))
    if not profile:
        build_cli(verbose)
    return lib


def build_cli(verbose: bool = True) -> str:
    """gpu_rules_counter: host C++ over the C ABI, linked to the library next to it."""
    src = os.path.join(CSRC, "cli", "gpu_rules_counter.cpp")
    tmp = CLI + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"), src, "-o", tmp, "-L" + PKG,
           "-lgpu_rules", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("gpu_rules_counter build failed")
    os.replace(tmp, CLI)
    if verbose:
        print("built", os.path.relpath(CLI, ROOT))
    return CLI


if __name__ == "__main__":
    build(profile="--profile" in sys.argv[1:])
