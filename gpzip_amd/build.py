"""Build the native libraries of gpzip_amd in-tree.

    python -m gpzip_amd.build          # libgpzip_hip.so + libgpzip_synth.so

libgpzip_hip.so   HIP kernels (gfx950) + host runtime behind include/gpzip_hip.h
                  + the C++ GpuBackend/ParallelChunkedWriter mirror.
libgpzip_synth.so seeded synthetic workloads for tests and bench.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")

HIP_SOURCES = ["gpz_kernels.hip", "gpz_runtime.hip", "gpz_batch.hip", "gpz_server.hip", "gpz_backend.hip"]
# host-only C++ (the CPU codec and decode side, the hybrid scheduler, the tar
# front-end), built with g++
CPP_SOURCES = ["gpz_cpu.cpp", "gpz_hybrid.cpp", "gpz_archive.cpp"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    "-Wno-unused-value",
    # keep packed FP32 instructions out of the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
    "-I" + CSRC,
]

HIP_LIB = os.path.join(HERE, "libgpzip_hip.so")
SYNTH_LIB = os.path.join(HERE, "libgpzip_synth.so")


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{msg}")


def build_hip(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(BUILD, src + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", path, "-o", obj])
    for src in CPP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        objs.append(obj)
        cmds.append(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        list(ex.map(_run, cmds))
    tmp = HIP_LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread", "-lz"])
    os.replace(tmp, HIP_LIB)
    if verbose:
        print("built", HIP_LIB)
    return HIP_LIB


def build_synth(verbose: bool = False) -> str:
    tmp = SYNTH_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", tmp, os.path.join(CSRC, "gpz_synth.cpp")])
    os.replace(tmp, SYNTH_LIB)
    if verbose:
        print("built", SYNTH_LIB)
    return SYNTH_LIB


def build_all(verbose: bool = False) -> None:
    build_synth(verbose)
    build_hip(verbose)
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    build_all(verbose=True)
    sys.exit(0)
