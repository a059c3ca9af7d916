"""Build recipe for libhpxfft_hip.so (gfx950) and the C++ drop-in programs.

Compiles every HIP translation unit under ``csrc/`` with hipcc for gfx950
(in parallel), links ``lib/libhpxfft_hip.so``, and builds the C++ programs
under ``cpp/`` (drop-in class harness and tests) against it with g++.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libhpxfft_hip.so")
BIN_DIR = os.path.join(HERE, "bin")
INCLUDE = os.path.join(REPO, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 (v_pk_*_f32) code generation stays off
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", INCLUDE,
]

SOURCES = ["inst_f64.hip", "inst_f32.hip", "inst_f16.hip", "inst_bf16.hip", "inst_batched.hip", "capi.hip"]
CPP_PROGRAMS = {
    # binary name : source under cpp/
    "shared_loop_2d_gpu": "shared_loop_2d_gpu.cpp",
    "test_shared_loop_gpu": "test_shared_loop_gpu.cpp",
    "test_vector_2d": "test_vector_2d.cpp",
}


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build_library(jobs: int | None = None) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_programs() -> list[str]:
    os.makedirs(BIN_DIR, exist_ok=True)
    cpp = os.path.join(HERE, "cpp")
    out = []
    for name, src in CPP_PROGRAMS.items():
        path = os.path.join(cpp, src)
        if not os.path.exists(path):
            continue
        dst = os.path.join(BIN_DIR, name)
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", INCLUDE, path, "-o", dst,
              f"-L{LIB_DIR}", "-lhpxfft_hip", f"-Wl,-rpath,{LIB_DIR}", "-Wl,-rpath,$ORIGIN/../lib"])
        out.append(dst)
    return out


def build_all() -> None:
    build_library()
    build_programs()


if __name__ == "__main__":
    build_all()
    print(LIB)
    sys.exit(0)
