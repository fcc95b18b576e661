"""Builds later_amd/liblater_amd.so from the HIP sources with hipcc for gfx950.

Every HIP source is compiled to an object and linked into one shared library
whose exported symbols are the C-ABI of include/later.h.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liblater_amd.so")
SOURCES = ["gen.hip", "panel.hip", "gemm.hip", "check.hip", "rgsqrf.hip", "prof.hip", "ls.hip", "serve.hip", "plan.hip", "batch.hip", "panel_chol.hip", "noalloc.hip", "batch_wg.hip", "ooc.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-fvisibility=hidden",
    # packed-FP32 VALU forms are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _compile(src: str) -> str:
    obj = os.path.join(CSRC, "_build", src.replace(".hip", ".o"))
    os.makedirs(os.path.dirname(obj), exist_ok=True)
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed for {src}")
    return obj


def build(verbose: bool = False) -> str:
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_driver(verbose)
    return LIB


DRIVER_SRC = os.path.join(HERE, "test_qr.c")
DRIVER = os.path.join(HERE, "test_qr")


def build_driver(verbose: bool = False) -> str:
    """The plain-C command-line driver (test_qr algo m n [-check]), compiled
    by gcc against include/later.h and linked to the library."""
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cmd = ["gcc", "-O2", "-std=c11", "-Wall", "-D__HIP_PLATFORM_AMD__", f"-I{rocm}/include", DRIVER_SRC,
           "-o", DRIVER + ".tmp", f"-L{HERE}", "-llater_amd", f"-L{rocm}/lib", "-lamdhip64",
           "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{rocm}/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("driver build failed")
    os.replace(DRIVER + ".tmp", DRIVER)
    if verbose:
        print("built", DRIVER)
    return DRIVER


if __name__ == "__main__":
    build(verbose=True)
