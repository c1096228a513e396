"""Builds libbpf_amd.so (gfx950) in place: hipcc on every source of csrc/.

Usage: python build.py [--jobs N] [--debug]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbpf_amd.so")
SOURCES = ["api.cpp", "host_common.cpp", "profile.cpp", "analysis.cpp", "krylov.cpp", "poly.cpp", "solvers.cpp", "bicgstab.cpp",
           "refactor.cpp", "ordering.cpp", "lu_host.cpp", "batch.cpp", "dense_api.cpp", "precond.cpp", "relabel.cpp", "noalloc.hip", "block_jacobi.hip", "dense_lu.hip", "wide.hip",
           "sparse_kernels.hip", "dense_kernels.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("BPF_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-parameter",
          "-Wno-unused-result", "-Wno-deprecated-declarations", "-I", CSRC]


def _compile(src: str, debug: bool) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    flags = list(COMMON)
    if debug:
        flags = [f for f in flags if f != "-O3"] + ["-O1", "-g"]
    cmd = [HIPCC] + flags + ["-x", "hip", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    warn = (r.stderr or "").strip()
    if warn:
        lines = [ln for ln in warn.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if lines:
            sys.stderr.write(f"[{src}]\n" + "\n".join(lines[:40]) + "\n")
    return obj


def build(jobs: int = 8, debug: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda s: _compile(s, debug), srcs))
    tmp = LIB + ".tmp"
    # rocBLAS / rocSOLVER: the dense algebra of the one-block wide solvers
    rocm_lib = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + [
        "-L" + rocm_lib, "-lrocsolver", "-lrocblas", "-Wl,-rpath," + rocm_lib]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--debug", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.debug))
