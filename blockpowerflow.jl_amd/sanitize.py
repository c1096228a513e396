"""Sanitizer builds of the host-only code of libbpf_amd.

The orderings (csrc/ordering.cpp), the host sparse LU (csrc/lu_host.cpp), the
level analysis and elimination plans and the batched solve's thread pool
(csrc/host_common.cpp, csrc/host_common.h) use no HIP, so they build with g++
alone.  This script builds them with their self-check (csrc/host_check.cpp)
under the address + undefined-behaviour sanitizers and, separately, under the
thread sanitizer, and runs the checks.

Usage: python sanitize.py [--kind address|thread|all] [--rounds N]
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
HOST_SOURCES = ["host_check.cpp", "host_common.cpp", "ordering.cpp", "lu_host.cpp"]
CXX = os.environ.get("CXX_HOST", "g++")
SANITIZERS = {
    "address": ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"],
    "thread": ["-fsanitize=thread"],
}


def build(kind: str) -> str:
    """Compiles the host self-check under one sanitizer; returns the binary's path."""
    os.makedirs(OUT, exist_ok=True)
    exe = os.path.join(OUT, f"host_check_{kind}")
    cmd = ([CXX, "-std=c++17", "-O1", "-g", "-Wall", "-Wextra", "-Wno-unused-parameter", "-Werror"]
           + SANITIZERS[kind] + [os.path.join(CSRC, s) for s in HOST_SOURCES] + ["-o", exe])
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host sanitizer build failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return exe


def run(kind: str, rounds: int) -> subprocess.CompletedProcess:
    """Builds and runs the self-check; a sanitizer report makes the exit status non-zero."""
    exe = build(kind)
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0"
    env["TSAN_OPTIONS"] = "halt_on_error=1"
    return subprocess.run([exe, str(rounds)], capture_output=True, text=True, env=env)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--kind", default="all", choices=["address", "thread", "all"])
    ap.add_argument("--rounds", type=int, default=3)
    a = ap.parse_args()
    rc = 0
    for k in (["address", "thread"] if a.kind == "all" else [a.kind]):
        r = run(k, a.rounds)
        sys.stdout.write(f"[{k}] {r.stdout}")
        sys.stderr.write(r.stderr)
        rc = rc or r.returncode
    sys.exit(rc)
