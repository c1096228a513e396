"""Host-only code under sanitizers (CPU).

The orderings, the host sparse LU behind the refactorization
(src/cusolverRF.jl:127-204), the level analysis of the triangular operators,
the elimination plans and the batched solve's host thread pool are built with
g++ under AddressSanitizer + UndefinedBehaviorSanitizer and under
ThreadSanitizer, and their self-check (csrc/host_check.cpp) runs: every result
against the property that defines it, and no sanitizer report.
"""
import os
import shutil
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "blockpowerflow.jl_amd")


def _sanitize():
    sys.path.insert(0, PKG)
    try:
        import sanitize
    finally:
        sys.path.remove(PKG)
    return sanitize


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")
@pytest.mark.parametrize("kind,rounds", [("address", 3), ("thread", 1)])
def test_host_code_under_sanitizer(kind, rounds):
    r = _sanitize().run(kind, rounds)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "0 failures" in r.stdout, out[-4000:]
    assert "Sanitizer" not in r.stderr, r.stderr[-4000:]
