"""CPU checks of the C ABI: the library loads and exports every declared symbol."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "bpf.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bpf_[a-z0-9_]+)\s*\(", src)))


def test_header_symbols_exported(bpf):
    names = _declared()
    assert len(names) > 30
    for n in names:
        assert hasattr(bpf.L, n), n


def test_python_signatures_cover_header(bpf):
    assert set(_declared()) == set(bpf._lib.SIGNATURES)


def test_version_and_error_string(bpf):
    assert b"gfx950" in bpf.L.bpf_version()
    assert isinstance(bpf.L.bpf_last_error(), bytes)
