"""CPU checks of the boundary's documentation: every reference citation
``path.jl:a[-b][,c[-d]...]`` in the header, the package, the oracle and the
tests points inside the cited file, and the Julia ``Params`` / ``Stats``
structs of INTEGRATION.md list the fields of ``bpf_solver_params`` /
``bpf_solver_stats`` in order (a ccall passes them by reference)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
CITE = re.compile(r"((?:src|test|benchmarks)/[A-Za-z0-9_/]+\.jl):(\d+(?:-\d+)?(?:,\s?\d+(?:-\d+)?)*)")
SHORT = re.compile(r"(?<![\w/.])(block_utils|block_gmres|block_bicgstab|cusolverRF|preconditioners_nvidia|"
                   r"preconditioners_amd)\.jl:(\d+(?:-\d+)?(?:,\s?\d+(?:-\d+)?)*)")


def _sources():
    for top in ("include", "blockpowerflow.jl_amd", "oracle", "tests", "bench.py", "__graft_entry__.py",
                "DESIGN.md", "INTEGRATION.md"):
        path = os.path.join(ROOT, top)
        if os.path.isfile(path):
            yield path
            continue
        for d, _, files in os.walk(path):
            if "_build" in d or "__pycache__" in d:
                continue
            for f in files:
                if f.endswith((".py", ".h", ".cpp", ".hip", ".c", ".md")):
                    yield os.path.join(d, f)


def _lines(path, cache={}):
    if path not in cache:
        with open(path, encoding="utf-8", errors="replace") as fh:
            cache[path] = sum(1 for _ in fh)
    return cache[path]


def _numbers(spec):
    out = []
    for part in spec.replace(" ", "").split(","):
        out += [int(x) for x in part.split("-")]
    return out


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources not mounted")
def test_reference_citations_in_range():
    bad = []
    short_map = {"block_utils": "src/block_utils.jl", "block_gmres": "src/block_gmres.jl",
                 "block_bicgstab": "src/block_bicgstab.jl", "cusolverRF": "src/cusolverRF.jl",
                 "preconditioners_nvidia": "src/preconditioners_nvidia.jl",
                 "preconditioners_amd": "src/preconditioners_amd.jl"}
    checked = 0
    for path in _sources():
        if os.path.basename(path) == "test_citations.py":
            continue
        text = open(path, encoding="utf-8", errors="replace").read()
        found = [(m.group(1), m.group(2)) for m in CITE.finditer(text)]
        found += [(short_map[m.group(1)], m.group(2)) for m in SHORT.finditer(text)]
        for rel, spec in found:
            ref = os.path.join(REF, rel)
            if not os.path.isfile(ref):
                bad.append(f"{os.path.relpath(path, ROOT)}: {rel} does not exist")
                continue
            n = _lines(ref)
            checked += 1
            for v in _numbers(spec):
                if not 1 <= v <= n:
                    bad.append(f"{os.path.relpath(path, ROOT)}: {rel}:{spec} (file has {n} lines)")
                    break
    assert checked > 100
    assert not bad, "\n".join(bad)


def _c_struct_fields(name):
    src = open(os.path.join(ROOT, "include", "bpf.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    body = re.search(r"typedef struct \{([^{}]*)\}\s*" + name + r"\s*;", src, flags=re.S).group(1)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        ctype, names = decl.split(None, 1)
        fields += [(n.strip(), ctype) for n in names.split(",")]
    return fields


def _julia_struct_fields(name):
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    body = re.search(r"struct " + name + r"\b[^\n]*\n(.*?)\nend", text, flags=re.S).group(1)
    return re.findall(r"(\w+)::(\w+)", body)


@pytest.mark.parametrize("cname,jname", [("bpf_solver_params", "Params"), ("bpf_solver_stats", "Stats")])
def test_integration_structs_match_header(cname, jname):
    jl = {"int": "Cint", "double": "Cdouble"}
    c = [(n, jl[t]) for n, t in _c_struct_fields(cname)]
    assert c == _julia_struct_fields(jname)


@pytest.mark.parametrize("cname,pyname", [("bpf_solver_params", "SolverParams"), ("bpf_solver_stats", "SolverStats")])
def test_ctypes_structs_match_header(bpf, cname, pyname):
    import ctypes as C
    py = {"int": C.c_int, "double": C.c_double}
    c = [(n, py[t]) for n, t in _c_struct_fields(cname)]
    assert c == list(getattr(bpf._lib, pyname)._fields_)
