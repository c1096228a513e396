"""Interpreter exit with live handles (GPU).

A Python process that ends while a context, matrices, an ILU(0) handle, a
relabelled system, a captured inner-solve graph, a batch and a
refactorization handle are all alive must exit cleanly: at interpreter
teardown the destructors run in no fixed order, the default context may be
gone before the handles it owns, and the library's own process-wide state is
released after that.  Each case runs in a child process (a crash there would
end the test runner); the exit status must be the script's own.
"""
import os
import subprocess
import sys
import textwrap

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = textwrap.dedent("""
    import importlib.util, os, sys
    root, mode = sys.argv[1], sys.argv[2]
    pkg = os.path.join(root, "blockpowerflow.jl_amd")
    def load(name, path, search=None):
        spec = importlib.util.spec_from_file_location(name, path, submodule_search_locations=search)
        mod = importlib.util.module_from_spec(spec)
        sys.modules[name] = mod
        spec.loader.exec_module(mod)
        return mod
    bpf = load("blockpowerflow_jl_amd", os.path.join(pkg, "__init__.py"), [pkg])
    gen = load("bpf_generator", os.path.join(pkg, "generator.py"))
    bpf.set_option("relabel_min", 100)
    _, S = gen.make_system(300, 8, seed=3)
    _, S2 = gen.make_system(500, 4, seed=4)
    A = bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data)
    A2 = bpf.HipSparseMatrixCSR(S2.indptr, S2.indices, S2.data)
    B = bpf.HipMatrix.from_numpy(S.B)
    B2 = bpf.HipMatrix.from_numpy(S2.B)
    P = bpf.ilu0(A, 8)
    X, st = bpf.block_gmres(A, B, N=P, ldiv=True, precision="mixed", inner_its=4, inner_width=8,
                            berr_tol=1e-10, itmax=200)
    assert st.solved
    batch = bpf.BatchedSystems([A, A2], [8, 4])
    Xs, stb, solved = batch.solve([B, B2], precision="mixed", inner_its=4, berr_tol=1e-10, itmax=200)
    assert solved.all()
    rf = bpf.CUSOLVERRF.CusolverRfLU(A, nrhs=8)
    Xr = bpf.HipMatrix.from_numpy(S.B)
    bpf.ldiv_(rf, Xr)
    solver = bpf.BlockGMRESSolver(A, B)
    print("handles live", flush=True)
    if mode == "close_ctx":
        bpf.default_context().close()
    elif mode == "raise":
        raise SystemExit(3)
""")


@pytest.mark.gpu
@pytest.mark.parametrize("mode,rc", [("exit", 0), ("close_ctx", 0), ("raise", 3)])
def test_exit_with_live_handles(mode, rc):
    print(f"[exit test] mode {mode}", flush=True)
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT, mode], capture_output=True, text=True, timeout=240)
    out = f"rc {r.returncode}\nstdout:\n{r.stdout[-3000:]}\nstderr:\n{r.stderr[-3000:]}"
    assert "handles live" in r.stdout, out
    assert r.returncode == rc, out
    assert "Segmentation" not in r.stderr and "core dumped" not in r.stderr, out
