"""Writes the committed golden fixtures of tests/golden/ (run from the
repository root: python tests/golden/make_golden.py).

Everything here is computed by straightforward restatements that share no
code with oracle/ or the HIP library, so the fixtures check both:

  * system:  a 30-bus synthetic polar Newton-Raphson Jacobian (the package's
             generator, seed 7) with 3 N(0,1) right-hand sides;
  * ilu0:    ILU(0) factor values by dense IKJ elimination restricted to the
             pattern (Saad, Iterative Methods, Alg. 10.4);
  * spmm:    A @ B with dense NumPy;
  * xdirect: dense LU solve (LAPACK via numpy.linalg.solve);
  * gmres:   residual norms of 12 steps of textbook unpreconditioned GMRES
             (Saad Alg. 6.9: MGS Arnoldi, Givens rotations) on column 0,
             the p = 1 special case of block GMRES (src/block_gmres.jl).
"""
import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def _generator():
    spec = importlib.util.spec_from_file_location(
        "bpf_generator", os.path.join(ROOT, "blockpowerflow.jl_amd", "generator.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["bpf_generator"] = mod      # dataclasses look their module up
    spec.loader.exec_module(mod)
    return mod


def dense_ilu0(indptr, indices, data):
    n = len(indptr) - 1
    A = np.zeros((n, n))
    P = np.zeros((n, n), dtype=bool)
    for i in range(n):
        for k in range(indptr[i], indptr[i + 1]):
            A[i, indices[k]] = data[k]
            P[i, indices[k]] = True
    for i in range(1, n):
        for k in range(i):
            if not P[i, k]:
                continue
            A[i, k] /= A[k, k]
            for j in range(k + 1, n):
                if P[i, j]:
                    A[i, j] -= A[i, k] * A[k, j]
    return np.array([A[i, indices[k]] for i in range(n) for k in range(indptr[i], indptr[i + 1])])


def textbook_gmres(A, b, m):
    n = len(b)
    beta = np.linalg.norm(b)
    V = np.zeros((n, m + 1))
    H = np.zeros((m + 1, m))
    V[:, 0] = b / beta
    cs, sn = np.zeros(m), np.zeros(m)
    g = np.zeros(m + 1)
    g[0] = beta
    res = [beta]
    for j in range(m):
        w = A @ V[:, j]
        for i in range(j + 1):
            H[i, j] = w @ V[:, i]
            w = w - H[i, j] * V[:, i]
        H[j + 1, j] = np.linalg.norm(w)
        V[:, j + 1] = w / H[j + 1, j]
        for i in range(j):
            t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
            H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
            H[i, j] = t
        r = np.hypot(H[j, j], H[j + 1, j])
        cs[j], sn[j] = H[j, j] / r, H[j + 1, j] / r
        H[j, j] = r
        H[j + 1, j] = 0.0
        g[j + 1] = -sn[j] * g[j]
        g[j] = cs[j] * g[j]
        res.append(abs(g[j + 1]))
    return np.array(res)


def main():
    gen = _generator()
    _, S = gen.make_system(30, 3, seed=7)
    n = S.n
    Ad = np.zeros((n, n))
    for i in range(n):
        for k in range(S.indptr[i], S.indptr[i + 1]):
            Ad[i, S.indices[k]] = S.data[k]
    np.savez(os.path.join(HERE, "case30.npz"),
             indptr=S.indptr, indices=S.indices, data=S.data, B=S.B,
             ilu0=dense_ilu0(S.indptr, S.indices, S.data),
             spmm=Ad @ S.B,
             xdirect=np.linalg.solve(Ad, S.B),
             gmres=textbook_gmres(Ad, S.B[:, 0], 12))
    print("wrote", os.path.join(HERE, "case30.npz"), "n =", n)


if __name__ == "__main__":
    main()
