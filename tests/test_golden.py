"""Committed golden fixtures (tests/golden/case30.npz, written by
tests/golden/make_golden.py from restatements that share no code with the
oracle or the HIP library): checked against the oracle on the CPU and against
the device path on the GPU."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "case30.npz")


@pytest.fixture(scope="module")
def g():
    return np.load(GOLD)


def test_generator_reproduces_fixture(gen, g):
    _, S = gen.make_system(30, 3, seed=7)
    np.testing.assert_array_equal(S.indptr, g["indptr"])
    np.testing.assert_array_equal(S.indices, g["indices"])
    np.testing.assert_array_equal(S.data, g["data"])
    np.testing.assert_array_equal(S.B, g["B"])


def test_oracle_matches_golden(oracle, g):
    ip, ix, dv, B = g["indptr"], g["indices"], g["data"], g["B"]
    np.testing.assert_allclose(oracle.ilu0(ip, ix, dv), g["ilu0"], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(oracle.spmm(ip, ix, dv, B), g["spmm"], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(oracle.direct_solve(ip, ix, dv, B), g["xdirect"], rtol=1e-9)
    import scipy.sparse as sp
    n = len(ip) - 1
    A = sp.csr_matrix((dv, ix, ip), shape=(n, n))
    _, st = oracle.block_gmres(A, B[:, :1], memory=12, atol=0.0, rtol=1e-30, itmax=12, history=True)
    np.testing.assert_allclose(np.array(st.residuals), g["gmres"][: len(st.residuals)], rtol=1e-8)


@pytest.mark.gpu
def test_device_matches_golden(bpf, g):
    ip, ix, dv, B = g["indptr"], g["indices"], g["data"], g["B"]
    A = bpf.HipSparseMatrixCSR(ip, ix, dv)
    P = bpf.ilu0(A)
    np.testing.assert_allclose(P.factor_values(), g["ilu0"], rtol=1e-12, atol=1e-14)
    Bd = bpf.HipMatrix.from_numpy(B)
    np.testing.assert_allclose((A @ Bd).to_numpy(), g["spmm"], rtol=1e-13, atol=1e-13)
    # block GMRES with p = 1 is textbook GMRES: same residual norms
    b1 = bpf.HipMatrix.from_numpy(B[:, :1])
    _, st = bpf.block_gmres(A, b1, memory=12, atol=0.0, rtol=1e-30, itmax=12, history=True)
    np.testing.assert_allclose(np.array(st.residuals), g["gmres"][: len(st.residuals)], rtol=1e-7)
    # direct refactorization solve
    rf = bpf.CUSOLVERRF.CusolverRfLU(A, nrhs=3)
    X = bpf.HipMatrix(B.shape)
    bpf.ldiv_(X, rf, Bd)
    np.testing.assert_allclose(X.to_numpy(), g["xdirect"], rtol=1e-9, atol=1e-12)
