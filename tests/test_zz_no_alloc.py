"""The no-allocation backward-error entry point against bpf_backward_error,
through the C ABI: two calls in a row, the second on a smaller system with
fewer right-hand sides (the scratch keeps the first call's contents)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(bpf, A, X, B):
    out = bpf.HipMatrix((B.ncols,), np.float64, A.ctx)
    bpf.check(bpf.L.bpf_backward_error_noalloc(A.ctx.handle, A.handle, B.ncols, X.ptr, X.nrows, B.ptr,
                                               B.nrows, out.ptr))
    A.ctx.synchronize()
    return out.to_numpy()


def _case(bpf, gen, nbus, p, seed, solved):
    _, S = gen.make_system(nbus, p, seed=seed)
    A = bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data)
    B = bpf.HipMatrix.from_numpy(S.B)
    if solved:
        P = bpf.ilu0(A, p)
        X, st = bpf.block_gmres(A, B, N=P, ldiv=True, memory=40, berr_tol=1e-12, itmax=500)
        assert st.solved
    else:
        X = bpf.HipMatrix.from_numpy(np.random.default_rng(seed).standard_normal(S.B.shape))
    return A, X, B


@pytest.mark.parametrize("solved", [False, True])
def test_noalloc_backward_error_matches(bpf, gen, solved):
    for nbus, p, seed in ((2000, 96, 61), (300, 7, 62)):
        A, X, B = _case(bpf, gen, nbus, p, seed, solved)
        ref = bpf.backward_error(A, X, B)
        got = _noalloc(bpf, A, X, B)
        # same quantities, residual rows summed in another order: within
        # rounding of the residual, relative 1e-12 or 1e-14 of the denominator
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-14)
        assert got.shape == (p,)


def test_noalloc_rejects_too_many_rhs(bpf, gen):
    _, S = gen.make_system(9, 1, seed=63)
    A = bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data)
    X = bpf.HipMatrix((S.n, 4097), np.float64, A.ctx)
    out = bpf.HipMatrix((4097,), np.float64, A.ctx)
    rc = bpf.L.bpf_backward_error_noalloc(A.ctx.handle, A.handle, 4097, X.ptr, S.n, X.ptr, S.n, out.ptr)
    assert rc == bpf._lib.BPF_E_UNSUPPORTED
    assert b"4096" in bpf.L.bpf_last_error()
