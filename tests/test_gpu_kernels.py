"""GPU parity of the sparse kernels against the oracle (through the C ABI)."""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _system(gen, nbus, p, seed):
    grid, S = gen.make_system(nbus, p, seed=seed)
    return grid, S


@pytest.mark.parametrize("nbus,p", [(300, 1), (300, 16), (300, 37), (2000, 64), (2000, 128), (2000, 200)])
def test_spmm_colmajor_f64(bpf, gen, oracle, nbus, p):
    _, S = _system(gen, nbus, p, 11)
    A = bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data)
    X = bpf.HipMatrix.from_numpy(S.B)
    Y = bpf.HipMatrix((S.n, p))
    bpf.mul_(Y, A, X)
    ref = oracle.spmm(S.indptr, S.indices, S.data, S.B)
    got = Y.to_numpy()
    # Same summation order (row-sequential FMA chain) -> equal to rounding of fma.
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())


@pytest.mark.parametrize("p", [1, 8, 16, 64, 128, 256])
def test_spmm_rowmajor(bpf, gen, oracle, p):
    _, S = _system(gen, 2000, p, 12)
    A = bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data)
    ref = oracle.spmm(S.indptr, S.indices, S.data, S.B)
    ctx = A.ctx
    # row-major n x p == column-major p x n
    Xr = bpf.HipMatrix.from_numpy(np.asfortranarray(S.B.T))
    Yr = bpf.HipMatrix((p, S.n))
    bpf._lib.check(bpf.L.bpf_spmm(ctx.handle, A.handle, 0, 1, p, Xr.ptr, p, Yr.ptr, p, 1.0, 0.0))
    np.testing.assert_allclose(Yr.to_numpy().T, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
    # FP32: within gamma_k |A||X| with k = max row length
    Xf = bpf.HipMatrix.from_numpy(np.asfortranarray(S.B.T), dtype=np.float32)
    Yf = bpf.HipMatrix((p, S.n), np.float32)
    bpf._lib.check(bpf.L.bpf_spmm(ctx.handle, A.handle, 1, 1, p, Xf.ptr, p, Yf.ptr, p, 1.0, 0.0))
    Aabs = sp.csr_matrix((np.abs(S.data), S.indices, S.indptr), shape=(S.n, S.n))
    kmax = int(np.diff(S.indptr).max())
    bound = (kmax + 4) * np.finfo(np.float32).eps * (Aabs @ np.abs(S.B))
    assert np.all(np.abs(Yf.to_numpy().T.astype(np.float64) - ref) <= bound + 1e-30)


def test_spmm_alpha_beta(bpf, gen, oracle):
    _, S = _system(gen, 300, 5, 13)
    A = bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data)
    X = bpf.HipMatrix.from_numpy(S.B)
    Y0 = np.asfortranarray(np.random.default_rng(0).standard_normal(S.B.shape))
    Y = bpf.HipMatrix.from_numpy(Y0)
    bpf.mul_(Y, A, X, -2.0, 0.5)
    ref = -2.0 * oracle.spmm(S.indptr, S.indices, S.data, S.B) + 0.5 * Y0
    np.testing.assert_allclose(Y.to_numpy(), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("nbus", [300, 2000, 9000])
def test_ilu0_factor_matches_oracle(bpf, gen, oracle, nbus):
    _, S = _system(gen, nbus, 1, 21)
    A = bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data)
    P = bpf.ilu0(A)
    got = P.factor_values()
    ref = oracle.ilu0(S.indptr, S.indices, S.data)
    # Same elimination order per row (IKJ); only FMA contraction differs.
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-12 * np.abs(ref).max())
    nl, nu = P.levels()
    assert nl >= 1 and nu >= 1


@pytest.mark.parametrize("p", [1, 16, 64, 128])
def test_ilu0_apply_matches_oracle(bpf, gen, oracle, p):
    _, S = _system(gen, 2000, p, 22)
    A = bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data)
    P = bpf.ilu0(A, p)
    X = bpf.HipMatrix.from_numpy(S.B)
    Y = bpf.HipMatrix(S.B.shape)
    bpf.ldiv_(Y, P, X)
    F = oracle.ilu0(S.indptr, S.indices, S.data)
    ref = oracle.ilu0_apply(S.indptr, S.indices, F, S.B)
    np.testing.assert_allclose(Y.to_numpy(), ref, rtol=1e-10, atol=1e-11 * np.abs(ref).max())


def _spd(n, seed):
    """Graph Laplacian of a random sparse graph + diagonal (SPD, DC power-flow like)."""
    rng = np.random.default_rng(seed)
    m = 3 * n
    i = rng.integers(0, n, m)
    j = rng.integers(0, n, m)
    keep = i != j
    i, j = i[keep], j[keep]
    w = rng.uniform(1, 10, len(i))
    W = sp.coo_matrix((w, (i, j)), shape=(n, n)).tocsr()
    W = W + W.T
    Lap = sp.diags(np.asarray(W.sum(axis=1)).ravel()) - W + sp.diags(rng.uniform(0.1, 1.0, n))
    Lap = sp.csr_matrix(Lap)
    Lap.sort_indices()
    return Lap


@pytest.mark.parametrize("n,p", [(500, 1), (3000, 32)])
def test_ic0_factor_and_apply(bpf, oracle, n, p):
    M = _spd(n, 5)
    A = bpf.HipSparseMatrixCSR(M.indptr, M.indices, M.data)
    P = bpf.ic0(A, p)
    got = P.factor_values()
    ref = oracle.ic0(M.indptr, M.indices, M.data)
    lower = np.zeros(len(M.data), dtype=bool)
    for r in range(n):
        lower[M.indptr[r]:M.indptr[r + 1]] = M.indices[M.indptr[r]:M.indptr[r + 1]] <= r
    np.testing.assert_allclose(got[lower], ref[lower], rtol=1e-11, atol=1e-12)
    B = np.asfortranarray(np.random.default_rng(1).standard_normal((n, p)))
    X = bpf.HipMatrix.from_numpy(B)
    Y = bpf.HipMatrix(B.shape)
    bpf.ldiv_(Y, P, X)
    refy = oracle.ic0_apply(M.indptr, M.indices, ref, B)
    np.testing.assert_allclose(Y.to_numpy(), refy, rtol=1e-10, atol=1e-11 * np.abs(refy).max())


def test_ilu0_refactor_new_values(bpf, gen, oracle):
    _, S = _system(gen, 2000, 1, 23)
    A = bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data)
    P = bpf.ilu0(A)
    vals = S.data * 2.0
    A.set_values(vals)
    P.factor_(A)
    ref = oracle.ilu0(S.indptr, S.indices, vals)
    np.testing.assert_allclose(P.factor_values(), ref, rtol=1e-11, atol=1e-12 * np.abs(ref).max())


def test_backward_error_matches_host(bpf, gen, oracle):
    _, S = _system(gen, 300, 7, 24)
    A = bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data)
    Xh = np.asfortranarray(np.random.default_rng(3).standard_normal(S.B.shape))
    got = bpf.backward_error(A, bpf.HipMatrix.from_numpy(Xh), bpf.HipMatrix.from_numpy(S.B))
    ref = oracle.backward_error(S.indptr, S.indices, S.data, Xh, S.B)
    np.testing.assert_allclose(got, ref, rtol=1e-13)


@pytest.mark.parametrize("p", [16, 128])
def test_syncfree_matches_level_schedule(bpf, gen, oracle, p):
    # The optional single-launch sync-free factorization computes every row
    # with the same operation order as the level-by-level launches (bitwise);
    # the solves agree to rounding (the narrow-level kernel splits long rows
    # over several wavefronts, which reorders their sums).
    _, S = _system(gen, 2000, p, 23)
    A = bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data)
    X = bpf.HipMatrix.from_numpy(S.B)
    out = {}
    try:
        for sf in (0, 1):
            bpf.set_option("syncfree", sf)
            P = bpf.ilu0(A, p)
            Y = bpf.HipMatrix(S.B.shape)
            bpf.ldiv_(Y, P, X)
            out[sf] = (P.factor_values(), Y.to_numpy())
    finally:
        bpf.set_option("syncfree", 0)
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_allclose(out[0][1], out[1][1], rtol=1e-13, atol=1e-15 * np.abs(out[0][1]).max())
    F = oracle.ilu0(S.indptr, S.indices, S.data)
    ref = oracle.ilu0_apply(S.indptr, S.indices, F, S.B)
    np.testing.assert_allclose(out[1][1], ref, rtol=1e-10, atol=1e-11 * np.abs(ref).max())


@pytest.mark.parametrize("p", [8, 64, 128, 256])
def test_spmm_autotuned_shape_is_exact(bpf, gen, oracle, p):
    # Every SpMM launch shape sums each row's products in the same order, so
    # the autotuner's choice cannot change the result.
    _, S = _system(gen, 2000, p, 14)
    A = bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data)
    ctx = A.ctx
    Xr = bpf.HipMatrix.from_numpy(np.asfortranarray(S.B.T))
    out = {}
    try:
        for tune in (0, 1):
            bpf.set_option("autotune", tune)
            Yr = bpf.HipMatrix((p, S.n))
            bpf._lib.check(bpf.L.bpf_spmm(ctx.handle, A.handle, 0, 1, p, Xr.ptr, p, Yr.ptr, p, 1.0, 0.0))
            out[tune] = Yr.to_numpy()
    finally:
        bpf.set_option("autotune", 1)
    np.testing.assert_array_equal(out[0], out[1])
    ref = oracle.spmm(S.indptr, S.indices, S.data, S.B)
    np.testing.assert_allclose(out[1].T, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())


@pytest.mark.parametrize("p", [1, 16, 128, 200])
def test_fp32_triangular_solve_backward_error(bpf, gen, p):
    # The FP32 level-scheduled triangular solve (the mixed form's inner
    # preconditioner and the refactorization's FP32 solves) has a
    # componentwise backward error of a few FP32 roundings per row.  On a
    # lower-triangular A, ILU(0) is exact (L = A D^-1, U = D), so its FP32
    # application y satisfies |x - A y| <= (k + 4) eps32 |A| |y| (k the
    # longest row, x the FP32 input; residual in FP64).  A lost or reordered
    # term, a wrong level order or a wrong diagonal breaks it by orders of
    # magnitude.
    _, S = _system(gen, 2000, p, 24)
    T = sp.tril(S.scipy()).tocsr()
    T.sort_indices()
    A = bpf.HipSparseMatrixCSR(T.indptr.astype(np.int32), T.indices.astype(np.int32), T.data)
    P = bpf.ilu0(A, p)
    x32 = np.asfortranarray(S.B.astype(np.float32))
    X = bpf.HipMatrix.from_numpy(x32, dtype=np.float32)
    Y = bpf.HipMatrix(S.B.shape, np.float32)
    bpf.ldiv_(Y, P, X)
    y = Y.to_numpy().astype(np.float64)
    r = x32.astype(np.float64) - T @ y
    kmax = int(np.diff(T.indptr).max())
    bound = (kmax + 4) * np.finfo(np.float32).eps * (abs(T) @ np.abs(y))
    assert np.all(np.abs(r) <= bound + 1e-300), np.max(np.abs(r) / (bound + 1e-300))
    # and the FP64 application of the same factor solves T y = x to FP64 rounding
    X64 = bpf.HipMatrix.from_numpy(x32.astype(np.float64))
    Y64 = bpf.HipMatrix(S.B.shape)
    bpf.ldiv_(Y64, P, X64)
    y64 = Y64.to_numpy()
    r64 = x32.astype(np.float64) - T @ y64
    assert np.all(np.abs(r64) <= (kmax + 4) * np.finfo(np.float64).eps * (abs(T) @ np.abs(y64)) + 1e-300)
