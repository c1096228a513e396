"""GPU refactorization path (CUSOLVERRF replacement) against the oracle.

Replays the reference's own known-answer tests (test/cusolverRF.jl) on a
synthetic 300-bus Jacobian, checks the device factor against the oracle's
fixed-pattern refactorization, and runs the multi-RHS Newton sequence."""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _sys(gen, nbus, p, seed):
    grid, S = gen.make_system(nbus, p, seed=seed)
    return grid, S


@pytest.mark.parametrize("ordering", [":AMD", ":MDQ", ":METIS", ":RCM"])
def test_rf_known_answers(bpf, gen, oracle, ordering):
    _, S = _sys(gen, 300, 1, 41)
    A = bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data)
    b = np.random.default_rng(0).random(S.n)
    sol = oracle.direct_solve(S.indptr, S.indices, S.data, b)          # J \ b (:12)
    gb = bpf.HipMatrix.from_numpy(b.reshape(-1, 1))
    x = bpf.HipMatrix((S.n, 1))
    rf = bpf.CUSOLVERRF.CusolverRfLU(A, ordering=ordering)
    bpf.ldiv_(x, rf, gb)
    np.testing.assert_allclose(x.to_numpy().ravel(), sol, rtol=1.5e-8)   # isapprox (:21)
    A.set_values(2.0 * S.data)                                          # nzVal .*= 2 (:24)
    bpf.CUSOLVERRF.update_(rf, A)
    bpf.ldiv_(x, rf, gb)
    np.testing.assert_allclose(x.to_numpy().ravel(), 0.5 * sol, rtol=1.5e-8)   # (:28)
    bpf.CUSOLVERRF.cudestroy_(rf)


def test_rf_factor_matches_oracle_refactor(bpf, gen, oracle):
    _, S = _sys(gen, 300, 1, 42)
    A = bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data)
    rf = bpf.CUSOLVERRF.lu(A)
    P, Q, rp, ci, v = rf.factors()
    assert sorted(P) == list(range(S.n)) and sorted(Q) == list(range(S.n))
    ref = oracle.lu_refactor(S.indptr, S.indices, S.data, P, Q, rp, ci)
    np.testing.assert_allclose(v, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())
    # L U = A[P, Q]
    F = sp.csr_matrix((v, ci, rp), shape=(S.n, S.n))
    Lm = sp.tril(F, -1) + sp.eye(S.n)
    U = sp.triu(F)
    B = S.scipy()[P][:, Q]
    assert abs(Lm @ U - B).max() <= 1e-10 * abs(B).max()
    # new values on the same pattern: device refactor == oracle refactor
    vals = S.data * (1.0 + 0.01 * np.random.default_rng(1).standard_normal(S.nnz))
    A.set_values(vals)
    bpf.CUSOLVERRF.lu_(rf, A)
    _, _, _, _, v2 = rf.factors()
    ref2 = oracle.lu_refactor(S.indptr, S.indices, vals, P, Q, rp, ci)
    np.testing.assert_allclose(v2, ref2, rtol=1e-10, atol=1e-12 * np.abs(ref2).max())


@pytest.mark.parametrize("nbus,p", [(300, 1), (2000, 64), (9000, 256)])
def test_rf_multi_rhs(bpf, gen, oracle, nbus, p):
    _, S = _sys(gen, nbus, p, 43)
    A = bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data)
    rf = bpf.CUSOLVERRF.CusolverRfLU(A, nrhs=p)
    B = bpf.HipMatrix.from_numpy(S.B)
    X = bpf.HipMatrix(S.B.shape)
    bpf.ldiv_(X, rf, B)
    Xd = oracle.direct_solve(S.indptr, S.indices, S.data, S.B)
    np.testing.assert_allclose(X.to_numpy(), Xd, rtol=1e-8, atol=1e-10 * np.abs(Xd).max())
    # in place
    bpf.ldiv_(rf, B)
    np.testing.assert_allclose(B.to_numpy(), Xd, rtol=1e-8, atol=1e-10 * np.abs(Xd).max())


def test_rf_refined_newton_sequence(bpf, gen, oracle):
    grid, S = gen.make_system(2000, 32, seed=44)
    A = bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data)
    rf = bpf.CUSOLVERRF.CusolverRfLU(A)
    B = bpf.HipMatrix.from_numpy(S.B)
    for vals in gen.newton_value_sets(grid, 4, seed=45):
        A.set_values(vals)
        bpf.CUSOLVERRF.update_(rf, A)
        X, berr = rf.solve_refined(A, B, refine=2)
        assert berr.max() <= 1e-10, berr.max()
        assert rf.refine_steps == 2
        ref = oracle.backward_error(S.indptr, S.indices, vals, X.to_numpy(), S.B)
        # both at rounding level: compare against the target, not each other
        assert ref.max() <= 1e-10
        np.testing.assert_allclose(berr, ref, rtol=0, atol=1e-14)
        Xd = oracle.direct_solve(S.indptr, S.indices, vals, S.B)
        np.testing.assert_allclose(X.to_numpy(), Xd, rtol=1e-7, atol=1e-9 * np.abs(Xd).max())


def test_rf_batch(bpf, gen, oracle):
    # test/cusolverRF.jl:34-68: batch of 32, RHS i*b -> i*(J\b)
    _, S = _sys(gen, 300, 1, 46)
    A = bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data)
    nbatch = 32
    rfb = bpf.CUSOLVERRF.CusolverRfLUBatch(A, nbatch)
    b = np.random.default_rng(2).random(S.n)
    sol = oracle.direct_solve(S.indptr, S.indices, S.data, b)
    gB = [bpf.HipMatrix.from_numpy((i + 1) * b.reshape(-1, 1)) for i in range(nbatch)]
    gX = [bpf.HipMatrix((S.n, 1)) for _ in range(nbatch)]
    bpf.CUSOLVERRF.solve_(gX, rfb, gB)
    for i in range(nbatch):
        np.testing.assert_allclose(gX[i].to_numpy().ravel(), (i + 1) * sol, rtol=1.5e-8)
    bpf.CUSOLVERRF.update_(rfb, A)
    Bm = bpf.HipMatrix.from_numpy(np.column_stack([(i + 1) * b for i in range(nbatch)]))
    Xm = bpf.HipMatrix((S.n, nbatch))
    bpf.ldiv_(Xm, rfb, Bm)
    np.testing.assert_allclose(Xm.to_numpy(), np.outer(sol, np.arange(1, nbatch + 1)), rtol=1.5e-8)
    # rf_batch_refactor!: one value set per batch entry
    scales = np.linspace(1.0, 2.0, nbatch)
    rf_vals = np.column_stack([s * S.data for s in scales])
    bpf.CUSOLVERRF.rf_batch_refactor_(rfb, rf_vals)
    bpf.ldiv_(Xm, rfb, Bm)
    expect = np.outer(sol, np.arange(1, nbatch + 1) / scales)
    np.testing.assert_allclose(Xm.to_numpy(), expect, rtol=1.5e-8)


def test_rf_batchsize_one_rejected(bpf, gen):
    _, S = _sys(gen, 300, 1, 47)
    A = bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data)
    with pytest.raises(ValueError):
        bpf.CUSOLVERRF.CusolverRfLUBatch(A, 1)


def test_rf_singular_raises(bpf):
    n = 6
    M = sp.csr_matrix(np.diag([1.0, 2.0, 0.0, 4.0, 5.0, 6.0]) + np.diag(np.ones(n - 1), 1) * 0.0)
    M = sp.csr_matrix(M + sp.csr_matrix(([1.0], ([2], [3])), shape=(n, n)))
    M.eliminate_zeros()
    # row 2 has only an entry in column 3 and column 2 is empty: structurally singular
    M.sort_indices()
    A = bpf.HipSparseMatrixCSR(M.indptr, M.indices, M.data)
    with pytest.raises(bpf.SingularException):
        bpf.CUSOLVERRF.CusolverRfLU(A)


def test_rf_dense_tail_matches_oracle(bpf, gen, oracle):
    # n >= 16384: the last rows of B = A[P, Q] are refactored and solved as a
    # dense block (blocked LU without pivoting on matrix cores); the factor
    # values must equal the fixed-pattern sparse refactorization's.
    grid, S = gen.make_system(10000, 64, seed=46)
    assert S.n >= 16384
    A = bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data)
    rf = bpf.CUSOLVERRF.CusolverRfLU(A, nrhs=64)
    P, Q, rp, ci, v = rf.factors()
    ref = oracle.lu_refactor(S.indptr, S.indices, S.data, P, Q, rp, ci)
    np.testing.assert_allclose(v, ref, rtol=1e-9, atol=1e-11 * np.abs(ref).max())
    B = bpf.HipMatrix.from_numpy(S.B)
    X = bpf.HipMatrix(S.B.shape)
    bpf.ldiv_(X, rf, B)
    Xd = oracle.direct_solve(S.indptr, S.indices, S.data, S.B)
    np.testing.assert_allclose(X.to_numpy(), Xd, rtol=1e-8, atol=1e-10 * np.abs(Xd).max())
    # Newton value set on the same pattern; FP32 solves + 2 FP64 refinements
    vals = gen.newton_value_sets(grid, 1, seed=47)[0]
    A.set_values(vals)
    bpf.CUSOLVERRF.update_(rf, A)
    _, _, _, _, v2 = rf.factors()
    ref2 = oracle.lu_refactor(S.indptr, S.indices, vals, P, Q, rp, ci)
    np.testing.assert_allclose(v2, ref2, rtol=1e-9, atol=1e-11 * np.abs(ref2).max())
    Xr, berr = rf.solve_refined(A, B, refine=2)
    assert berr.max() <= 1e-10
    ref_x = oracle.direct_solve(S.indptr, S.indices, vals, S.B)
    np.testing.assert_allclose(Xr.to_numpy(), ref_x, rtol=1e-7, atol=1e-9 * np.abs(ref_x).max())


def test_rf_refined_stops_at_tolerance(bpf, gen, oracle):
    # refinement with a backward-error target: stops at the first step whose
    # residual meets it, and X is bitwise the fixed-step result of that many steps
    grid, S = gen.make_system(2000, 24, seed=48)
    A = bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data)
    rf = bpf.CUSOLVERRF.CusolverRfLU(A)
    B = bpf.HipMatrix.from_numpy(S.B)
    for vals in gen.newton_value_sets(grid, 3, seed=49):
        A.set_values(vals)
        bpf.CUSOLVERRF.update_(rf, A)
        X, berr = rf.solve_refined(A, B, refine=4, berr_tol=1e-10)
        k = rf.refine_steps
        assert 1 <= k <= 4 and berr.max() <= 1e-10, (k, berr.max())
        ref = oracle.backward_error(S.indptr, S.indices, vals, X.to_numpy(), S.B)
        np.testing.assert_allclose(berr, ref, rtol=0, atol=1e-14)
        Xk, berr_k = rf.solve_refined(A, B, refine=k)
        assert rf.refine_steps == k
        np.testing.assert_array_equal(X.to_numpy(), Xk.to_numpy())
        np.testing.assert_array_equal(berr, berr_k)
        if k > 1:
            _, berr_less = rf.solve_refined(A, B, refine=k - 1)
            assert berr_less.max() > 1e-10
        # a tolerance no column meets runs every step
        _, berr_all = rf.solve_refined(A, B, refine=3, berr_tol=1e-300)
        assert rf.refine_steps == 3


def _families(bpf, ctx):
    import ctypes as C
    out = []
    for f in range(6):
        n = C.c_longlong()
        ms, by, fl = C.c_double(), C.c_double(), C.c_double()
        bpf._lib.check(bpf.L.bpf_profile_read(ctx.handle, f, C.byref(n), C.byref(ms), C.byref(by), C.byref(fl)))
        out.append((n.value, ms.value, by.value, fl.value))
    return out


def test_kernel_timers_cover_factor_and_dense_tail(bpf, gen):
    # the in-process timers count one sparse-factor launch for the head rows
    # and one for the tail rows, and the dense-tail LU's factorization and
    # solves, with their compulsory bytes / flops
    grid, S = gen.make_system(10000, 16, seed=51)
    assert S.n >= 16384
    A = bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data)
    rf = bpf.CUSOLVERRF.CusolverRfLU(A, nrhs=16)
    B = bpf.HipMatrix.from_numpy(S.B)
    ctx = A.ctx
    bpf._lib.check(bpf.L.bpf_profile_enable(ctx.handle, 1))
    try:
        bpf.CUSOLVERRF.update_(rf, A)
        _, berr = rf.solve_refined(A, B, refine=2)
        fam = _families(bpf, ctx)
    finally:
        bpf._lib.check(bpf.L.bpf_profile_enable(ctx.handle, 0))
    assert berr.max() <= 1e-10
    P, Q, rp, ci, v = rf.factors()
    n_fac, ms_fac, by_fac, _ = fam[4]
    assert n_fac == 2 and ms_fac > 0
    # every factor entry read and written once (8 + 8 bytes) plus its column index
    # (4), and three 4-byte row entries per row
    assert by_fac == pytest.approx(20.0 * len(v) + 12.0 * S.n)
    n_dlu, ms_dlu, by_dlu, fl_dlu = fam[5]
    assert n_dlu == 1 + 3 and ms_dlu > 0 and by_dlu > 0 and fl_dlu > 0   # factor + three solves
    assert fam[1][0] > 0   # level-scheduled triangular solves


@pytest.mark.parametrize("p", [100, 256])
def test_tail_slabs_are_exact(bpf, gen, oracle, p):
    # The narrow-level part of a triangular solve over more than 64
    # right-hand sides runs as one workgroup per column slab: every column's
    # arithmetic is the same as with one workgroup for all of them, so FP64
    # and FP32 (refined) LU solves and ILU(0) applications are bitwise equal
    # whatever the slab width, and the FP64 solve is the direct solution.
    grid, S = gen.make_system(10000, p, seed=52)
    assert S.n >= 16384
    A = bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data)
    rf = bpf.CUSOLVERRF.CusolverRfLU(A, nrhs=p)
    Pi = bpf.ilu0(A, p)
    B = bpf.HipMatrix.from_numpy(S.B)
    old = bpf.get_option("tail_slab")
    out = {}
    try:
        for slab in (0, 64):
            bpf.set_option("tail_slab", slab)
            X = bpf.HipMatrix(S.B.shape)
            bpf.ldiv_(X, rf, B)
            Xr, berr = rf.solve_refined(A, B, refine=1)
            Y = bpf.HipMatrix(S.B.shape)
            bpf.ldiv_(Y, Pi, B)
            out[slab] = (X.to_numpy(), Xr.to_numpy(), berr, Y.to_numpy())
    finally:
        bpf.set_option("tail_slab", old)
    for a, b in zip(out[0], out[64]):
        np.testing.assert_array_equal(a, b)
    Xd = oracle.direct_solve(S.indptr, S.indices, S.data, S.B)
    np.testing.assert_allclose(out[64][0], Xd, rtol=1e-8, atol=1e-10 * np.abs(Xd).max())


def test_dense_tail_inverse_factors(bpf, gen, oracle):
    # The dense tail solved by explicit inverses of its triangular factors
    # (recursive doubling on matrix cores, by the first solve that uses them
    # after a refactorization): with the option at 2 every solve uses them and the FP64 solve is the direct
    # solution at a backward error of 1e-10 or better; at 1 only the FP32
    # solves do (the FP64 solve is bitwise the blocked substitution's); the
    # refined solutions meet the tolerance in every mode.
    grid, S = gen.make_system(10000, 48, seed=53)
    assert S.n >= 16384
    A = bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data)
    rf = bpf.CUSOLVERRF.CusolverRfLU(A, nrhs=48)
    B = bpf.HipMatrix.from_numpy(S.B)
    vals = gen.newton_value_sets(grid, 1, seed=54)[0]
    A.set_values(vals)
    Xd = oracle.direct_solve(S.indptr, S.indices, vals, S.B)
    old = bpf.get_option("dense_inverse")
    out = {}
    try:
        for mode in (0, 1, 2):
            bpf.set_option("dense_inverse", mode)
            bpf.CUSOLVERRF.update_(rf, A)
            X = bpf.HipMatrix(S.B.shape)
            bpf.ldiv_(X, rf, B)
            Xr, berr = rf.solve_refined(A, B, refine=2, berr_tol=1e-10)
            out[mode] = (X.to_numpy(), Xr.to_numpy(), berr)
    finally:
        bpf.set_option("dense_inverse", old)
    np.testing.assert_array_equal(out[0][0], out[1][0])
    for mode in (0, 1, 2):
        X, Xr, berr = out[mode]
        assert berr.max() <= 1e-10, (mode, berr.max())
        np.testing.assert_allclose(Xr, Xd, rtol=1e-7, atol=1e-9 * np.abs(Xd).max())
        np.testing.assert_allclose(X, Xd, rtol=1e-8, atol=1e-10 * np.abs(Xd).max())
        be = oracle.backward_error(S.indptr, S.indices, vals, X, S.B)
        assert be.max() <= 1e-10, (mode, be.max())
