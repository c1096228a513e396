"""GPU block solvers against the FP64 sparse direct oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(bpf, gen, oracle, nbus, p, seed):
    _, S = gen.make_system(nbus, p, seed=seed)
    A = bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data)
    B = bpf.HipMatrix.from_numpy(S.B)
    Xd = oracle.direct_solve(S.indptr, S.indices, S.data, S.B)
    return S, A, B, Xd


def _rel(X, Xd):
    return np.abs(X - Xd).max(axis=0) / np.abs(Xd).max(axis=0)


@pytest.mark.parametrize("p", [1, 16, 32])
def test_block_gmres_fp64_reference_criterion(bpf, gen, oracle, p):
    S, A, B, Xd = _setup(bpf, gen, oracle, 300, p, 31)
    P = bpf.ilu0(A, p)
    X, st = bpf.block_gmres(A, B, N=P, ldiv=True, memory=30, atol=0.0, rtol=1e-10, history=True)
    Xh = X.to_numpy()
    assert st.solved, st
    assert st.status == "solution good enough given atol and rtol"
    r = np.linalg.norm(S.B - S.scipy() @ Xh)
    assert r <= 1e-10 * np.linalg.norm(S.B) * 10
    assert len(st.residuals) == st.niter + 1
    assert np.all(_rel(Xh, Xd) < 1e-6)


def test_block_gmres_unpreconditioned_finite_termination(bpf, gen, oracle):
    # Without restart and preconditioner, block GMRES terminates in ceil(n/p) steps.
    S, A, B, Xd = _setup(bpf, gen, oracle, 9, 2, 32)
    n = S.n
    X, st = bpf.block_gmres(A, B, memory=100, atol=0.0, rtol=1e-12, itmax=(n + 1) // 2)
    assert st.niter <= (n + 1) // 2
    np.testing.assert_allclose(X.to_numpy(), Xd, rtol=1e-7, atol=1e-9 * np.abs(Xd).max())


@pytest.mark.parametrize("nbus,p", [(300, 16), (2000, 64), (9000, 64)])
def test_block_gmres_fp64_backward_error(bpf, gen, oracle, nbus, p):
    S, A, B, Xd = _setup(bpf, gen, oracle, nbus, p, 33)
    P = bpf.ilu0(A, p)
    X, st = bpf.block_gmres(A, B, N=P, ldiv=True, memory=30, berr_tol=1e-10, itmax=500)
    Xh = X.to_numpy()
    be = oracle.backward_error(S.indptr, S.indices, S.data, Xh, S.B)
    assert st.solved, st
    assert be.max() <= 1e-10
    assert np.all(_rel(Xh, Xd) < 1e-5)


@pytest.mark.parametrize("nbus,p,s,prec,deg", [
    (300, 16, 0, "mixed", 0), (2000, 64, 32, "mixed", 8), (9000, 128, 0, "mixed", 8),
    (300, 16, 0, "mixed32", 0), (2000, 64, 32, "mixed32", 4), (9000, 128, 0, "mixed32", 0)])
def test_block_gmres_mixed(bpf, gen, oracle, nbus, p, s, prec, deg):
    S, A, B, Xd = _setup(bpf, gen, oracle, nbus, p, 34)
    P = bpf.ilu0(A, p)
    X, st = bpf.block_gmres(A, B, N=P, ldiv=True, memory=40, precision=prec, berr_tol=1e-10,
                            block_width=s, itmax=1000, poly_degree=deg)
    Xh = X.to_numpy()
    be = oracle.backward_error(S.indptr, S.indices, S.data, Xh, S.B)
    assert st.solved, (st, be.max())
    assert be.max() <= 1e-10
    if prec == "mixed32":
        assert st.refine_steps >= 1
    assert np.all(_rel(Xh, Xd) < 1e-5)


@pytest.mark.parametrize("nbus,p,s,prec,g", [(2000, 64, 64, "fp64", 2), (9000, 128, 0, "mixed", 4),
                                              (2000, 128, 64, "fp64", 3), (300, 48, 0, "mixed", 8)])
def test_grouped_bmgs_matches_one_block_steps(bpf, gen, oracle, nbus, p, s, prec, g):
    # Projecting against g basis blocks per block-MGS step is the same
    # Gram-Schmidt in exact arithmetic (the blocks are orthonormal): the
    # residual histories agree to rounding, and so do the solutions.
    S, A, B, Xd = _setup(bpf, gen, oracle, nbus, p, 35)
    P = bpf.ilu0(A, p)
    out = {}
    for gg in (1, g):
        X, st = bpf.block_gmres(A, B, N=P, ldiv=True, memory=30, atol=0.0, rtol=1e-9, itmax=40,
                                history=True, block_width=s, precision=prec, bmgs_blocks=gg)
        out[gg] = (X.to_numpy(), np.array(st.residuals), st.niter)
    assert out[1][2] == out[g][2]
    # Residual norms agree to rounding relative to the initial residual.  In
    # the mixed form the FP32 preconditioner sees the rounding difference of
    # its FP64 input, so the histories agree to FP32 level there.
    rt = 1e-6 if prec == "fp64" else 1e-4
    np.testing.assert_allclose(out[g][1], out[1][1], rtol=rt, atol=1e-12 * out[1][1][0])
    np.testing.assert_allclose(out[g][0], out[1][0], rtol=rt, atol=1e-8 * np.abs(out[1][0]).max())
    be = oracle.backward_error(S.indptr, S.indices, S.data, out[g][0], S.B)
    assert be.max() <= 1e-6


@pytest.mark.parametrize("nbus,p,s,prec,restart", [(2000, 64, 32, "fp64", False), (9000, 128, 0, "mixed", False),
                                                   (300, 16, 0, "fp64", True)])
def test_overlapped_hessenberg_update_is_exact(bpf, gen, oracle, nbus, p, s, prec, restart):
    # The Hessenberg update on the second stream runs the same kernels on the
    # same data as the one-iteration-after-another order: bitwise equal
    # solutions, histories and iteration counts.
    S, A, B, Xd = _setup(bpf, gen, oracle, nbus, p, 36)
    P = bpf.ilu0(A, p)
    out = {}
    try:
        for ovl in (0, 1):
            bpf.set_option("overlap", ovl)
            X, st = bpf.block_gmres(A, B, N=P, ldiv=True, memory=12 if restart else 30, restart=restart,
                                    atol=0.0, rtol=1e-9, itmax=60, history=True, block_width=s,
                                    precision=prec, berr_tol=1e-10 if prec == "mixed" else 0.0)
            out[ovl] = (X.to_numpy(), np.array(st.residuals), st.niter, st.solved)
    finally:
        bpf.set_option("overlap", 1)
    assert out[0][2] == out[1][2] and out[0][3] == out[1][3]
    np.testing.assert_array_equal(out[1][1], out[0][1])
    np.testing.assert_array_equal(out[1][0], out[0][0])


def test_block_gmres_restart_and_warm_start(bpf, gen, oracle):
    S, A, B, Xd = _setup(bpf, gen, oracle, 300, 8, 35)
    P = bpf.ilu0(A)
    X, st = bpf.block_gmres(A, B, N=P, ldiv=True, memory=20, restart=True, atol=0.0, rtol=1e-9,
                            itmax=400)
    assert st.solved
    X0 = bpf.HipMatrix.from_numpy(Xd * (1 + 1e-3))
    X2, st2 = bpf.block_gmres(A, B, X0, N=P, ldiv=True, memory=30, atol=0.0, rtol=1e-9)
    assert st2.solved
    assert st2.niter <= st.niter
    np.testing.assert_allclose(X2.to_numpy(), Xd, rtol=1e-5, atol=1e-7 * np.abs(Xd).max())


def test_block_gmres_itmax_status(bpf, gen, oracle):
    S, A, B, Xd = _setup(bpf, gen, oracle, 300, 4, 36)
    X, st = bpf.block_gmres(A, B, atol=0.0, rtol=1e-14, itmax=3)
    assert not st.solved
    assert st.niter == 3
    assert st.status == "maximum number of iterations exceeded"


def _diag_dominant(n, seed):
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    M = sp.random(n, n, density=6.0 / n, random_state=seed, format="csr")
    M = M + sp.diags(np.abs(M).sum(axis=1).A1 * 1.5 + 1.0)
    M = sp.csr_matrix(M)
    M.sort_indices()
    return M, np.asfortranarray(rng.standard_normal((n, 1)))


@pytest.mark.parametrize("p", [1, 8, 32])
def test_block_bicgstab_converges(bpf, oracle, p):
    import scipy.sparse.linalg as spla
    M, _ = _diag_dominant(2000, 7)
    Bh = np.asfortranarray(np.random.default_rng(p).standard_normal((M.shape[0], p)))
    A = bpf.HipSparseMatrixCSR(M.indptr, M.indices, M.data)
    B = bpf.HipMatrix.from_numpy(Bh)
    P = bpf.ilu0(A)
    X, st = bpf.block_bicgstab(A, B, N=P, atol=0.0, rtol=1e-10, itmax=500, history=True)
    Xh = X.to_numpy()
    assert st.solved, st
    Xd = spla.splu(M.tocsc()).solve(Bh)
    assert np.all(_rel(Xh, Xd) < 1e-8)
    assert abs(st.residuals[-1] - np.linalg.norm(Bh - M @ Xh)) <= 1e-8 * np.linalg.norm(Bh)


def test_block_bicgstab_power_flow_p1(bpf, gen, oracle):
    # On the power-flow Jacobian the reference recurrence converges for p = 1;
    # for p > 1 it diverges on CPU as well (checked against the restatement in
    # test_block_bicgstab_matches_cpu_restatement).
    S, A, B, Xd = _setup(bpf, gen, oracle, 300, 1, 37)
    P = bpf.ilu0(A)
    X, st = bpf.block_bicgstab(A, B, N=P, atol=0.0, rtol=1e-10, itmax=2000, history=True)
    assert st.solved, st
    assert np.all(_rel(X.to_numpy(), Xd) < 1e-6)


@pytest.mark.parametrize("p", [1, 4, 16])
def test_block_bicgstab_matches_cpu_restatement(bpf, gen, oracle, p):
    # Same preconditioner: the GPU recurrence and the CPU restatement of
    # src/block_bicgstab.jl follow the same iterates to rounding (first steps).
    S, A, B, Xd = _setup(bpf, gen, oracle, 300, p, 38)
    P = bpf.ilu0(A)
    X, st = bpf.block_bicgstab(A, B, N=P, atol=0.0, rtol=1e-8, itmax=200, history=True)
    F = oracle.ilu0(S.indptr, S.indices, S.data)
    Nf = lambda Z: oracle.ilu0_apply(S.indptr, S.indices, F, Z)
    Xc, stc = oracle.block_bicgstab(S.scipy(), S.B, N=Nf, atol=0.0, rtol=1e-8, itmax=200, history=True)
    k = min(len(st.residuals), len(stc.residuals), 8)
    np.testing.assert_allclose(st.residuals[:k], stc.residuals[:k], rtol=1e-6)


# ------------------------------------------------ left preconditioning (M)
@pytest.mark.parametrize("side", ["left", "both"])
def test_block_gmres_left_preconditioner_matches_oracle(bpf, gen, oracle, side):
    # block_gmres(A, B; M, N) (src/block_gmres.jl:194-203,226,300): the run
    # estimates ||M (B - A X)||_F; the oracle restates the same recurrence, so
    # the residual histories agree to rounding over the first iterations and
    # both reach the direct solution.
    S, A, B, Xd = _setup(bpf, gen, oracle, 300, 8, 41)
    P = bpf.ilu0(A, 8)
    F = oracle.ilu0(S.indptr, S.indices, S.data)
    apply = lambda Z: oracle.ilu0_apply(S.indptr, S.indices, F, Z)
    kw = dict(M=P) if side == "left" else dict(M=P, N=P)
    okw = dict(M=apply) if side == "left" else dict(M=apply, N=apply)
    X, st = bpf.block_gmres(A, B, memory=60, atol=0.0, rtol=1e-11, history=True, ldiv=True, **kw)
    Xo, sto = oracle.block_gmres(S.scipy(), S.B, memory=60, atol=0.0, rtol=1e-11, history=True, **okw)
    assert st.solved and sto.solved
    k = min(6, len(sto.residuals))
    np.testing.assert_allclose(st.residuals[:k], sto.residuals[:k], rtol=1e-7)
    assert abs(st.niter - sto.niter) <= 1
    # stopping rule on the M-weighted residual, as the reference
    MR = apply(S.B - S.scipy() @ X.to_numpy())
    assert np.linalg.norm(MR) <= 1e-11 * np.linalg.norm(apply(S.B)) * 10
    assert np.all(_rel(X.to_numpy(), Xd) < 1e-6)


def test_block_gmres_left_preconditioner_backward_error(bpf, gen, oracle):
    S, A, B, Xd = _setup(bpf, gen, oracle, 2000, 32, 42)
    P = bpf.ilu0(A, 32)
    X, st = bpf.block_gmres(A, B, M=P, memory=40, berr_tol=1e-10, itmax=500)
    be = oracle.backward_error(S.indptr, S.indices, S.data, X.to_numpy(), S.B)
    assert st.solved and be.max() <= 1e-10
    assert np.all(_rel(X.to_numpy(), Xd) < 1e-5)


# ------------------------------------------------------ Float32 blocks
def test_block_gmres_float32_blocks(bpf, gen, oracle):
    # The reference is generic over the element type of B (default tolerances
    # sqrt(eps(Float32))); X comes back as Float32.
    S, A, B, Xd = _setup(bpf, gen, oracle, 300, 16, 43)
    B32 = bpf.HipMatrix.from_numpy(S.B.astype(np.float32))
    P = bpf.ilu0(A, 16)
    X, st = bpf.block_gmres(A, B32, N=P, ldiv=True, memory=30)
    assert X.dtype == np.float32 and st.solved
    Xh = X.to_numpy().astype(np.float64)
    r = np.linalg.norm(S.B.astype(np.float32).astype(np.float64) - S.scipy() @ Xh)
    eps32 = np.sqrt(np.finfo(np.float32).eps)
    assert r <= 2 * (eps32 + eps32 * np.linalg.norm(S.B)) + 1e-6 * np.linalg.norm(S.B)
    # warm start with a Float32 guess, and an explicit tolerance
    X2, st2 = bpf.block_gmres(A, B32, X, N=P, ldiv=True, memory=30, atol=0.0, rtol=1e-3)
    assert X2.dtype == np.float32 and st2.solved
    r2 = np.linalg.norm(S.B.astype(np.float32).astype(np.float64) - S.scipy() @ X2.to_numpy().astype(np.float64))
    assert r2 <= 1.01 * r + 1e-6 * np.linalg.norm(S.B)
    with pytest.raises(TypeError):
        bpf.block_gmres(A, B32, bpf.HipMatrix.from_numpy(np.zeros(S.B.shape)), N=P, ldiv=True)


def test_block_bicgstab_float32_blocks(bpf, gen, oracle):
    S, A, B, Xd = _setup(bpf, gen, oracle, 118, 4, 44)
    P = bpf.ilu0(A, 4)
    X, st = bpf.block_bicgstab(A, bpf.HipMatrix.from_numpy(S.B.astype(np.float32)), N=P,
                               atol=0.0, rtol=1e-6, itmax=200)
    assert X.dtype == np.float32 and st.solved
    assert np.all(_rel(X.to_numpy().astype(np.float64), Xd) < 1e-3)


def test_ldiv_float32_block(bpf, gen, oracle):
    S, A, B, Xd = _setup(bpf, gen, oracle, 300, 5, 45)
    P = bpf.ilu0(A, 5)
    x32 = bpf.HipMatrix.from_numpy(S.B.astype(np.float32))
    y32 = bpf.HipMatrix((S.n, 5), np.float32)
    bpf.ldiv_(y32, P, x32)
    F = oracle.ilu0(S.indptr, S.indices, S.data)
    ref = oracle.ilu0_apply(S.indptr, S.indices, F, S.B.astype(np.float32).astype(np.float64))
    np.testing.assert_allclose(y32.to_numpy(), ref.astype(np.float32), rtol=1e-5,
                               atol=1e-5 * np.abs(ref).max())


# ------------------------------------- inner FP32 block GMRES preconditioner
@pytest.mark.parametrize("nbus,p,its,width", [(2000, 64, 8, 16), (9000, 128, 8, 16), (300, 16, 4, 0),
                                              (2000, 48, 6, 32)])
def test_block_gmres_inner_gmres_preconditioner(bpf, gen, oracle, nbus, p, its, width):
    # Flexible block GMRES whose preconditioner step is `its` iterations of
    # an FP32 ILU(0)-preconditioned block GMRES: the FP64 outer run still
    # reaches the backward-error target and the direct solution, in fewer
    # outer iterations than with ILU(0) alone.
    S, A, B, Xd = _setup(bpf, gen, oracle, nbus, p, 46)
    P = bpf.ilu0(A, p)
    kw = dict(N=P, ldiv=True, memory=40, precision="mixed", berr_tol=1e-10, itmax=1000)
    X, st = bpf.block_gmres(A, B, inner_its=its, inner_width=width, **kw)
    Xh = X.to_numpy()
    be = oracle.backward_error(S.indptr, S.indices, S.data, Xh, S.B)
    assert st.solved and be.max() <= 1e-10, (st, be.max())
    assert np.all(_rel(Xh, Xd) < 1e-5)
    _, st0 = bpf.block_gmres(A, B, **kw)
    assert st.niter < st0.niter, (st.niter, st0.niter)


def test_inner_gmres_overlapped_is_exact(bpf, gen, oracle):
    # The inner runs queue on the main stream only; the outer Hessenberg
    # update overlapping the next iteration changes nothing, bit for bit.
    S, A, B, Xd = _setup(bpf, gen, oracle, 2000, 32, 47)
    P = bpf.ilu0(A, 32)
    out = {}
    try:
        for ovl in (0, 1):
            bpf.set_option("overlap", ovl)
            X, st = bpf.block_gmres(A, B, N=P, ldiv=True, memory=40, precision="mixed", berr_tol=1e-10,
                                    itmax=400, inner_its=6, inner_width=16, history=True)
            out[ovl] = (X.to_numpy(), np.array(st.residuals), st.niter)
    finally:
        bpf.set_option("overlap", 1)
    assert out[0][2] == out[1][2]
    np.testing.assert_array_equal(out[1][1], out[0][1])
    np.testing.assert_array_equal(out[1][0], out[0][0])


def test_inner_gmres_graph_replay_is_exact(bpf, gen, oracle):
    # The inner FP32 runs are captured into a HIP graph on their first call
    # and replayed on every later outer iteration: the same kernels in the
    # same order, so the result is bitwise that of issuing every launch from
    # the host, and the replays did happen.
    S, A, B, Xd = _setup(bpf, gen, oracle, 2000, 32, 48)
    P = bpf.ilu0(A, 32)
    out = {}
    try:
        for graph in (0, 1):
            bpf.set_option("graph", graph)
            bpf.set_option("graph_replays", 0)
            X, st = bpf.block_gmres(A, B, N=P, ldiv=True, memory=40, precision="mixed", berr_tol=1e-10,
                                    itmax=400, inner_its=6, inner_width=16, history=True)
            out[graph] = (X.to_numpy(), np.array(st.residuals), st.niter, bpf.get_option("graph_replays"))
    finally:
        bpf.set_option("graph", 1)
    assert out[0][3] == 0
    # one eager call (which also captures), then one replay per outer iteration
    assert out[1][3] >= out[1][2] - 1 > 0, out[1][3]
    assert out[0][2] == out[1][2]
    np.testing.assert_array_equal(out[1][1], out[0][1])
    np.testing.assert_array_equal(out[1][0], out[0][0])
    be = oracle.backward_error(S.indptr, S.indices, S.data, out[1][0], S.B)
    assert be.max() <= 1e-10


def test_relabelled_solve(bpf, gen, oracle):
    # The mixed form on systems of at least `relabel_min` unknowns runs in a
    # reverse Cuthill-McKee relabelling (products and ILU(0) triangular solves
    # on relabelled copies of A and of the factor's operators, blocks permuted
    # at the layout conversion): the same solution, in the same number of
    # block iterations to within one (the Gram sums run over the rows in
    # another order); the relabelled copies follow new values of A and a new
    # factor.  The FP64 form keeps A's order: bitwise the same result.
    S, A, B, Xd = _setup(bpf, gen, oracle, 2000, 32, 49)
    P = bpf.ilu0(A, 32)
    kw = dict(N=P, ldiv=True, memory=40, berr_tol=1e-10, itmax=400)
    mixed = dict(precision="mixed", inner_its=6, inner_width=16)
    old = bpf.get_option("relabel_min")
    try:
        out = {}
        for rl in (0, 1000):
            bpf.set_option("relabel_min", rl)
            X, st = bpf.block_gmres(A, B, **kw, **mixed)
            out[rl] = (X.to_numpy(), st.niter, st.solved)
            Xp, _ = bpf.block_gmres(A, B, precision="fp64", **kw)
            out[rl, "fp64"] = Xp.to_numpy()
        assert out[1000][2] and abs(out[1000][1] - out[0][1]) <= 1, (out[0][1], out[1000][1])
        be = oracle.backward_error(S.indptr, S.indices, S.data, out[1000][0], S.B)
        assert be.max() <= 1e-10
        assert np.all(_rel(out[1000][0], Xd) < 1e-5)
        np.testing.assert_array_equal(out[1000, "fp64"], out[0, "fp64"])
        bpf.set_option("relabel_min", 1000)
        A.set_values(2.0 * S.data)
        P.factor_(A)
        X2, st2 = bpf.block_gmres(A, B, **kw, **mixed)
        assert st2.solved
        np.testing.assert_allclose(X2.to_numpy(), 0.5 * Xd, rtol=0, atol=1e-7 * np.abs(Xd).max())
    finally:
        bpf.set_option("relabel_min", old)
