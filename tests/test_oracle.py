"""CPU tests that pin the oracle (oracle/oracle.c, oracle/oracle.py).

Each check is chosen so that a plausible slip (dropped term, wrong sign,
transposed operand, off-by-one in an index) fails it: closed forms,
invariants, special cases that reduce to a library routine (SciPy / NumPy),
brute force on tiny inputs, and the reference's own known answers
(test/cusolverRF.jl) on its fixture when /root/reference is mounted.
"""
import os

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla

REF = "/root/reference"


def _rand_csr(n, density, seed, diag=5.0, sym_pattern=False):
    rng = np.random.default_rng(seed)
    M = sp.random(n, n, density=density, random_state=seed, format="csr")
    if sym_pattern:
        M = M + sp.random(n, n, density=0, format="csr") + M.T.multiply(0.5)
    M = M + sp.diags(diag + rng.random(n))
    M = sp.csr_matrix(M)
    M.sort_indices()
    return M


def _dense_ilu0(A, mask):
    """Brute-force ILU(0): Gaussian elimination that discards fill outside mask."""
    n = A.shape[0]
    a = A.copy()
    for i in range(1, n):
        for k in range(i):
            if not mask[i, k]:
                continue
            a[i, k] /= a[k, k]
            for j in range(k + 1, n):
                if mask[i, j]:
                    a[i, j] -= a[i, k] * a[k, j]
    return a


# ----------------------------------------------------------------- SpMM
@pytest.mark.parametrize("p", [1, 3, 17])
def test_spmm_matches_scipy(oracle, p):
    M = _rand_csr(60, 0.1, 1)
    X = np.random.default_rng(2).standard_normal((60, p))
    np.testing.assert_allclose(oracle.spmm(M.indptr, M.indices, M.data, X), M @ X, rtol=1e-14, atol=1e-14)


def test_spmm_row_sequential_bitwise(oracle):
    M = _rand_csr(40, 0.15, 3)
    X = np.random.default_rng(4).standard_normal((40, 2))
    ref = np.zeros((40, 2))
    for c in range(2):
        for i in range(40):
            acc = 0.0
            for k in range(M.indptr[i], M.indptr[i + 1]):
                acc += M.data[k] * X[M.indices[k], c]
            ref[i, c] = acc
    assert np.array_equal(oracle.spmm(M.indptr, M.indices, M.data, X), ref)


# ----------------------------------------------------------------- ILU(0)
def test_ilu0_matches_bruteforce(oracle):
    M = _rand_csr(30, 0.15, 5)
    F = oracle.ilu0(M.indptr, M.indices, M.data)
    mask = M.toarray() != 0
    ref = _dense_ilu0(M.toarray(), mask)
    got = sp.csr_matrix((F, M.indices, M.indptr), shape=M.shape).toarray()
    np.testing.assert_allclose(got[mask], ref[mask], rtol=1e-12, atol=1e-12)


def test_ilu0_invariant_on_pattern(oracle, gen):
    _, S = gen.make_system(300, 1, seed=3)
    F = oracle.ilu0(S.indptr, S.indices, S.data)
    Fm = sp.csr_matrix((F, S.indices, S.indptr), shape=(S.n, S.n))
    L = sp.tril(Fm, -1) + sp.eye(S.n)
    U = sp.triu(Fm)
    LU = (L @ U).tocsr()
    A = S.scipy()
    # (LU)_ij = a_ij for every (i, j) in pattern(A)  (Saad, Sec. 10.3.2)
    rows = np.repeat(np.arange(S.n), np.diff(S.indptr))
    lu_on = np.asarray(LU[rows, S.indices]).ravel()
    np.testing.assert_allclose(lu_on, S.data, rtol=1e-11, atol=1e-11 * np.abs(S.data).max())
    # ... but not an exact factorization (fill is discarded)
    assert abs(LU - A).max() > 1e-6


def test_ilu0_is_exact_lu_on_tridiagonal(oracle):
    n = 50
    rng = np.random.default_rng(7)
    T = sp.diags([rng.random(n - 1), 4 + rng.random(n), rng.random(n - 1)], [-1, 0, 1], format="csr")
    T.sort_indices()
    F = oracle.ilu0(T.indptr, T.indices, T.data)
    P, L, U = sla.lu(T.toarray())
    assert np.allclose(P, np.eye(n))   # diagonally dominant: no pivoting
    Fm = sp.csr_matrix((F, T.indices, T.indptr), shape=T.shape).toarray()
    np.testing.assert_allclose(np.tril(Fm, -1), np.tril(L, -1), atol=1e-13)
    np.testing.assert_allclose(np.triu(Fm), U, atol=1e-13)


def test_ilu0_zero_pivot_raises(oracle):
    A = sp.csr_matrix(np.array([[0.0, 1.0], [1.0, 1.0]]))
    A.sort_indices()
    with pytest.raises(ZeroDivisionError):
        oracle.ilu0(A.indptr, A.indices, A.data)


# ----------------------------------------------------------------- IC(0)
def test_ic0_is_cholesky_on_dense_spd(oracle):
    rng = np.random.default_rng(8)
    G = rng.standard_normal((12, 12))
    S = G @ G.T + 12 * np.eye(12)
    A = sp.csr_matrix(S)
    A.sort_indices()
    F = oracle.ic0(A.indptr, A.indices, A.data)
    Lg = np.tril(sp.csr_matrix((F, A.indices, A.indptr), shape=A.shape).toarray())
    np.testing.assert_allclose(Lg, np.linalg.cholesky(S), rtol=1e-12, atol=1e-12)


def test_ic0_invariant_on_lower_pattern(oracle):
    n = 200
    rng = np.random.default_rng(9)
    W = sp.random(n, n, density=0.03, random_state=9)
    W = W + W.T
    A = sp.csr_matrix(sp.diags(np.asarray(abs(W).sum(1)).ravel() + 1.0) - W)
    A.sort_indices()
    F = oracle.ic0(A.indptr, A.indices, A.data)
    L = sp.tril(sp.csr_matrix((F, A.indices, A.indptr), shape=A.shape)).tocsr()
    LLt = (L @ L.T).tocsr()
    lower = sp.tril(A).tocoo()
    np.testing.assert_allclose(np.asarray(LLt[lower.row, lower.col]).ravel(), lower.data, rtol=1e-11, atol=1e-11)


# -------------------------------------------------------- triangular solves
def test_trsm_modes_match_scipy(oracle):
    M = _rand_csr(80, 0.08, 10)
    X = np.random.default_rng(11).standard_normal((80, 4))
    F = M.data
    Lu = (sp.tril(M, -1) + sp.eye(80)).tocsr()
    U = sp.triu(M).tocsr()
    Ld = sp.tril(M).tocsr()
    np.testing.assert_allclose(oracle.trsm(0, M.indptr, M.indices, F, X),
                               spla.spsolve_triangular(Lu, X, lower=True), rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(oracle.trsm(1, M.indptr, M.indices, F, X),
                               spla.spsolve_triangular(U, X, lower=False), rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(oracle.trsm(2, M.indptr, M.indices, F, X),
                               spla.spsolve_triangular(Ld, X, lower=True), rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(oracle.trsm(3, M.indptr, M.indices, F, X),
                               spla.spsolve_triangular(Ld.T.tocsr(), X, lower=False), rtol=1e-11, atol=1e-12)


def test_ilu0_apply_order(oracle):
    # y = U^{-1} (L^{-1} x): L first (src/preconditioners_nvidia.jl:56-57)
    M = _rand_csr(50, 0.1, 12)
    F = oracle.ilu0(M.indptr, M.indices, M.data)
    Fm = sp.csr_matrix((F, M.indices, M.indptr), shape=M.shape).toarray()
    L = np.tril(Fm, -1) + np.eye(50)
    U = np.triu(Fm)
    x = np.random.default_rng(13).standard_normal((50, 2))
    np.testing.assert_allclose(oracle.ilu0_apply(M.indptr, M.indices, F, x),
                               np.linalg.solve(U, np.linalg.solve(L, x)), rtol=1e-10)


# ------------------------------------------------------- LU refactorization
def _symbolic_fill(B):
    """Pattern of L+U for row-wise elimination of B without pivoting (the
    closure the fixed-pattern refactorization needs), by brute force."""
    import heapq
    n = B.shape[0]
    rows = []
    for i in range(n):
        pat = set(B.indices[B.indptr[i]:B.indptr[i + 1]].tolist()) | {i}
        heap = [c for c in pat if c < i]
        heapq.heapify(heap)
        done = set()
        while heap:
            k = heapq.heappop(heap)
            if k in done:
                continue
            done.add(k)
            for j in rows[k]:
                if j > k and j not in pat:
                    pat.add(j)
                    if j < i:
                        heapq.heappush(heap, j)
        rows.append(sorted(pat))
    indptr = np.zeros(n + 1, dtype=np.int32)
    indptr[1:] = np.cumsum([len(r) for r in rows])
    indices = np.concatenate([np.array(r, dtype=np.int32) for r in rows])
    return sp.csr_matrix((np.ones(len(indices)), indices, indptr), shape=B.shape)


def _splu_pattern(A):
    """Row/column permutations of SuperLU with threshold 1 (partial pivoting)
    and the complete fill pattern of P A Q."""
    lu = spla.splu(sp.csc_matrix(A), permc_spec="COLAMD", diag_pivot_thresh=1.0)
    n = A.shape[0]
    Pr = np.empty(n, dtype=np.int32)
    Pr[lu.perm_r] = np.arange(n)          # row i of B = row Pr[i] of A
    Pc = np.empty(n, dtype=np.int32)
    Pc[lu.perm_c] = np.arange(n)          # column j of B = column Pc[j] of A
    B = sp.csr_matrix(A)[Pr][:, Pc].tocsr()
    B.sort_indices()
    F = _symbolic_fill(B)
    return Pr, Pc, F, lu


def test_lu_refactor_reproduces_factorization(oracle, gen):
    _, S = gen.make_system(118, 1, seed=4)
    A = S.scipy()
    Pr, Pc, F, lu = _splu_pattern(A)
    vals = oracle.lu_refactor(S.indptr, S.indices, S.data, Pr, Pc, F.indptr, F.indices)
    # same pivots as SuperLU: its U diagonal agrees with ours
    Fm = sp.csr_matrix((vals, F.indices, F.indptr), shape=A.shape)
    np.testing.assert_allclose(Fm.diagonal(), lu.U.diagonal(), rtol=1e-9)
    L = sp.tril(Fm, -1) + sp.eye(S.n)
    U = sp.triu(Fm)
    B = A[Pr][:, Pc]
    assert abs(L @ U - B).max() <= 1e-10 * abs(B).max()


def test_lu_refactor_known_answers(oracle, gen):
    # Replays test/cusolverRF.jl:21-28 on a synthetic Jacobian: the refactored
    # solve equals the direct solve; values x2 give half the solution.
    _, S = gen.make_system(300, 1, seed=5)
    A = S.scipy()
    Pr, Pc, F, _ = _splu_pattern(A)
    b = np.random.default_rng(5).random(S.n)
    x_direct = oracle.direct_solve(S.indptr, S.indices, S.data, b)

    def solve(vals):
        f = oracle.lu_refactor(S.indptr, S.indices, vals, Pr, Pc, F.indptr, F.indices)
        y = oracle.trsm(0, F.indptr, F.indices, f, b[Pr])
        z = oracle.trsm(1, F.indptr, F.indices, f, y)
        x = np.empty(S.n)
        x[Pc] = z[:, 0]
        return x

    np.testing.assert_allclose(solve(S.data), x_direct, rtol=1.5e-8)
    np.testing.assert_allclose(solve(2.0 * S.data), 0.5 * x_direct, rtol=1.5e-8)


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "data", "case300.txt")),
                    reason="reference fixture not mounted")
def test_reference_fixture_case300_known_answers(oracle):
    import scipy.io
    J = sp.csr_matrix(scipy.io.mmread(os.path.join(REF, "data", "case300.txt")))
    J.sort_indices()
    assert J.shape == (530, 530) and J.nnz == 3736
    Pr, Pc, F, _ = _splu_pattern(J)
    rng = np.random.default_rng(0)
    b = rng.random(530)
    sol = spla.spsolve(J.tocsc(), b)                  # J \ b (test/cusolverRF.jl:12)
    f = oracle.lu_refactor(J.indptr, J.indices, J.data, Pr, Pc, F.indptr, F.indices)
    z = oracle.trsm(1, F.indptr, F.indices, f, oracle.trsm(0, F.indptr, F.indices, f, b[Pr]))
    x = np.empty(530)
    x[Pc] = z[:, 0]
    np.testing.assert_allclose(x, sol, rtol=1.5e-8)   # isapprox (:21)
    f2 = oracle.lu_refactor(J.indptr, J.indices, 2 * J.data, Pr, Pc, F.indptr, F.indices)
    z2 = oracle.trsm(1, F.indptr, F.indices, f2, oracle.trsm(0, F.indptr, F.indices, f2, b[Pr]))
    x2 = np.empty(530)
    x2[Pc] = z2[:, 0]
    np.testing.assert_allclose(x2, 0.5 * sol, rtol=1.5e-8)   # (:24-28)
    # batch: i*b -> i*x (:42-68)
    for i in (1, 7, 32):
        np.testing.assert_allclose(oracle.direct_solve(J.indptr, J.indices, J.data, i * b), i * sol,
                                   rtol=1.5e-8)


# ---------------------------------------------------------- Krylov solvers
def _textbook_gmres(A, b, N, m):
    """Right-preconditioned GMRES(m) without restart, Givens rotations
    (Saad, Iterative Methods, Alg. 6.9 / 9.5); returns residual norms."""
    n = len(b)
    beta = np.linalg.norm(b)
    V = np.zeros((n, m + 1))
    H = np.zeros((m + 1, m))
    V[:, 0] = b / beta
    g = np.zeros(m + 1)
    g[0] = beta
    cs, sn = np.zeros(m), np.zeros(m)
    res = [beta]
    for j in range(m):
        w = A @ N(V[:, j:j + 1])[:, 0]
        for i in range(j + 1):
            H[i, j] = V[:, i] @ w
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


def test_block_gmres_p1_is_textbook_gmres(oracle, gen):
    _, S = gen.make_system(118, 1, seed=6)
    A = S.scipy()
    F = oracle.ilu0(S.indptr, S.indices, S.data)
    N = lambda Z: oracle.ilu0_apply(S.indptr, S.indices, F, Z)
    X, st = oracle.block_gmres(A, S.B, memory=30, N=N, atol=0.0, rtol=1e-12, itmax=12, history=True)
    ref = _textbook_gmres(A, S.B[:, 0], N, len(st.residuals) - 1)
    np.testing.assert_allclose(st.residuals, ref, rtol=1e-8)


def test_block_gmres_finite_termination(oracle, gen):
    # n = 14, p = 2, no preconditioner, no restart: exact in ceil(n/p) = 7 steps
    _, S = gen.make_system(9, 2, seed=7)
    assert S.n == 14
    X, st = oracle.block_gmres(S.scipy(), S.B, memory=20, atol=0.0, rtol=1e-13, itmax=7)
    Xd = np.linalg.solve(S.scipy().toarray(), S.B)
    np.testing.assert_allclose(X, Xd, rtol=1e-8, atol=1e-10)
    assert st.niter <= 7


def test_block_gmres_estimate_equals_true_residual(oracle, gen):
    _, S = gen.make_system(118, 4, seed=8)
    A = S.scipy()
    for k in (1, 3, 6):
        X, st = oracle.block_gmres(A, S.B, memory=30, atol=0.0, rtol=1e-30, itmax=k, history=True)
        true = np.linalg.norm(S.B - A @ X)
        assert abs(st.residuals[-1] - true) <= 1e-9 * np.linalg.norm(S.B)
        # minimal residual: non-increasing within a cycle
        assert np.all(np.diff(st.residuals) <= 1e-12 * st.residuals[0])


def test_block_gmres_converges_to_direct_solution(oracle, gen):
    _, S = gen.make_system(300, 16, seed=9)
    A = S.scipy()
    F = oracle.ilu0(S.indptr, S.indices, S.data)
    N = lambda Z: oracle.ilu0_apply(S.indptr, S.indices, F, Z)
    X, st = oracle.block_gmres(A, S.B, memory=30, N=N, atol=0.0, rtol=1e-11, restart=False)
    assert st.solved and st.status == "solution good enough given atol and rtol"
    Xd = oracle.direct_solve(S.indptr, S.indices, S.data, S.B)
    np.testing.assert_allclose(X, Xd, rtol=1e-7, atol=1e-8 * np.abs(Xd).max())
    assert oracle.backward_error(S.indptr, S.indices, S.data, X, S.B).max() < 1e-12


def test_block_gmres_tired_status(oracle, gen):
    _, S = gen.make_system(118, 2, seed=10)
    X, st = oracle.block_gmres(S.scipy(), S.B, atol=0.0, rtol=1e-14, itmax=2)
    assert not st.solved and st.niter == 2
    assert st.status == "maximum number of iterations exceeded"


def _scalar_bicgstab(A, b, N, iters):
    """Preconditioned BiCGSTAB with shadow vector b and |omega| as in
    src/block_bicgstab.jl:146 (the block recurrence with p = 1)."""
    x = np.zeros_like(b)
    r = b.copy()
    p = b.copy()
    res = [np.linalg.norm(r)]
    for _ in range(iters):
        y = N(p)
        v = A @ y
        alpha = (b @ r) / (b @ v)
        u = r - alpha * v
        x = x + alpha * y
        y = N(u)
        q = A @ y
        omega = abs(q @ u) / abs(q @ q)
        x = x + omega * y
        r = u - omega * q
        beta = (b @ q) / (b @ v)
        p = r - beta * (p - omega * v)
        res.append(np.linalg.norm(r))
    return x, np.array(res)


def test_block_bicgstab_p1_is_scalar_recurrence(oracle, gen):
    _, S = gen.make_system(118, 1, seed=11)
    A = S.scipy()
    F = oracle.ilu0(S.indptr, S.indices, S.data)
    N = lambda Z: oracle.ilu0_apply(S.indptr, S.indices, F, Z)
    X, st = oracle.block_bicgstab(A, S.B, N=N, atol=0.0, rtol=1e-30, itmax=8, history=True)
    x, res = _scalar_bicgstab(A, S.B[:, 0], lambda v: N(v[:, None])[:, 0], 8)
    np.testing.assert_allclose(st.residuals, res, rtol=1e-8)
    np.testing.assert_allclose(X[:, 0], x, rtol=1e-8, atol=1e-12)


def test_block_bicgstab_uses_abs_omega(oracle):
    # A matrix where <Q,U> < 0 at the first step: |omega| differs from omega.
    A = np.array([[1.0, 2.0], [-3.0, 1.0]])
    b = np.array([[1.0], [0.0]])
    X, st = oracle.block_bicgstab(A, b, atol=0.0, rtol=1e-30, itmax=1, history=True)
    x, res = _scalar_bicgstab(A, b[:, 0], lambda v: v, 1)
    np.testing.assert_allclose(X[:, 0], x, rtol=1e-12)


def _textbook_bicgstab(A, b, N, iters, shadow=None):
    """Right-preconditioned BiCGSTAB in van der Vorst's rho-ratio form
    (van der Vorst 1992; Saad, Iterative Methods, Alg. 7.7): beta from
    rho_{i+1}/rho_i * alpha/omega, omega = (t, s)/(t, t) without absolute
    values.  Returns x, the residual norms and the omegas."""
    rhat = b.copy() if shadow is None else shadow
    x = np.zeros_like(b)
    r = b.copy()
    p = r.copy()
    rho = rhat @ r
    res, omegas = [np.linalg.norm(r)], []
    for _ in range(iters):
        y = N(p)
        v = A @ y
        alpha = rho / (rhat @ v)
        s = r - alpha * v
        z = N(s)
        t = A @ z
        omega = (t @ s) / (t @ t)
        x = x + alpha * y + omega * z
        r = s - omega * t
        rho_new = rhat @ r
        beta = (rho_new / rho) * (alpha / omega)
        p = r + beta * (p - omega * v)
        rho = rho_new
        res.append(np.linalg.norm(r))
        omegas.append(omega)
    return x, np.array(res), np.array(omegas)


def _pd_symmetric_part(n, seed):
    # A + A^T positive definite: (A s, s) > 0 for every s, so omega > 0 and
    # the reference's |omega| (src/block_bicgstab.jl:146) equals omega.
    rng = np.random.default_rng(seed)
    M = sp.random(n, n, density=0.05, random_state=seed, format="csr")
    M = M - M.T + sp.diags(4.0 + rng.random(n))
    return sp.csr_matrix(M)


def test_block_bicgstab_p1_is_textbook_bicgstab(oracle):
    A = _pd_symmetric_part(120, 21)
    b = np.random.default_rng(22).standard_normal((120, 1))
    X, st = oracle.block_bicgstab(A, b, atol=0.0, rtol=1e-30, itmax=10, history=True)
    x, res, om = _textbook_bicgstab(A, b[:, 0], lambda v: v, 10)
    assert np.all(om > 0)
    np.testing.assert_allclose(st.residuals, res, rtol=1e-8)
    np.testing.assert_allclose(X[:, 0], x, rtol=1e-8, atol=1e-12 * np.abs(x).max())


def test_block_bicgstab_p1_preconditioned_is_textbook(oracle, gen):
    _, S = gen.make_system(118, 1, seed=23)
    A = S.scipy()
    F = oracle.ilu0(S.indptr, S.indices, S.data)
    N = lambda Z: oracle.ilu0_apply(S.indptr, S.indices, F, Z)
    X, st = oracle.block_bicgstab(A, S.B, N=N, atol=0.0, rtol=1e-30, itmax=6, history=True)
    x, res, om = _textbook_bicgstab(A, S.B[:, 0], lambda v: N(v[:, None])[:, 0], 6)
    assert np.all(om > 0)      # on these 6 steps |omega| = omega
    np.testing.assert_allclose(st.residuals, res, rtol=1e-7)
    np.testing.assert_allclose(X[:, 0], x, rtol=1e-7, atol=1e-10 * np.abs(x).max())


@pytest.mark.parametrize("p", [1, 8])
def test_block_bicgstab_converges_to_direct_solution(oracle, p):
    M = _rand_csr(300, 0.02, 24, diag=8.0)
    B = np.random.default_rng(25).standard_normal((300, p))
    X, st = oracle.block_bicgstab(M, B, atol=0.0, rtol=1e-12, itmax=300)
    assert st.solved and st.status == "solution good enough given atol and rtol"
    Xd = spla.splu(sp.csc_matrix(M)).solve(B)
    np.testing.assert_allclose(X, Xd, rtol=1e-9, atol=1e-10 * np.abs(Xd).max())


def test_block_bicgstab_recursive_residual_is_true_residual(oracle):
    M = _rand_csr(200, 0.03, 26, diag=6.0)
    B = np.random.default_rng(27).standard_normal((200, 4))
    F = oracle.ilu0(M.indptr, M.indices, M.data)
    N = lambda Z: oracle.ilu0_apply(M.indptr, M.indices, F, Z)
    gaps = []

    def cb(it, X, R):
        gaps.append(np.linalg.norm(R - (B - M @ X)) / np.linalg.norm(B))

    X, st = oracle.block_bicgstab(M, B, N=N, atol=0.0, rtol=1e-12, itmax=60, callback=cb)
    assert st.solved and len(gaps) == st.niter
    assert max(gaps) < 1e-12


def test_block_bicgstab_shadow_block_C(oracle):
    # C = B is the default shadow; another C gives another (valid) iteration
    M = _rand_csr(150, 0.03, 28, diag=6.0)
    B = np.random.default_rng(29).standard_normal((150, 3))
    C = np.random.default_rng(30).standard_normal((150, 3))
    X1, s1 = oracle.block_bicgstab(M, B, atol=0.0, rtol=1e-12, itmax=200)
    X2, s2 = oracle.block_bicgstab(M, B, C=C, atol=0.0, rtol=1e-12, itmax=200, history=True)
    X3, s3 = oracle.block_bicgstab(M, B, C=B, atol=0.0, rtol=1e-12, itmax=200)
    assert s1.solved and s2.solved
    np.testing.assert_array_equal(X1, X3)
    np.testing.assert_allclose(X1, X2, rtol=1e-8, atol=1e-10 * np.abs(X1).max())
    x, res, _ = _textbook_bicgstab(M, B[:, 0], lambda v: v, 3, shadow=C[:, 0])
    Xs, ss = oracle.block_bicgstab(M, B[:, :1], C=C[:, :1], atol=0.0, rtol=1e-30, itmax=3, history=True)
    np.testing.assert_allclose(ss.residuals, res, rtol=1e-8)


# --------------------------------------------------- left preconditioning
def _textbook_left_gmres(A, b, Minv, m):
    """Left-preconditioned GMRES without restart (Saad, Alg. 9.4): Arnoldi on
    M^{-1} A from r0 = M^{-1} b; returns ||M^{-1}(b - A x_k)|| per step."""
    r0 = Minv(b)
    beta = np.linalg.norm(r0)
    n = len(b)
    V = np.zeros((n, m + 1))
    H = np.zeros((m + 1, m))
    V[:, 0] = r0 / beta
    res = [beta]
    for j in range(m):
        w = Minv(A @ V[:, j])
        for i in range(j + 1):
            H[i, j] = V[:, i] @ w
            w = w - H[i, j] * V[:, i]
        H[j + 1, j] = np.linalg.norm(w)
        V[:, j + 1] = w / H[j + 1, j]
        e = np.zeros(j + 2)
        e[0] = beta
        y = np.linalg.lstsq(H[:j + 2, :j + 1], e, rcond=None)[0]
        res.append(np.linalg.norm(e - H[:j + 2, :j + 1] @ y))
    return np.array(res)


def test_block_gmres_p1_left_is_textbook(oracle, gen):
    _, S = gen.make_system(118, 1, seed=31)
    A = S.scipy()
    F = oracle.ilu0(S.indptr, S.indices, S.data)
    Mv = lambda Z: oracle.ilu0_apply(S.indptr, S.indices, F, Z)
    X, st = oracle.block_gmres(A, S.B, memory=30, M=Mv, atol=0.0, rtol=1e-12, itmax=10, history=True)
    ref = _textbook_left_gmres(A, S.B[:, 0], lambda v: Mv(v[:, None])[:, 0], len(st.residuals) - 1)
    np.testing.assert_allclose(st.residuals, ref, rtol=1e-8)
    # the estimate is the M-weighted true residual (not ||b - A x||)
    true_m = np.linalg.norm(Mv(S.B - A @ X))
    assert abs(st.residuals[-1] - true_m) <= 1e-9 * np.linalg.norm(Mv(S.B))


def test_block_gmres_left_and_right_converge(oracle, gen):
    _, S = gen.make_system(300, 6, seed=32)
    A = S.scipy()
    F = oracle.ilu0(S.indptr, S.indices, S.data)
    P = lambda Z: oracle.ilu0_apply(S.indptr, S.indices, F, Z)
    Xd = oracle.direct_solve(S.indptr, S.indices, S.data, S.B)
    for kw in (dict(M=P), dict(M=P, N=P)):
        X, st = oracle.block_gmres(A, S.B, memory=60, atol=0.0, rtol=1e-12, **kw)
        assert st.solved
        np.testing.assert_allclose(X, Xd, rtol=1e-7, atol=1e-8 * np.abs(Xd).max())


def test_ic0_apply_order(oracle):
    # y = L^{-T} (L^{-1} x): L first, then its transpose
    # (src/preconditioners_nvidia.jl:19-22)
    n = 60
    W = sp.random(n, n, density=0.08, random_state=33)
    W = W + W.T
    A = sp.csr_matrix(sp.diags(np.asarray(abs(W).sum(1)).ravel() + 1.0) - W)
    A.sort_indices()
    F = oracle.ic0(A.indptr, A.indices, A.data)
    L = np.tril(sp.csr_matrix((F, A.indices, A.indptr), shape=A.shape).toarray())
    x = np.random.default_rng(34).standard_normal((n, 3))
    ref = np.linalg.solve(L.T, np.linalg.solve(L, x))
    np.testing.assert_allclose(oracle.ic0_apply(A.indptr, A.indices, F, x), ref, rtol=1e-10, atol=1e-12)
    wrong = np.linalg.solve(L, np.linalg.solve(L.T, x))
    assert np.abs(wrong - ref).max() > 1e-6


# ------------------------------------------------------------ block Jacobi
def test_block_jacobi_one_part_is_exact_inverse(oracle):
    M = _rand_csr(40, 0.1, 35)
    X = np.random.default_rng(36).standard_normal((40, 3))
    apply = oracle.block_jacobi(M.indptr, M.indices, M.data, np.zeros(40, int))
    np.testing.assert_allclose(apply(X), np.linalg.solve(M.toarray(), X), rtol=1e-11, atol=1e-12)


def test_block_jacobi_exact_on_block_diagonal(oracle):
    rng = np.random.default_rng(37)
    sizes = [5, 3, 7, 1]
    D = sp.block_diag([sp.csr_matrix(rng.standard_normal((k, k)) + 4 * np.eye(k)) for k in sizes]).tocsr()
    perm = rng.permutation(D.shape[0])        # parts need not be contiguous
    Pm = sp.eye(D.shape[0]).tocsr()[perm]
    A = (Pm @ D @ Pm.T).tocsr()
    A.sort_indices()
    part_d = np.repeat(np.arange(len(sizes)), sizes)
    part = np.empty_like(part_d)
    part[np.argsort(perm)] = part_d
    X = rng.standard_normal((A.shape[0], 2))
    apply = oracle.block_jacobi(A.indptr, A.indices, A.data, part)
    np.testing.assert_allclose(apply(X), np.linalg.solve(A.toarray(), X), rtol=1e-11, atol=1e-12)
    # dropping the coupling: on a matrix with off-block entries only the
    # diagonal blocks count
    E = A + sp.csr_matrix((np.ones(1), ([0], [(int(np.flatnonzero(part != part[0])[0]))])), shape=A.shape)
    apply2 = oracle.block_jacobi(E.indptr, E.indices, E.data, part)
    np.testing.assert_allclose(apply2(X), apply(X), rtol=1e-13)


# ---------------------------------------------------- direct solve, berr
def test_direct_solve_and_backward_error(oracle):
    M = _rand_csr(40, 0.1, 14)
    B = np.random.default_rng(15).standard_normal((40, 3))
    X = oracle.direct_solve(M.indptr, M.indices, M.data, B)
    np.testing.assert_allclose(X, np.linalg.solve(M.toarray(), B), rtol=1e-10)
    be = oracle.backward_error(M.indptr, M.indices, M.data, X, B)
    assert be.max() < 1e-15
    Xp = X.copy()
    Xp[0, 1] += 1e-6
    be2 = oracle.backward_error(M.indptr, M.indices, M.data, Xp, B)
    R = B[:, 1] - M @ Xp[:, 1]
    ref = np.abs(R).max() / (abs(M).sum(1).max() * np.abs(Xp[:, 1]).max() + np.abs(B[:, 1]).max())
    assert abs(be2[1] - ref) <= 1e-12 * ref


# ---------------------------------------------------------- generator
def test_generator_sizes_and_determinism(gen):
    _, S1 = gen.make_system(300, 16, seed=1)
    _, S2 = gen.make_system(300, 16, seed=1)
    assert S1.n == S2.n and np.array_equal(S1.data, S2.data) and np.array_equal(S1.B, S2.B)
    assert 500 <= S1.n <= 560 and 3000 <= S1.nnz <= 5000
    # structurally symmetric with a full diagonal (like the reference's dumps)
    A = S1.scipy()
    assert (abs(A) != 0).toarray().tolist() == (abs(A.T) != 0).toarray().tolist()
    assert np.all(A.diagonal() != 0)
    g, _ = gen.make_system(2000, 1, seed=2)
    assert len(g.pv) == 400 and len(g.pq) == 1599


def test_generator_injection_rhs(gen):
    grid, S = gen.make_system(300, 10, seed=3, rhs="injection", nunit=6)
    U = S.B[:, 4:]
    assert np.all((U != 0).sum(axis=0) == 1) and np.all(U.max(axis=0) == 1.0)
