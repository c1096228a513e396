"""Dense building blocks of block GMRES on the GPU, through the C ABI:
the thin QR of a block (householder!, src/block_utils.jl:38-53) and the
block-Hessenberg least-squares update + back substitution
(src/block_gmres.jl:318-375), against NumPy/LAPACK in FP64."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,p", [(3000, 1), (3000, 7), (5000, 16), (4000, 33), (6000, 64), (8000, 128)])
def test_householder_cholqr2(bpf, n, p):
    rng = np.random.default_rng(n + p)
    A = rng.standard_normal((n, p)) * np.logspace(0, 3, p)     # columns of uneven scale
    Qd, Rd = bpf.householder(bpf.HipMatrix.from_numpy(A))
    Q, R = Qd.to_numpy(), Rd.to_numpy()
    assert not Qd.shifted
    assert np.abs(Q.T @ Q - np.eye(p)).max() <= 1e-13 * max(1, p)
    assert np.linalg.norm(Q @ R - A) <= 1e-14 * p * np.linalg.norm(A)
    assert np.all(np.tril(R, -1) == 0.0) and np.all(np.diag(R) > 0)
    # R equals LAPACK's R up to row signs.
    Rl = np.linalg.qr(A, mode="r")
    np.testing.assert_allclose(np.abs(R), np.abs(Rl), rtol=1e-9, atol=1e-12 * np.abs(Rl).max())


def test_householder_ill_conditioned_and_rank_deficient(bpf):
    rng = np.random.default_rng(5)
    n, p = 4000, 32
    U, _ = np.linalg.qr(rng.standard_normal((n, p)))
    V, _ = np.linalg.qr(rng.standard_normal((p, p)))
    A = U @ np.diag(np.logspace(0, -7, p)) @ V.T               # cond 1e7: two passes suffice
    Qd, Rd = bpf.householder(bpf.HipMatrix.from_numpy(A))
    Q, R = Qd.to_numpy(), Rd.to_numpy()
    assert np.abs(Q.T @ Q - np.eye(p)).max() <= 1e-12
    assert np.linalg.norm(Q @ R - A) <= 1e-13 * np.linalg.norm(A)
    B = A.copy()
    B[:, 5] = B[:, 3]                                           # exactly rank deficient
    Qd, Rd = bpf.householder(bpf.HipMatrix.from_numpy(B))
    assert Qd.shifted                                           # reported as a breakdown
    Q, R = Qd.to_numpy(), Rd.to_numpy()
    assert np.all(np.isfinite(Q)) and np.all(np.isfinite(R))
    # the dependent column is deflated or re-orthogonalized: Q stays
    # orthonormal and Q R reproduces B to rounding
    assert np.abs(Q.T @ Q - np.eye(p)).max() <= 1e-12
    assert np.linalg.norm(Q @ R - B) <= 1e-12 * np.linalg.norm(B)
    assert abs(R[5, 5]) <= 1e-8 * np.linalg.norm(B[:, 5])
    # exactly dependent columns (pivots at rounding level): Householder QR
    # completes Q with orthonormal directions and R has zero diagonal entries
    rng = np.random.default_rng(6)
    C = rng.standard_normal((n, p))
    C[:, 7] = 0.0                                               # a zero column
    C[:, 9] = 2.0 * C[:, 2] - C[:, 4]                           # a combination
    Qd, Rd = bpf.householder(bpf.HipMatrix.from_numpy(C))
    Q, R = Qd.to_numpy(), Rd.to_numpy()
    assert Qd.shifted
    assert np.linalg.norm(Q @ R - C) <= 1e-12 * np.linalg.norm(C)
    assert R[7, 7] == 0.0
    assert np.all(np.isfinite(Q))
    # the combination column: a pivot at rounding level of either sign is
    # deflated (R[9, 9] = 0) or kept and re-projected; either way Q is
    # orthonormal and R[9, 9] is at rounding level
    assert np.abs(Q.T @ Q - np.eye(p)).max() <= 1e-12
    assert abs(R[9, 9]) <= 1e-8 * np.linalg.norm(C[:, 9])


def _hessenberg(rng, p, K, triangular_sub):
    H = np.zeros(((K + 1) * p, K * p))
    for k in range(K):
        H[: (k + 1) * p, k * p:(k + 1) * p] = rng.standard_normal(((k + 1) * p, p))
        C = rng.standard_normal((p, p)) + 3 * np.eye(p)
        H[(k + 1) * p:(k + 2) * p, k * p:(k + 1) * p] = np.triu(C) if triangular_sub else C
    G = np.triu(rng.standard_normal((p, p))) + 2 * np.eye(p)
    return H, G


@pytest.mark.parametrize("p,K,tri", [(1, 6, True), (5, 4, True), (16, 8, True), (20, 3, False),
                                     (32, 5, True), (64, 4, False), (128, 3, True)])
def test_block_hessenberg_lsq_matches_lstsq(bpf, p, K, tri):
    rng = np.random.default_rng(100 * p + K)
    H, G = _hessenberg(rng, p, K, tri)
    Y, rn = bpf.block_hessenberg_lsq(H, G)
    rhs = np.zeros(((K + 1) * p, p))
    rhs[:p] = G
    Yl = np.linalg.lstsq(H, rhs, rcond=None)[0]
    np.testing.assert_allclose(Y, Yl, rtol=0, atol=1e-10 * np.abs(Yl).max())
    # residual estimate after each block column = the true least-squares residual
    for k in range(K):
        Hk = H[: (k + 2) * p, : (k + 1) * p]
        r = rhs[: (k + 2) * p]
        Yk = np.linalg.lstsq(Hk, r, rcond=None)[0]
        true = np.linalg.norm(r - Hk @ Yk, axis=0)
        np.testing.assert_allclose(rn[k], true, rtol=1e-9, atol=1e-12 * np.linalg.norm(G))
    # minimal-residual invariant: the estimates never increase
    assert np.all(np.diff(rn, axis=0) <= 1e-12 * np.linalg.norm(G))
