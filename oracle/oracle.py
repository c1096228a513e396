"""CPU oracle for the multi-RHS solve path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py import this module, and only as the checker / CPU baseline.  The
product (the HIP library and its Python mirror) never imports it.

Contents
--------
* ctypes bindings of ``liboracle.so`` (oracle.c): SpMM, ILU(0), IC(0),
  triangular solves on a combined factor, fixed-pattern LU refactorization;
* ``block_gmres``: restatement of src/block_gmres.jl:170-411 on CPU arrays
  (Householder QR through LAPACK geqrf/orgqr/ormqr, exactly the calls the
  reference's CPU path makes through Julia's LAPACK wrappers);
* ``block_bicgstab``: restatement of the CPU variant src/block_bicgstab.jl:70-178
  (p x p solves by LU, omega with absolute values as at :146);
* ``direct_solve``: FP64 sparse direct solve (SuperLU via scipy), the known
  answer the reference's own test uses (UMFPACK ``J \\ b``,
  test/cusolverRF.jl:12);
* ``backward_error``: per-RHS ||A x - b||_inf / (||A||_inf ||x||_inf + ||b||_inf);
* ``block_jacobi``: dense per-block inverses for a given partition (ExaPF's
  BlockJacobiPreconditioner, benchmarks/benchmarks_nvidia.jl:46-51).

Pinning (tests/test_oracle.py): block_gmres reduces to textbook GMRES for p=1,
right-preconditioned (Saad Alg. 6.9 / 9.5 with Givens) and left-preconditioned
(Alg. 9.4), terminates exactly in ceil(n/p) iterations on tiny systems, and its
residual estimate equals the true (M-weighted) residual norm; block_bicgstab
with p=1 matches van der Vorst's rho-ratio BiCGSTAB where omega > 0, its
recursive residual equals B - A X at every step, and both solvers converge to
the direct solution; block_jacobi equals the exact inverse for one part and
on block-diagonal matrices.
"""
from __future__ import annotations

import ctypes
import os
import time

import numpy as np
import scipy.linalg
import scipy.linalg.lapack as lapack
import scipy.sparse as sp
import scipy.sparse.linalg as spla

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = ctypes.CDLL(path)
        ip = ctypes.POINTER(ctypes.c_int)
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_spmm.argtypes = [ctypes.c_int, ip, ip, dp, ctypes.c_int, dp, ctypes.c_int, dp, ctypes.c_int]
        L.orc_ilu0.argtypes = [ctypes.c_int, ip, ip, dp, dp]
        L.orc_ic0.argtypes = [ctypes.c_int, ip, ip, dp, dp]
        L.orc_trsm.argtypes = [ctypes.c_int, ctypes.c_int, ip, ip, dp, ctypes.c_int, dp, ctypes.c_int, dp, ctypes.c_int]
        L.orc_lu_refactor.argtypes = [ctypes.c_int, ip, ip, dp, ip, ip, ip, ip, dp]
        for f in (L.orc_spmm, L.orc_ilu0, L.orc_ic0, L.orc_trsm, L.orc_lu_refactor):
            f.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _csr(indptr, indices, data):
    return (np.ascontiguousarray(indptr, dtype=np.int32),
            np.ascontiguousarray(indices, dtype=np.int32),
            np.ascontiguousarray(data, dtype=np.float64))


def spmm(indptr, indices, data, X):
    """Y = A X (oracle.c orc_spmm; reference mul!(W,A,P), src/block_gmres.jl:299)."""
    ip_, ix, dv = _csr(indptr, indices, data)
    n = len(ip_) - 1
    X = np.asfortranarray(X, dtype=np.float64)
    if X.ndim == 1:
        X = X.reshape(-1, 1, order="F")
    Y = np.zeros((n, X.shape[1]), order="F")
    lib().orc_spmm(n, _ip(ip_), _ip(ix), _dp(dv), X.shape[1], _dp(X), max(1, X.shape[0]), _dp(Y), max(1, n))
    return Y


def ilu0(indptr, indices, data):
    """Combined ILU(0) factor on A's pattern (oracle.c orc_ilu0)."""
    ip_, ix, dv = _csr(indptr, indices, data)
    out = np.empty_like(dv)
    rc = lib().orc_ilu0(len(ip_) - 1, _ip(ip_), _ip(ix), _dp(dv), _dp(out))
    if rc != 0:
        raise ZeroDivisionError(f"ILU0 zero pivot at row {-rc - 1}")
    return out


def ic0(indptr, indices, data):
    ip_, ix, dv = _csr(indptr, indices, data)
    out = np.empty_like(dv)
    rc = lib().orc_ic0(len(ip_) - 1, _ip(ip_), _ip(ix), _dp(dv), _dp(out))
    if rc != 0:
        raise ValueError(f"IC0 non-positive pivot at row {-rc - 1}")
    return out


TRSM_UNIT_LOWER, TRSM_UPPER, TRSM_LOWER, TRSM_LOWER_T = 0, 1, 2, 3


def trsm(mode, indptr, indices, fval, X):
    ip_, ix, dv = _csr(indptr, indices, fval)
    n = len(ip_) - 1
    X = np.asfortranarray(X, dtype=np.float64)
    if X.ndim == 1:
        X = X.reshape(-1, 1, order="F")
    Y = np.zeros_like(X, order="F")
    rc = lib().orc_trsm(mode, n, _ip(ip_), _ip(ix), _dp(dv), X.shape[1], _dp(X), max(1, n), _dp(Y), max(1, n))
    if rc != 0:
        raise ZeroDivisionError(f"zero diagonal at row {-rc - 1}")
    return Y


def ilu0_apply(indptr, indices, fval, X):
    """y = U^{-1} L^{-1} x (src/preconditioners_nvidia.jl:55-58)."""
    return trsm(TRSM_UPPER, indptr, indices, fval, trsm(TRSM_UNIT_LOWER, indptr, indices, fval, X))


def ic0_apply(indptr, indices, fval, X):
    """y = L^{-T} L^{-1} x (src/preconditioners_nvidia.jl:19-22)."""
    return trsm(TRSM_LOWER_T, indptr, indices, fval, trsm(TRSM_LOWER, indptr, indices, fval, X))


def block_jacobi(indptr, indices, data, part):
    """Block-Jacobi preconditioner for a given partition (ExaPF's
    BlockJacobiPreconditioner as passed by benchmarks/benchmarks_nvidia.jl:46-51):
    returns a callable Y = apply(X) with Y[idx] = inv(A[idx, idx]) X[idx] for
    the rows idx of each part (dense inverses by LAPACK)."""
    ip, ix, dv = _csr(indptr, indices, data)
    A = sp.csr_matrix((dv, ix, ip), shape=(len(ip) - 1, len(ip) - 1))
    part = np.asarray(part)
    blocks = []
    for b in range(int(part.max()) + 1):
        idx = np.flatnonzero(part == b)
        blocks.append((idx, np.linalg.inv(A[idx][:, idx].toarray())))

    def apply(X):
        X = np.asarray(X, dtype=np.float64)
        Y = np.empty_like(X)
        for idx, Binv in blocks:
            Y[idx] = Binv @ X[idx]
        return Y
    return apply


def lu_refactor(indptr, indices, data, P, Q, frow, fcol):
    ip_, ix, dv = _csr(indptr, indices, data)
    P = np.ascontiguousarray(P, dtype=np.int32)
    Q = np.ascontiguousarray(Q, dtype=np.int32)
    frow = np.ascontiguousarray(frow, dtype=np.int32)
    fcol = np.ascontiguousarray(fcol, dtype=np.int32)
    fval = np.zeros(len(fcol))
    rc = lib().orc_lu_refactor(len(ip_) - 1, _ip(ip_), _ip(ix), _dp(dv), _ip(P), _ip(Q), _ip(frow), _ip(fcol), _dp(fval))
    if rc != 0:
        raise ValueError(f"refactorization failed ({rc})")
    return fval


def direct_solve(indptr, indices, data, B):
    """FP64 sparse direct solve (SuperLU), the known answer of the reference's
    test (UMFPACK ``J \\ b``, test/cusolverRF.jl:12)."""
    n = len(indptr) - 1
    A = sp.csc_matrix(sp.csr_matrix((data, indices, indptr), shape=(n, n)))
    lu = spla.splu(A, permc_spec="COLAMD")
    return lu.solve(np.asarray(B, dtype=np.float64))


def backward_error(indptr, indices, data, X, B):
    """Per-column ||A x - b||_inf / (||A||_inf ||x||_inf + ||b||_inf) in FP64."""
    n = len(indptr) - 1
    A = sp.csr_matrix((data, indices, indptr), shape=(n, n))
    X = np.asarray(X, dtype=np.float64).reshape(n, -1)
    B = np.asarray(B, dtype=np.float64).reshape(n, -1)
    R = B - A @ X
    normA = abs(A).sum(axis=1).max()
    return np.abs(R).max(axis=0) / (normA * np.abs(X).max(axis=0) + np.abs(B).max(axis=0))


# --------------------------------------------------------------------------
# Block Krylov restatements (CPU arrays)
# --------------------------------------------------------------------------

class Stats:
    def __init__(self):
        self.niter = 0
        self.solved = False
        self.residuals = []
        self.timer = 0.0
        self.status = "unknown"


def _householder(Q, compact=False):
    """Q, R, tau = householder!(Q, R, tau; compact) -- src/block_utils.jl:46-53
    (LAPACK geqrf, copy of the upper triangle, orgqr unless compact)."""
    qr, tau, work, info = lapack.dgeqrf(np.asfortranarray(Q))
    k = Q.shape[1]
    R = np.triu(qr[:k, :k])
    if compact:
        return qr, R, tau
    q, work, info = lapack.dorgqr(qr, tau)
    return q, R, tau


def _ormqr_t(H, tau, D):
    """LAPACK.ormqr!('L','T', H, tau, D) (src/block_gmres.jl:323,336)."""
    out, work, info = lapack.dormqr("L", "T", H, tau, np.asfortranarray(D), max(1, D.shape[1] * 64))
    return out


def block_gmres(A, B, *, memory=20, M=None, N=None, restart=False, atol=None, rtol=None,
                itmax=0, history=False):
    """Restatement of block_gmres!(solver, A, B; M, N, ldiv, restart, ...)
    (src/block_gmres.jl:170-411).

    ``A`` is anything with ``A @ X``; ``M`` (left preconditioner) and ``N``
    (right preconditioner) are ``None`` (identity) or callables applying the
    preconditioner (what mulorldiv! does with the operator, src/block_utils.jl:113).
    The residual it estimates and stops on is ||M (B - A X)||_F (:226-227,:341).
    Returns ``(X, stats)``.
    """
    t0 = time.perf_counter()
    n, p = B.shape
    eps = np.finfo(np.float64).eps
    atol = np.sqrt(eps) if atol is None else atol
    rtol = np.sqrt(eps) if rtol is None else rtol
    mem = min(n // p, memory)                                   # :42
    stats = Stats()
    X = np.zeros((n, p), order="F")                              # :216
    W = np.array(B, dtype=np.float64, order="F")                 # :224
    if M is not None:
        W = np.asfortranarray(M(W))                              # :226
    RNorm = np.linalg.norm(W)                                    # :227
    if history:
        stats.residuals.append(RNorm)
    eps_stop = atol + rtol * RNorm                               # :230
    itmax = 2 * (n // p) if itmax == 0 else itmax                # :238
    inner_itmax = itmax
    it = 0
    solved = RNorm <= eps_stop
    tired = it >= itmax
    Xr = np.zeros((n, p), order="F") if restart else X
    npass = 0
    while not (solved or tired):                                 # :251
        V, Z, Rb, Hs, taus = [], [], {}, [], []
        if restart:
            Xr[:] = 0.0
            if npass >= 1:
                W = B - A @ X                                     # :268-269
                if M is not None:
                    W = M(W)                                      # :270
        Vk, Z1, _ = _householder(W)                               # :275-276
        V.append(Vk)
        Z.append(Z1)
        npass += 1
        inner = 0
        inner_tired = False
        nr = 0
        while not (solved or inner_tired):                       # :282
            inner += 1
            P = V[inner - 1] if N is None else N(V[inner - 1])   # :297-298
            Q = np.asfortranarray(A @ P)                         # :299
            if M is not None:
                Q = np.asfortranarray(M(Q))                      # :300
            for i in range(inner):                               # :301-304
                Psi = V[i].T @ Q
                Q = Q - V[i] @ Psi
                Rb[nr + i] = Psi
            Qn, C, tau_new = _householder(Q)                     # :316
            for i in range(inner - 1):                           # :320-326
                D = np.vstack([Rb[nr + i], Rb[nr + i + 1]])
                D = _ormqr_t(Hs[i], taus[i], D)
                Rb[nr + i] = D[:p]
                Rb[nr + i + 1] = D[p:]
            Hk = np.vstack([Rb[nr + inner - 1], C])              # :329-331
            hqr, Rkk, tk = _householder(Hk, compact=True)
            Hs.append(hqr)
            taus.append(tk)
            Rb[nr + inner - 1] = Rkk
            D = np.vstack([Z[inner - 1], np.zeros((p, p))])      # :334-337
            D = _ormqr_t(hqr, tk, D)
            Z[inner - 1] = D[:p]
            Cbar = D[p:]
            RNorm = np.linalg.norm(Cbar)                         # :341-342
            if history:
                stats.residuals.append(RNorm)
            nr += inner                                          # :346
            solved = RNorm <= eps_stop                           # :349
            inner_tired = inner >= min(mem, inner_itmax) if restart else inner >= inner_itmax
            if not (solved or inner_tired):                      # :356-363
                V.append(Qn)
                Z.append(Cbar.copy())
        # block back-substitution R Y = Z                         # :367-375
        Y = [z.copy() for z in Z[:inner]]
        for i in range(inner - 1, -1, -1):
            pos = nr + i - inner
            for j in range(inner - 1, i, -1):
                Y[i] = Y[i] - Rb[pos] @ Y[j]
                pos = pos - (j + 1) + 1
            Y[i] = scipy.linalg.solve_triangular(Rb[pos], Y[i], lower=False)
        acc = np.zeros((n, p))
        for i in range(inner):                                   # :378-380
            acc += V[i] @ Y[i]
        if N is not None:                                        # :381-384
            acc = N(acc)
        if restart:
            Xr[:] = acc
            X += Xr                                              # :385
        else:
            X[:] = acc
        inner_itmax -= inner
        it += inner
        tired = it >= itmax
    if tired:
        stats.status = "maximum number of iterations exceeded"
    if solved:
        stats.status = "solution good enough given atol and rtol"
    stats.niter = it
    stats.solved = bool(solved)
    stats.timer = time.perf_counter() - t0
    return X, stats


def block_bicgstab(A, B, *, C=None, N=None, atol=None, rtol=None, itmax=0, history=False,
                   callback=None):
    """Restatement of the CPU block_bicgstab! (src/block_bicgstab.jl:70-178).
    ``N`` is None or a callable applying N^{-1} (the reference calls
    mul!(Y, N, P) with an operator that applies the inverse, :134).
    ``callback(iter, X, R)``, if given, sees the iterate and the recursively
    updated residual after every iteration."""
    t0 = time.perf_counter()
    n, p = B.shape
    eps = np.finfo(np.float64).eps
    atol = np.sqrt(eps) if atol is None else atol
    rtol = np.sqrt(eps) if rtol is None else rtol
    C = B if C is None else C
    Ct = np.asarray(C).T
    stats = Stats()
    X = np.zeros((n, p))
    P = np.array(B, dtype=np.float64)
    R = np.array(B, dtype=np.float64)
    rNorm = np.linalg.norm(R)                                    # :114
    itmax = 2 * n if itmax == 0 else itmax                       # :117
    if history:
        stats.residuals.append(rNorm)
    eps_stop = atol + rtol * rNorm                               # :122
    it = 0
    solved = rNorm <= eps_stop
    tired = it >= itmax
    while not (solved or tired):
        Y = P if N is None else N(P)                             # :134
        V = A @ Y                                                # :135
        K = Ct @ V                                               # :136
        D = Ct @ R                                               # :137
        # (no finiteness check: a breakdown propagates as NaN into R, as
        # LAPACK getrf does in the reference, and ends the run below)
        Kf = scipy.linalg.lu_factor(K, check_finite=False)       # :138
        alpha = scipy.linalg.lu_solve(Kf, D, check_finite=False)  # :139
        U = R - V @ alpha                                        # :140-141
        X = X + Y @ alpha                                        # :142-143
        Y = U if N is None else N(U)                             # :144
        Q = A @ Y                                                # :145
        omega = abs(np.sum(Q * U)) / abs(np.sum(Q * Q))          # :146
        X = X + omega * Y                                        # :147
        R = U - omega * Q                                        # :148
        D = Ct @ Q                                               # :149
        beta = scipy.linalg.lu_solve(Kf, D, check_finite=False)  # :150
        U = P - omega * V                                        # :151
        P = R - U @ beta                                         # :152-153
        it += 1
        if callback is not None:
            callback(it, X, R)
        rNorm = np.linalg.norm(R)                                # :156
        if history:
            stats.residuals.append(rNorm)
        solved = rNorm <= eps_stop
        tired = it >= itmax
        if not np.isfinite(rNorm):
            # Breakdown (a singular C^T V): the reference would iterate on
            # NaN until itmax; the restatement stops here with status
            # "unknown", as the device solver does.
            tired = False
            break
    if tired:
        stats.status = "maximum number of iterations exceeded"
    if solved:
        stats.status = "solution good enough given atol and rtol"
    stats.niter = it
    stats.solved = bool(solved)
    stats.timer = time.perf_counter() - t0
    return X, stats
