"""CUSOLVERRF: fixed-pattern LU refactorization with multi-RHS solves.

Mirror of the reference's submodule (src/cusolverRF.jl) over libbpf_amd.so:

    CusolverRfLU(A; nrhs=1, ordering=:AMD, tol=1e-8, fast_mode=true)   :222-301
    update_(rflu, A)        update!(rflu, A)                           :304-314
    ldiv_(x, rflu, b)       ldiv!(x, rflu, b)                          :317-324
    ldiv_(rflu, x)          ldiv!(rflu, x)                             :325-331
    lu(A; kw...)            LinearAlgebra.lu(A)                        :333
    lu_(rflu, A)            LinearAlgebra.lu!(rflu, A)                 :334
    CusolverRfLUBatch(A, batchsize; ordering, tol, fast_mode)         :336-419
    update_(rflub, A), solve_(xs, rflub, bs), ldiv_(X, rflub, B),
    ldiv_(rflub, X), rf_batch_refactor_(rflub, Anzval)                :421-481
    cudestroy_(rf)          cudestroy!(rf)                             :54

Differences from the reference, each deliberate: any nrhs is accepted (the
reference errors for nrhs > 1, :228-230); a pivot below ``tol`` raises
SingularException (the reference constructs it without throwing, :164-166);
the batch ``update!`` sets every batch entry to A's values (the reference
passes one pointer for all entries, :423).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check

L = _lib.load()


def _ordering(o):
    key = str(o).lstrip(":").upper()
    if key not in _lib.ORDERINGS:
        raise ValueError(f"unknown ordering {o!r}; expected one of :AMD, :MDQ, :METIS, :RCM")
    return _lib.ORDERINGS[key]


class _RfBase:
    handle = None

    def _info(self):
        n, nf, nl, nu = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(L.bpf_rf_info(self.handle, C.byref(n), C.byref(nf), C.byref(nl), C.byref(nu)))
        return n.value, nf.value, nl.value, nu.value

    @property
    def nnz_factor(self):
        return self._info()[1]

    def levels(self):
        return self._info()[2:]

    def factors(self):
        """(P, Q, rowptr, colidx, values) of B = A[P, Q] = L U (combined L\\U CSR)."""
        n, nf, _, _ = self._info()
        P = np.empty(n, np.int32)
        Q = np.empty(n, np.int32)
        rp = np.empty(n + 1, np.int32)
        ci = np.empty(nf, np.int32)
        v = np.empty(nf)
        check(L.bpf_rf_get_factor(self.handle, P.ctypes.data_as(_lib.ip), Q.ctypes.data_as(_lib.ip),
                                  rp.ctypes.data_as(_lib.ip), ci.ctypes.data_as(_lib.ip),
                                  v.ctypes.data_as(_lib.dp)))
        return P, Q, rp, ci, v

    def destroy(self):
        if self.handle and self.ctx.handle:
            L.bpf_rf_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class CusolverRfLU(_RfBase):
    """Refactorization handle (src/cusolverRF.jl:26-37, 222-301)."""

    def __init__(self, A, *, nrhs=1, ordering=":AMD", tol=1e-8, fast_mode=True):
        if A.shape[0] != A.shape[1]:
            raise AssertionError("only square matrices are supported")
        self.ctx = A.ctx
        self.n = self.m = A.n
        self.nnzA = A.nnz()
        self.nrhs = nrhs
        h = C.c_void_p()
        check(L.bpf_rf_create(self.ctx.handle, A.handle, _ordering(ordering), float(tol), C.byref(h)))
        self.handle = h
        self.refine_steps = 0

    def solve_refined(self, A, B, refine=2, X=None, berr_tol=0.0):
        """FP32 triangular solves + at most `refine` FP64 refinement steps;
        with berr_tol > 0 refinement stops once every RHS's backward error is
        <= berr_tol.  Returns (X, per-RHS backward error); X (optional)
        receives the solution and `self.refine_steps` the steps taken."""
        from . import HipMatrix
        if X is None:
            X = HipMatrix(B.shape, np.float64, B.ctx)
        elif X.shape != B.shape or X.dtype != np.float64:
            raise ValueError("X must be a float64 matrix of B's shape")
        berr = np.zeros(B.ncols)
        steps = C.c_int(0)
        check(L.bpf_rf_solve_refined(self.handle, A.handle, B.ncols, B.ptr, B.nrows, X.ptr, X.nrows,
                                     int(refine), float(berr_tol), berr.ctypes.data_as(_lib.dp),
                                     C.byref(steps)))
        self.refine_steps = steps.value
        return X, berr


class CusolverRfLUBatch(_RfBase):
    """Batched refactorization handle: one pattern, `batchsize` value sets
    (src/cusolverRF.jl:41-52, 336-419)."""

    def __init__(self, A, batchsize, *, ordering=":AMD", tol=1e-8, fast_mode=False):
        if A.shape[0] != A.shape[1]:
            raise AssertionError("only square matrices are supported")
        if batchsize == 1:
            raise ValueError("Please specify a number of batch greater than 1 for `CuSolverRfLUBatch`")
        self.ctx = A.ctx
        self.n = self.m = A.n
        self.nnzA = A.nnz()
        self.batchsize = int(batchsize)
        h = C.c_void_p()
        check(L.bpf_rf_batch_create(self.ctx.handle, A.handle, self.batchsize, _ordering(ordering),
                                    float(tol), C.byref(h)))
        self.handle = h


def _ldiv(x, rf, b):
    if isinstance(rf, CusolverRfLUBatch):
        if not (x.ncols == b.ncols == rf.batchsize):
            raise AssertionError("X and B must have batchsize columns")
        check(L.bpf_rf_batch_solve(rf.handle, b.ptr, b.nrows, x.ptr, x.nrows))
        return x
    dt = _lib.BPF_F64
    check(L.bpf_rf_solve(rf.handle, dt, b.ncols, b.ptr, b.nrows, x.ptr, x.nrows))
    return x


def _ldiv_inplace(rf, x):
    if isinstance(rf, CusolverRfLUBatch):
        if x.ncols != rf.batchsize:
            raise AssertionError("X must have batchsize columns")
        check(L.bpf_rf_batch_solve(rf.handle, x.ptr, x.nrows, x.ptr, x.nrows))
        return x
    check(L.bpf_rf_solve(rf.handle, _lib.BPF_F64, x.ncols, x.ptr, x.nrows, x.ptr, x.nrows))
    return x


def update_(rf, A):
    """update!(rflu, A): new values of A, same pattern and pivots."""
    if isinstance(rf, CusolverRfLUBatch):
        vals = np.tile(A.nzVal, (rf.batchsize, 1)).T
        return rf_batch_refactor_(rf, vals)
    check(L.bpf_rf_refactor(rf.handle, A.handle))


def rf_batch_refactor_(rf, Anzval):
    """rf_batch_refactor!(rflu, Anzval): Anzval is nnz x batchsize (one value
    set per column), host array or HipMatrix."""
    from . import HipMatrix
    M = Anzval if isinstance(Anzval, HipMatrix) else HipMatrix.from_numpy(np.asfortranarray(Anzval), rf.ctx)
    check(L.bpf_rf_batch_refactor(rf.handle, M.ptr, M.nrows))


def solve_(xs, rf, bs):
    """solve!(x::Vector{CuVector}, rflu, b::Vector{CuVector}) for the batch."""
    from . import HipMatrix
    B = HipMatrix.from_numpy(np.column_stack([b.to_numpy().ravel() for b in bs]), rf.ctx)
    X = HipMatrix(B.shape, np.float64, rf.ctx)
    _ldiv(X, rf, B)
    Xh = X.to_numpy()
    for i, x in enumerate(xs):
        x.copy_(HipMatrix.from_numpy(Xh[:, i].reshape(x.shape), rf.ctx))
    return xs


def lu(A, **options):
    """LinearAlgebra.lu(A; options...) = CusolverRfLU(A; options...)."""
    return CusolverRfLU(A, **options)


def lu_(rf, A):
    """LinearAlgebra.lu!(rflu, A) = update!(rflu, A)."""
    return update_(rf, A)


def cudestroy_(rf):
    rf.destroy()


class CUSOLVERRF:
    """Namespace object mirroring the reference's `CUSOLVERRF` submodule."""
    _RfBase = _RfBase
    CusolverRfLU = CusolverRfLU
    CusolverRfLUBatch = CusolverRfLUBatch
    update_ = staticmethod(update_)
    rf_batch_refactor_ = staticmethod(rf_batch_refactor_)
    solve_ = staticmethod(solve_)
    lu = staticmethod(lu)
    lu_ = staticmethod(lu_)
    cudestroy_ = staticmethod(cudestroy_)
    _ldiv = staticmethod(_ldiv)
    _ldiv_inplace = staticmethod(_ldiv_inplace)
