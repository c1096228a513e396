"""Many independent systems A_i X_i = B_i of uneven size solved together.

No reference counterpart: the reference solves one system per call
(src/block_gmres.jl:127-411).  ``block_gmres_batched(As, Bs; kw...)`` returns
per-system solutions with the meaning of ``block_gmres(A_i, B_i; kw...)``:
each system has its own stopping test (eps_i = atol + rtol ||R0_i||_F, or a
per-RHS backward-error target when ``berr_tol > 0``).  Systems sharing an RHS
count are stacked into one block-diagonal operator on the device, so each
kernel launch covers all of them.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check

L = _lib.load()


class BatchedSystems:
    """Analysed batch: CSR patterns/values and RHS counts of every system,
    with the stacked operators and ILU(0) factors on the device."""

    def __init__(self, As, nrhs, ctx=None):
        from . import HipSparseMatrixCSR, default_context
        if len(As) == 0 or len(As) != len(nrhs):
            raise ValueError("need one RHS count per system")
        self.ctx = ctx or (As[0].ctx if isinstance(As[0], HipSparseMatrixCSR) else default_context())
        mats = []
        for A in As:
            if isinstance(A, HipSparseMatrixCSR):
                mats.append((A.indptr, A.indices, A.nzVal, A.n))
            else:
                A = A.tocsr()
                A.sort_indices()
                if A.shape[0] != A.shape[1]:
                    raise ValueError("System must be square")
                mats.append((np.ascontiguousarray(A.indptr, np.int32), np.ascontiguousarray(A.indices, np.int32),
                             np.ascontiguousarray(A.data, np.float64), A.shape[0]))
        self.nsys = len(mats)
        self.n = np.array([m[3] for m in mats], np.int32)
        self.p = np.array(nrhs, np.int32)
        nnz = np.array([int(m[0][-1]) for m in mats], np.int32)
        self._keep = mats
        rp = (_lib.ip * self.nsys)(*[m[0].ctypes.data_as(_lib.ip) for m in mats])
        ci = (_lib.ip * self.nsys)(*[m[1].ctypes.data_as(_lib.ip) for m in mats])
        va = (_lib.dp * self.nsys)(*[m[2].ctypes.data_as(_lib.dp) for m in mats])
        h = C.c_void_p()
        check(L.bpf_batch_create(self.ctx.handle, self.nsys, self.n.ctypes.data_as(_lib.ip),
                                 nnz.ctypes.data_as(_lib.ip), rp, ci, va, self.p.ctypes.data_as(_lib.ip),
                                 C.byref(h)))
        self.handle = h

    def solve(self, Bs, Xs=None, **kw):
        """Solves every system; returns (Xs, stats, solved) where solved[i]
        says whether system i met its stopping test."""
        from . import HipMatrix, SimpleStats, _fill_stats, _params
        if len(Bs) != self.nsys:
            raise ValueError("need one right-hand-side block per system")
        for i, B in enumerate(Bs):
            if B.shape != (int(self.n[i]), int(self.p[i])) or B.dtype != np.float64:
                raise ValueError(f"system {i}: B must be float64 of shape {(int(self.n[i]), int(self.p[i]))}")
        if Xs is None:
            Xs = [HipMatrix(B.shape, np.float64, self.ctx) for B in Bs]
        if kw.get("inner_width") == "auto":
            kw = dict(kw, inner_width=-1)
        prm = _params(**kw)
        st = _lib.SolverStats()
        solved = np.zeros(self.nsys, np.int32)
        bp = (C.c_void_p * self.nsys)(*[B.ptr for B in Bs])
        xp = (C.c_void_p * self.nsys)(*[X.ptr for X in Xs])
        check(L.bpf_batch_solve(self.handle, bp, xp, C.byref(prm), C.byref(st), solved.ctypes.data_as(_lib.ip)))
        stats = SimpleStats()
        _fill_stats(stats, st, np.zeros(0), False)
        return Xs, stats, solved.astype(bool)

    def inner_widths(self):
        """Columns per inner Krylov block the last solve used for each system
        (0 = no inner solve); with inner_width="auto" the width each bin's
        first solve chose by timing."""
        out = np.zeros(self.nsys, np.int32)
        check(L.bpf_batch_inner_widths(self.handle, out.ctypes.data_as(_lib.ip)))
        return out

    def destroy(self):
        # The context must still be alive: its stream and device own the
        # batch's buffers (at interpreter exit it may be gone first).
        if getattr(self, "handle", None) and self.ctx.handle:
            L.bpf_batch_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def block_gmres_batched(As, Bs, *, precision="mixed", **kw):
    """(Xs, stats, solved) = block_gmres_batched(As, Bs; kwargs...): As are
    HipSparseMatrixCSR or scipy sparse matrices, Bs are HipMatrix blocks
    (n_i x p_i).  Keywords as block_gmres (memory, restart, atol, rtol, itmax,
    timemax, berr_tol, precision "fp64" | "mixed", reorthogonalization,
    poly_degree, bmgs_blocks, inner_its, inner_width: columns per inner
    Krylov block or "auto" to time the candidates per RHS-count bin)."""
    for i, B in enumerate(Bs):
        if np.dtype(B.dtype) != np.float64:
            raise TypeError(f"system {i}: the batched solve takes float64 blocks, not {B.dtype}")
    batch = BatchedSystems(As, [B.shape[1] for B in Bs], ctx=Bs[0].ctx)
    try:
        return batch.solve(Bs, precision=precision, **kw)
    finally:
        batch.destroy()
