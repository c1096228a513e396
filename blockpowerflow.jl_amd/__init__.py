"""MI355X-native multi-right-hand-side sparse solves (host mirror of BlockPowerFlow.jl).

The entry points keep the reference's names, argument meaning and error
behaviour; Julia's mutating ``f!`` becomes ``f_`` here:

    block_gmres(A, B[, X0]; memory, M, N, ldiv, restart, reorthogonalization,
                atol, rtol, itmax, timemax, verbose, history)   src/block_gmres.jl:127-154
    block_gmres_(solver, A, B[, X0]; ...)                        src/block_gmres.jl:156-173
    BlockGMRESSolver(A, B; memory)                               src/block_gmres.jl:62-68
    block_bicgstab(A, B; C, N, atol, rtol, itmax, timemax, verbose, history)
                                                                 src/block_bicgstab.jl:182-190
    ilu0(A[, nrhs]), ic0(A[, nrhs]), ldiv_(y, P, x)              src/preconditioners_*.jl
    CUSOLVERRF.*                                                 src/cusolverRF.jl

Every computation runs in libbpf_amd.so (hand-written gfx950 kernels); this
module only moves arguments across the C ABI.  ``HipMatrix`` (device dense,
column-major) and ``HipSparseMatrixCSR`` (device CSR, int32 indices) play the
roles of CuMatrix / CuSparseMatrixCSR (or ROCMatrix / ROCSparseMatrixCSR).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import BpfError, SingularException, check

__all__ = [
    "Context", "default_context", "HipMatrix", "HipSparseMatrixCSR",
    "ilu0", "ic0", "ldiv_", "mul_", "block_gmres", "block_gmres_", "BlockGMRESSolver",
    "block_bicgstab", "block_bicgstab_", "BlockBicgstabSolver", "BlockGMRESStats", "SimpleStats",
    "backward_error", "spmm", "CUSOLVERRF", "BpfError", "SingularException", "BatchedSystems",
    "block_gmres_batched", "householder", "block_hessenberg_lsq", "set_option", "get_option",
    "HostCSR", "mmread", "mmread_dense", "mmwrite", "BlockJacobiPreconditioner", "update",
]

L = _lib.load()


class Context:
    """A device + stream (the implicit task-local device of CUDA.jl/AMDGPU.jl)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(L.bpf_ctx_create(device, C.byref(h)))
        self.handle = h
        self.device = device

    def synchronize(self):
        check(L.bpf_ctx_synchronize(self.handle))

    def timer_start(self):
        check(L.bpf_timer_start(self.handle))

    def timer_stop(self) -> float:
        ms = C.c_double()
        check(L.bpf_timer_stop(self.handle, C.byref(ms)))
        return ms.value

    def close(self):
        if self.handle:
            L.bpf_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_DEFAULT = None


def default_context() -> Context:
    global _DEFAULT
    if _DEFAULT is None:
        _DEFAULT = Context(0)
    return _DEFAULT


_OPTIONS = {"syncfree": 1, "autotune": 2, "overlap": 3, "batch_threads": 4, "wide_min": 5, "graph": 6,
            "graph_replays": 7, "relabel_min": 8, "tail_slab": 9,
            "dense_inverse": 10, "dense_blocks": 11, "tail_rounds": 12}


def set_option(name: str, value) -> None:
    """Process-wide execution option: "syncfree" (default False) runs each
    triangular solve / factorization as one sync-free launch instead of one
    launch per wide dependency level; "autotune" (default True) times the
    SpMM launch shapes on the first product of each matrix and width;
    "overlap" (default True) runs block GMRES's Hessenberg update of each
    iteration on a second stream beside the next iteration; "batch_threads"
    (default 4) is the number of RHS-count bins of a batched solve solved at
    the same time, each from its own host thread on its own streams;
    "wide_min" (default 129) is the smallest number of right-hand sides that
    FP64 block GMRES runs as one n x p Krylov block with its dense algebra on
    rocBLAS / rocSOLVER, as the reference runs every p
    (src/block_gmres.jl:46-56); narrower blocks run on the library's own
    kernels; "graph" (default True) captures the mixed form's inner FP32
    block GMRES run (its fixed chain of inner_its iterations) into a HIP graph
    on its first call and replays it on later outer iterations;
    "graph_replays" counts those replays (reset with 0); "relabel_min"
    (default 32768, 0 = never) is the smallest system that block GMRES in
    the "mixed" precision solves in a reverse Cuthill-McKee relabelling of
    A's pattern, for the locality of the rows its products and triangular
    solves gather (the "fp64" form keeps A's row order); "tail_slab" (default
    64, or 0) is the column width of the slabs that the narrow-level part of
    a triangular solve over more right-hand sides than that is split into, one
    workgroup per slab (0 = one workgroup for all columns; results are
    identical); "dense_inverse" (default 1) solves the dense tail of a
    refactorization handle by explicit inverses of its triangular factors,
    built by the first solve that uses them after each refactorization: 0 =
    never (blocked substitution), 1 = in the FP32 solves that refinement
    follows, 2 = in every solve; "dense_blocks" (default 12, 0 = none) is the
    most 128-row blocks of the dense tail of a refactorization handle created
    afterwards (at most about n / 16 rows); "tail_rounds" (default 2) admits a
    triangular-solve level to the single narrow-level launch when its rows fit
    in that many rounds of the launch's wavefronts (results are identical)."""
    check(L.bpf_set_option(_OPTIONS[name], int(value)))


def get_option(name: str) -> int:
    v = C.c_longlong()
    check(L.bpf_get_option(_OPTIONS[name], C.byref(v)))
    return v.value


def device_count() -> int:
    c = C.c_int()
    check(L.bpf_device_count(C.byref(c)))
    return c.value


class HipMatrix:
    """Dense device matrix, column-major (Fortran order), float64 or float32."""

    def __init__(self, shape, dtype=np.float64, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        if isinstance(shape, int):
            shape = (shape,)
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        nbytes = max(1, int(np.prod(self.shape))) * self.dtype.itemsize
        p = C.c_void_p()
        check(L.bpf_malloc(self.ctx.handle, C.byref(p), nbytes))
        self.ptr = p
        self.nbytes = nbytes

    @classmethod
    def from_numpy(cls, a, ctx: Context | None = None, dtype=None):
        a = np.asarray(a)
        dt = np.dtype(dtype) if dtype is not None else (a.dtype if a.dtype in (np.float32, np.float64) else np.float64)
        a = np.asfortranarray(a, dtype=dt)
        m = cls(a.shape, dt, ctx)
        if a.size:
            check(L.bpf_memcpy_h2d(m.ctx.handle, m.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes))
        return m

    def to_numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype, order="F")
        if out.size:
            check(L.bpf_memcpy_d2h(self.ctx.handle, out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes))
        return out

    def copy_(self, other: "HipMatrix"):
        if other.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        check(L.bpf_memcpy_d2d(self.ctx.handle, self.ptr, other.ptr, self.nbytes))
        return self

    def copy(self) -> "HipMatrix":
        m = HipMatrix(self.shape, self.dtype, self.ctx)
        return m.copy_(self)

    @property
    def nrows(self):
        return self.shape[0]

    @property
    def ncols(self):
        return self.shape[1] if len(self.shape) > 1 else 1

    def free(self):
        if getattr(self, "ptr", None):
            L.bpf_free(self.ctx.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ctx.handle:
                self.free()
        except Exception:
            pass

    def __len__(self):
        return self.shape[0]


Array = HipMatrix.to_numpy


class HipSparseMatrixCSR:
    """Square device CSR matrix (int32 indices, float64 values)."""

    def __init__(self, indptr, indices, data, n: int | None = None, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        indptr = np.ascontiguousarray(indptr, dtype=np.int32)
        indices = np.ascontiguousarray(indices, dtype=np.int32)
        data = np.ascontiguousarray(data, dtype=np.float64)
        n = len(indptr) - 1 if n is None else n
        h = C.c_void_p()
        check(L.bpf_csr_create(self.ctx.handle, n, int(indptr[-1]) if len(indptr) else 0,
                               indptr.ctypes.data_as(_lib.ip), indices.ctypes.data_as(_lib.ip),
                               data.ctypes.data_as(_lib.dp), C.byref(h)))
        self.handle = h
        self.n = n
        self.indptr = indptr
        self.indices = indices
        self._data = data.copy()
        self._dev_src = None

    @classmethod
    def from_scipy(cls, A, ctx=None):
        A = A.tocsr()
        A.sort_indices()
        return cls(A.indptr, A.indices, A.data, A.shape[0], ctx)

    @property
    def shape(self):
        return (self.n, self.n)

    def nnz(self):
        return int(self.indptr[-1])

    @property
    def nzVal(self) -> np.ndarray:
        """Host copy of the values (A.nzVal)."""
        if self._data is None:       # last set from a device array
            M, col = self._dev_src
            self._data = np.ascontiguousarray(M.to_numpy()[:, col])
            self._dev_src = None
        return self._data.copy()

    def set_values_device(self, M, col: int = 0):
        """New values from column `col` of the device matrix M (nnz x k,
        float64): a Newton step's Jacobian values already resident in HBM."""
        if M.dtype != np.float64 or M.nrows != self.nnz() or not 0 <= col < M.ncols:
            raise ValueError("values must be column col of an nnz-row float64 device matrix")
        ptr = C.c_void_p(M.ptr.value + col * M.nrows * 8)
        check(L.bpf_csr_set_values_device(self.handle, ptr))
        self._data = None
        self._dev_src = (M, col)

    def set_values(self, data):
        data = np.ascontiguousarray(data, dtype=np.float64)
        if data.shape != (self.nnz(),):
            raise ValueError("values must have nnz entries")
        check(L.bpf_csr_set_values(self.handle, data.ctypes.data_as(_lib.dp)))
        self._data = data.copy()
        self._dev_src = None

    def norm_inf(self) -> float:
        v = C.c_double()
        check(L.bpf_csr_info(self.handle, None, None, C.byref(v)))
        return v.value

    def __matmul__(self, X: HipMatrix) -> HipMatrix:
        Y = HipMatrix(X.shape, X.dtype, self.ctx)
        mul_(Y, self, X)
        return Y

    def __del__(self):
        try:
            if self.handle and self.ctx.handle:
                L.bpf_csr_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def _dtype_code(dt):
    return _lib.BPF_F32 if np.dtype(dt) == np.float32 else _lib.BPF_F64


def mul_(Y: HipMatrix, A, X: HipMatrix, alpha: float = 1.0, beta: float = 0.0):
    """mul!(Y, A, X[, alpha, beta]) for a CSR A (SpMM) or a preconditioner
    (applies its inverse, the meaning block_bicgstab's N has)."""
    if isinstance(A, _Preconditioner):
        return ldiv_(Y, A, X)
    if X.dtype != Y.dtype:
        raise TypeError("X and Y must have the same element type")
    p = X.ncols
    check(L.bpf_spmm(A.ctx.handle, A.handle, _dtype_code(X.dtype), _lib.BPF_COL_MAJOR, p, X.ptr,
                     X.nrows, Y.ptr, Y.nrows, alpha, beta))
    return Y


spmm = mul_


# ---------------------------------------------------------------- preconditioners
class _Preconditioner:
    kind = _lib.BPF_PREC_ILU0

    def __init__(self, A: HipSparseMatrixCSR, nrhs: int | None = None):
        self.ctx = A.ctx
        self.n = A.n
        self.nrhs = nrhs
        h = C.c_void_p()
        check(L.bpf_prec_create(self.ctx.handle, A.handle, self.kind, C.byref(h)))
        self.handle = h
        self.pattern = (A.indptr, A.indices)

    def factor_(self, A: HipSparseMatrixCSR):
        """Numeric refactorization with A's current values (same pattern)."""
        check(L.bpf_prec_factor(self.handle, A.handle))
        return self

    def factor_values(self) -> np.ndarray:
        out = np.empty(int(self.pattern[0][-1]), dtype=np.float64)
        check(L.bpf_prec_get_factor(self.handle, out.ctypes.data_as(_lib.dp)))
        return out

    @property
    def P(self):
        """The combined factor as (indptr, indices, values) on the host."""
        return self.pattern[0], self.pattern[1], self.factor_values()

    def levels(self):
        a, b = C.c_int(), C.c_int()
        check(L.bpf_prec_levels(self.handle, C.byref(a), C.byref(b)))
        return a.value, b.value

    def __del__(self):
        try:
            if self.handle and self.ctx.handle:
                L.bpf_prec_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class ILU0(_Preconditioner):
    """ILU(0) (NVIDIA_ILU0 / AMD_ILU0, src/preconditioners_nvidia.jl:37-47,
    src/preconditioners_amd.jl:39-46)."""
    kind = _lib.BPF_PREC_ILU0


class IC0(_Preconditioner):
    """IC(0) (NVIDIA_IC0 / AMD_IC0, src/preconditioners_nvidia.jl:1-11,
    src/preconditioners_amd.jl:1-9)."""
    kind = _lib.BPF_PREC_IC0


class BlockJacobiPreconditioner(_Preconditioner):
    """LinearSolvers.BlockJacobiPreconditioner(J, nblocks, device) of ExaPF, the
    right preconditioner the reference's benchmark and tests hand to
    block_gmres / block_bicgstab (benchmarks/benchmarks_nvidia.jl:46-51,71;
    test/block_gmres.jl:5,18-25): nblocks parts of the symmetrized graph of
    A (at most 128 unknowns each), dense inverses of the diagonal blocks on
    the device.  ``mul_(y, P, x)`` applies them; ``update(P, A)`` recomputes
    them from A's current values."""
    kind = _lib.BPF_PREC_BLOCK_JACOBI

    def __init__(self, A: HipSparseMatrixCSR, nblocks: int, device=None):
        self.ctx = A.ctx
        self.n = A.n
        self.nrhs = None
        self.nblocks = int(nblocks)
        h = C.c_void_p()
        check(L.bpf_block_jacobi_create(self.ctx.handle, A.handle, self.nblocks, C.byref(h)))
        self.handle = h
        self.pattern = (A.indptr, A.indices)

    @property
    def partition(self) -> np.ndarray:
        """Part (0..nblocks-1) of every unknown."""
        out = np.empty(self.n, dtype=np.int32)
        check(L.bpf_block_jacobi_partition(self.handle, out.ctypes.data_as(_lib.ip)))
        return out


def update(P, A: HipSparseMatrixCSR, device=None):
    """LinearSolvers.update(P, A, device): refactor a preconditioner with A's
    current values (same pattern)."""
    return P.factor_(A)


def ilu0(A: HipSparseMatrixCSR, nrhs: int | None = None) -> ILU0:
    """ilu0(A, nrhs) / ilu0(A): ILU(0) on A's pattern."""
    return ILU0(A, nrhs)


def ic0(A: HipSparseMatrixCSR, nrhs: int | None = None) -> IC0:
    """ic0(A, nrhs) / ic0(A): IC(0) on A's lower pattern."""
    return IC0(A, nrhs)


def ldiv_(y, P, x=None):
    """ldiv!(y, P, x): y = P^{-1} x for an ILU0/IC0 preconditioner or a
    CUSOLVERRF factorization (src/preconditioners_nvidia.jl:49-59,
    src/cusolverRF.jl:317-331).  ``ldiv_(rf, x)`` solves in place."""
    if isinstance(y, CUSOLVERRF._RfBase) and x is None:
        return CUSOLVERRF._ldiv_inplace(y, P)
    if isinstance(P, CUSOLVERRF._RfBase):
        return CUSOLVERRF._ldiv(y, P, x)
    if y.dtype != x.dtype:
        raise TypeError("x and y must have the same element type")
    dt = _block_dtype(x)
    if y.shape != x.shape or x.nrows != P.n:
        raise ValueError(f"blocks of size {x.shape} / {y.shape} for a preconditioner of size {P.n}")
    check(L.bpf_prec_apply(P.handle, _dtype_code(dt), _lib.BPF_COL_MAJOR, x.ncols, x.ptr, x.nrows,
                           y.ptr, y.nrows))
    return y


# ---------------------------------------------------------------- stats
class BlockGMRESStats:
    """BlockGMRESStats (src/block_utils.jl:62-68) plus backward-error fields."""

    def __init__(self):
        self.niter = 0
        self.solved = False
        self.residuals: list[float] = []
        self.timer = 0.0
        self.status = "unknown"
        self.refine_steps = 0
        self.max_berr = math.nan

    def __repr__(self):
        return (f"{type(self).__name__}(niter={self.niter}, solved={self.solved}, "
                f"status={self.status!r}, timer={self.timer:.3g}s)")


class SimpleStats(BlockGMRESStats):
    """SimpleStats (src/block_bicgstab.jl:15-21)."""


_STATUS = {
    _lib.BPF_STATUS_SOLVED: "solution good enough given atol and rtol",
    _lib.BPF_STATUS_TIRED: "maximum number of iterations exceeded",
    _lib.BPF_STATUS_OVERTIMED: "time limit exceeded",
    _lib.BPF_STATUS_UNKNOWN: "unknown",
}


def _fill_stats(stats, st: _lib.SolverStats, hist, history):
    stats.niter = st.niter
    stats.solved = bool(st.solved)
    stats.status = _STATUS.get(st.status, "unknown")
    stats.timer = st.timer
    stats.refine_steps = st.refine_steps
    stats.max_berr = st.max_berr
    stats.residuals = list(hist[: min(st.nresiduals, len(hist))]) if history else []


def _params(memory=20, restart=False, reorthogonalization=False, atol=None, rtol=None, itmax=0,
            timemax=math.inf, verbose=0, history=False, berr_tol=0.0, precision="fp64",
            block_width=0, max_refine=10, inner_rtol=1e-6, poly_degree=0, bmgs_blocks=1, inner_its=0,
            inner_width=0):
    prm = _lib.SolverParams()
    L.bpf_solver_params_default(C.byref(prm))
    eps = math.sqrt(np.finfo(np.float64).eps)
    prm.memory = int(memory)
    prm.restart = int(bool(restart))
    prm.reorthogonalization = int(bool(reorthogonalization))
    prm.atol = eps if atol is None else float(atol)
    prm.rtol = eps if rtol is None else float(rtol)
    prm.itmax = int(itmax)
    prm.timemax = 0.0 if (timemax is None or math.isinf(timemax)) else float(timemax)
    prm.verbose = int(verbose)
    prm.history = int(bool(history))
    prm.berr_tol = float(berr_tol)
    prm.precision = {"fp64": _lib.BPF_PRECISION_FP64, "mixed": _lib.BPF_PRECISION_MIXED,
                     "mixed32": _lib.BPF_PRECISION_MIXED32}[precision]
    prm.block_width = int(block_width)
    prm.max_refine = int(max_refine)
    prm.inner_rtol = float(inner_rtol)
    prm.poly_degree = int(poly_degree)
    prm.bmgs_blocks = int(bmgs_blocks)
    prm.inner_its = int(inner_its)
    prm.inner_width = int(inner_width)
    return prm


# ---------------------------------------------------------------- block GMRES
class BlockGMRESSolver:
    """Workspace of block_gmres! (src/block_gmres.jl:21-68).  ``memory`` is
    clamped to div(n, p) as in the reference; X aliases the returned solution."""

    def __init__(self, A, B, memory: int = 5):
        n, p = A.n, B.ncols
        self.m, self.n, self.p = n, n, p
        self.memory = max(1, min(n // max(p, 1), memory))
        self.X = HipMatrix((n, p), _block_dtype(B), A.ctx)
        self.stats = BlockGMRESStats()
        self.warm_start = False
        self._X0 = None

    def warm_start_(self, X0: HipMatrix):
        """warm_start!(solver, X0) (src/block_gmres.jl:70-79; the reference's
        field typo Δx/ΔX is not reproduced)."""
        if X0.shape != self.X.shape:
            raise ValueError(f"X0 should have size ({self.n}, {self.p})")
        if X0.dtype != self.X.dtype:
            raise TypeError(f"X0 has element type {X0.dtype}, the solver {self.X.dtype}")
        self._X0 = X0
        self.warm_start = True
        return self


def _block_dtype(B) -> np.dtype:
    """Element type of a right-hand-side block: Float64 or Float32 (the
    reference's solvers are generic over it, src/block_gmres.jl:170-173);
    anything else raises before any native call."""
    dt = np.dtype(B.dtype)
    if dt not in (np.float64, np.float32):
        raise TypeError(f"element type {dt} is not supported (float64 or float32)")
    return dt


def _check_system(A, B, *others):
    """Shape and element-type checks of block_gmres! / block_bicgstab!
    (src/block_gmres.jl:180-191): A square, B with n rows, and every other
    block (X, X0, C) of B's shape and element type."""
    if A.shape[0] != A.shape[1]:
        raise ValueError("System must be square")
    if A.n != B.nrows:
        raise ValueError("Inconsistent problem size")
    dt = _block_dtype(B)
    for M in others:
        if M is None:
            continue
        if M.shape != B.shape:
            raise ValueError(f"block of size {M.shape} differs from B's {B.shape}")
        if np.dtype(M.dtype) != dt:
            raise TypeError(f"block of element type {M.dtype} differs from B's {dt}")
    return dt


def _default_tol(dt) -> float:
    """sqrt(eps(T)), the reference's default atol and rtol."""
    return math.sqrt(float(np.finfo(dt).eps))


def block_gmres_(solver: BlockGMRESSolver, A: HipSparseMatrixCSR, B: HipMatrix, X0=None, *, M=None,
                 N=None, ldiv=False, restart=False, reorthogonalization=False, atol=None, rtol=None,
                 itmax=0, timemax=math.inf, verbose=0, history=False, berr_tol=0.0,
                 precision="fp64", block_width=0, max_refine=10, inner_rtol=1e-6, poly_degree=0,
                 bmgs_blocks=1, inner_its=0, inner_width=0):
    """block_gmres!(solver, A, B[, X0]; kwargs...) -> solver (src/block_gmres.jl:156-411)."""
    dt = _check_system(A, B, solver.X, X0, solver._X0 if solver.warm_start else None)
    tol = _default_tol(dt)
    atol = tol if atol is None else atol
    rtol = tol if rtol is None else rtol
    for name, P in (("M", M), ("N", N)):
        if P is not None and not isinstance(P, _Preconditioner):
            raise TypeError(f"{name} must be a preconditioner object of this package")
        if P is not None and P.n != A.n:
            raise ValueError(f"{name} has size {P.n}, the system {A.n}")
    if X0 is not None:
        solver.warm_start_(X0)
    prm = _params(solver.memory, restart, reorthogonalization, atol, rtol, itmax, timemax, verbose,
                  history, berr_tol, precision, block_width, max_refine, inner_rtol, poly_degree,
                  bmgs_blocks, inner_its, inner_width)
    st = _lib.SolverStats()
    cap = 1 + (itmax if itmax > 0 else 2 * (A.n // max(1, B.ncols))) + 16
    hist = np.zeros(cap)
    x0 = solver._X0 if solver.warm_start else None
    # mulorldiv!(y, P, x, ldiv) (src/block_utils.jl:113): ldiv! and mul! of a
    # preconditioner object both apply its inverse approximation here, so
    # ldiv only matters to the Julia caller's dispatch.
    Mh = M.handle if M is not None else None
    Nh = N.handle if N is not None else None
    check(L.bpf_block_gmres_ex(A.ctx.handle, A.handle, Mh, Nh, _dtype_code(dt), B.ncols, B.ptr, B.nrows,
                              x0.ptr if x0 is not None else None, x0.nrows if x0 is not None else 0,
                              solver.X.ptr, solver.X.nrows, C.byref(prm), C.byref(st),
                              hist.ctypes.data_as(_lib.dp), cap))
    solver.warm_start = False
    solver._X0 = None
    _fill_stats(solver.stats, st, hist, history)
    return solver


def block_gmres(A: HipSparseMatrixCSR, B: HipMatrix, X0: HipMatrix | None = None, *, memory=20,
                **kw):
    """(X, stats) = block_gmres(A, B[, X0]; memory, M, N, ldiv, restart,
    reorthogonalization, atol, rtol, itmax, timemax, verbose, history)
    (src/block_gmres.jl:127-154).  Extra keywords: berr_tol, precision
    ("fp64" | "mixed" | "mixed32"), block_width, max_refine, inner_rtol,
    poly_degree, bmgs_blocks (basis blocks projected together per block-MGS
    step; 1 = the reference's one-block loop), inner_its / inner_width (mixed:
    the preconditioner step is inner_its iterations of an FP32 block GMRES on
    groups of inner_width columns)."""
    solver = BlockGMRESSolver(A, B, memory=memory)
    block_gmres_(solver, A, B, X0, **kw)
    return solver.X, solver.stats


# ---------------------------------------------------------------- block BiCGSTAB
class BlockBicgstabSolver:
    """Workspace of block_bicgstab! (src/block_bicgstab.jl:25-58)."""

    def __init__(self, A, B):
        self.X = HipMatrix((A.n, B.ncols), _block_dtype(B), A.ctx)
        self.stats = SimpleStats()


def block_bicgstab_(solver, A, B, *, C=None, N=None, atol=None, rtol=None, itmax=0,
                    timemax=math.inf, verbose=0, history=False):
    """block_bicgstab!(solver, A, B; C=B, N, ...) -> (X, stats)
    (src/block_bicgstab.jl:187-299).  N (an ILU0/IC0 object) is applied as
    mul!(Y, N, P), i.e. as its inverse."""
    dt = _check_system(A, B, solver.X, C)
    tol = _default_tol(dt)
    atol = tol if atol is None else atol
    rtol = tol if rtol is None else rtol
    prm = _params(atol=atol, rtol=rtol, itmax=itmax, timemax=timemax, verbose=verbose, history=history)
    st = _lib.SolverStats()
    cap = 1 + (itmax if itmax > 0 else 2 * A.n) + 16
    cap = min(cap, 1 << 20)
    hist = np.zeros(cap)
    Ch = C.ptr if C is not None else None
    check(L.bpf_block_bicgstab_t(A.ctx.handle, A.handle, N.handle if N is not None else None,
                                 _dtype_code(dt), B.ncols, B.ptr, B.nrows, Ch,
                                 C.nrows if C is not None else 0, solver.X.ptr, solver.X.nrows,
                                 _ctypes_byref(prm), _ctypes_byref(st), hist.ctypes.data_as(_lib.dp),
                                 cap))
    _fill_stats(solver.stats, st, hist, history)
    return solver.X, solver.stats


def _ctypes_byref(x):
    return C.byref(x)


def block_bicgstab(A, B, **kw):
    """(X, stats) = block_bicgstab(A, B; C=B, N, atol, rtol, itmax, timemax,
    verbose, history) (src/block_bicgstab.jl:182-185)."""
    solver = BlockBicgstabSolver(A, B)
    return block_bicgstab_(solver, A, B, **kw)


def hessenberg_eigenvalues(H) -> np.ndarray:
    """Eigenvalues of an upper Hessenberg matrix (host utility of the library)."""
    H = np.ascontiguousarray(H, dtype=np.float64)
    d = H.shape[0]
    re, im = np.zeros(d), np.zeros(d)
    check(L.bpf_hessenberg_eigenvalues(H.ctypes.data_as(_lib.dp), d, re.ctypes.data_as(_lib.dp),
                                       im.ctypes.data_as(_lib.dp)))
    return re + 1j * im


def backward_error(A: HipSparseMatrixCSR, X: HipMatrix, B: HipMatrix) -> np.ndarray:
    """Per-RHS FP64 backward error ||Ax-b||inf / (||A||inf ||x||inf + ||b||inf)."""
    out = np.zeros(B.ncols)
    check(L.bpf_backward_error(A.ctx.handle, A.handle, B.ncols, X.ptr, X.nrows, B.ptr, B.nrows,
                               out.ctypes.data_as(_lib.dp)))
    return out


def householder(A: HipMatrix):
    """(Q, R) = householder(A) (src/block_utils.jl:38-53): thin QR of a
    device n x p block by CholQR2; R upper triangular with positive diagonal.
    Returns new device matrices Q (n x p) and R (p x p)."""
    n, p = A.shape
    if A.dtype != np.float64:
        raise ValueError("householder needs a float64 block")
    Q = HipMatrix((n, p), np.float64, A.ctx)
    check(L.bpf_memcpy_d2d(A.ctx.handle, Q.ptr, A.ptr, A.nbytes))
    R = HipMatrix((p, p), np.float64, A.ctx)
    flag = C.c_int()
    check(L.bpf_householder(A.ctx.handle, n, p, Q.ptr, n, R.ptr, p, C.byref(flag)))
    Q.shifted = bool(flag.value)
    return Q, R


def block_hessenberg_lsq(H: np.ndarray, G: np.ndarray, ctx: Context | None = None):
    """Block-Hessenberg least squares as block GMRES runs it on the device
    (src/block_gmres.jl:318-375): H is (K+1)p x Kp, G is p x p.  Returns
    (Y, rnorm) with Y (Kp x p) minimising ||[G; 0] - H Y|| per column and
    rnorm[k] the per-column residual norms after block column k."""
    ctx = ctx or default_context()
    G = np.asfortranarray(G, dtype=np.float64)
    H = np.asfortranarray(H, dtype=np.float64)
    p = G.shape[0]
    K = H.shape[1] // p
    if G.shape != (p, p) or H.shape != ((K + 1) * p, K * p):
        raise ValueError("H must be (K+1)p x Kp and G p x p")
    Y = np.zeros((K * p, p), order="F")
    rn = np.zeros((K, p))
    check(L.bpf_block_hessenberg_lsq(ctx.handle, p, K, H.ctypes.data_as(_lib.dp), H.shape[0],
                                     G.ctypes.data_as(_lib.dp), p, Y.ctypes.data_as(_lib.dp), K * p,
                                     rn.ctypes.data_as(_lib.dp)))
    return Y, rn


from .cusolverrf import CUSOLVERRF  # noqa: E402
from .batched import BatchedSystems, block_gmres_batched  # noqa: E402
from .matrixmarket import HostCSR, mmread, mmread_dense, mmwrite  # noqa: E402
