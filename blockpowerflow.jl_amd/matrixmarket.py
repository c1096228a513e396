"""Matrix Market files on the host: the on-disk format of the reference's
benchmark inputs (benchmarks/benchmarks_nvidia.jl:29-35 and
benchmarks/benchmarks_amd.jl:32-38 read dump/Gx_<case>.mtx, the power-flow
Jacobian, and dump/Gu_<case>.mtx, its right-hand sides, with
MatrixMarket.mmread and densify the latter with Matrix(B)).

``mmread`` returns a host CSR (int32 row pointers and column indices, the
layout ``HipSparseMatrixCSR`` takes) with the semantics of
MatrixMarket.mmread + SparseArrays.sparse: 1-based coordinates, duplicate
entries summed, ``symmetric`` / ``skew-symmetric`` / ``hermitian`` files
mirrored, ``pattern`` entries set to 1.  ``mmread_dense`` returns the
column-major dense matrix (Matrix(mmread(path))).  Files ending in ``.gz``
are read through gzip.  ``mmwrite`` writes ``coordinate real general``.
"""
from __future__ import annotations

import gzip
from dataclasses import dataclass

import numpy as np


@dataclass
class HostCSR:
    """A host CSR matrix: indptr (m+1), indices (nnz, sorted per row), data."""
    shape: tuple
    indptr: np.ndarray
    indices: np.ndarray
    data: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.indptr[-1])

    def to_dense(self) -> np.ndarray:
        m, n = self.shape
        out = np.zeros((m, n), order="F")
        rows = np.repeat(np.arange(m), np.diff(self.indptr))
        out[rows, self.indices] = self.data
        return out


def _open(path):
    path = str(path)
    if path.endswith(".gz"):
        return gzip.open(path, "rt")
    return open(path, "r")


def _header(f):
    first = f.readline()
    tok = first.split()
    if len(tok) != 5 or tok[0].lower() != "%%matrixmarket" or tok[1].lower() != "matrix":
        raise ValueError(f"not a Matrix Market matrix header: {first.strip()!r}")
    fmt, field, sym = (t.lower() for t in tok[2:])
    if fmt not in ("coordinate", "array"):
        raise ValueError(f"unknown Matrix Market format {fmt!r}")
    if field not in ("real", "integer", "pattern", "double"):
        raise ValueError(f"unsupported Matrix Market field {field!r} (real, integer, pattern)")
    if sym not in ("general", "symmetric", "skew-symmetric", "hermitian"):
        raise ValueError(f"unknown Matrix Market symmetry {sym!r}")
    if fmt == "array" and field == "pattern":
        raise ValueError("array format cannot hold a pattern matrix")
    line = f.readline()
    while line and (line.startswith("%") or not line.strip()):
        line = f.readline()
    if not line:
        raise ValueError("missing size line")
    size = [int(t) for t in line.split()]
    return fmt, field, sym, size, f.read()


def _entries(path):
    """(shape, rows, cols, vals) of the stored entries, 0-based, with the
    symmetric part mirrored."""
    with _open(path) as f:
        fmt, field, sym, size, body = _header(f)
    tok = body.split()
    if fmt == "coordinate":
        if len(size) != 3:
            raise ValueError("coordinate size line must hold rows, columns, entries")
        m, n, nnz = size
        per = 2 if field == "pattern" else 3
        if len(tok) != per * nnz:
            raise ValueError(f"expected {nnz} entries of {per} fields, found {len(tok)} fields")
        a = np.array(tok, dtype=np.float64).reshape(nnz, per) if nnz else np.zeros((0, per))
        r = a[:, 0].astype(np.int64) - 1
        c = a[:, 1].astype(np.int64) - 1
        v = np.ones(nnz) if field == "pattern" else a[:, 2].copy()
        if nnz and (r.min() < 0 or c.min() < 0 or r.max() >= m or c.max() >= n):
            raise ValueError("entry index outside the matrix")
    else:
        if len(size) != 2:
            raise ValueError("array size line must hold rows and columns")
        m, n = size
        vals = np.array(tok, dtype=np.float64)
        if sym == "general":
            if vals.size != m * n:
                raise ValueError(f"expected {m * n} values, found {vals.size}")
            c, r = np.divmod(np.arange(m * n), m)      # column-major
            v = vals
        else:
            # lower triangle by columns (diagonal omitted when skew-symmetric)
            first = 1 if sym == "skew-symmetric" else 0
            cols, rows = [], []
            for j in range(n):
                i = np.arange(j + first, m)
                rows.append(i)
                cols.append(np.full(i.size, j))
            r = np.concatenate(rows) if rows else np.zeros(0, np.int64)
            c = np.concatenate(cols) if cols else np.zeros(0, np.int64)
            if vals.size != r.size:
                raise ValueError(f"expected {r.size} values, found {vals.size}")
            v = vals
    if sym != "general":
        off = r != c
        sign = -1.0 if sym == "skew-symmetric" else 1.0
        r, c, v = (np.concatenate([r, c[off]]), np.concatenate([c, r[off]]),
                   np.concatenate([v, sign * v[off]]))
    return (m, n), r, c, v


def mmread(path) -> HostCSR:
    """MatrixMarket.mmread(path) as a host CSR (duplicates summed, explicit
    zeros kept as stored entries, as sparse(I, J, V) keeps them)."""
    (m, n), r, c, v = _entries(path)
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    if r.size:
        new = np.ones(r.size, dtype=bool)
        new[1:] = (r[1:] != r[:-1]) | (c[1:] != c[:-1])
        start = np.flatnonzero(new)
        v = np.add.reduceat(v, start)
        r, c = r[start], c[start]
    if max(m, n, r.size) >= 2 ** 31:
        raise ValueError("matrix too large for 32-bit CSR indices")
    indptr = np.zeros(m + 1, dtype=np.int32)
    np.cumsum(np.bincount(r, minlength=m), out=indptr[1:])
    return HostCSR((m, n), indptr, c.astype(np.int32), v.astype(np.float64))


def mmread_dense(path) -> np.ndarray:
    """Matrix(mmread(path)): the dense column-major matrix."""
    return mmread(path).to_dense()


def mmwrite(path, A, comment: str = "") -> None:
    """Writes a HostCSR (or a dense 2-D array) as `coordinate real general`
    with 17 significant digits, so mmread gives back the same values."""
    if not isinstance(A, HostCSR):
        D = np.asarray(A, dtype=np.float64)
        if D.ndim != 2:
            raise ValueError("mmwrite takes a HostCSR or a 2-D array")
        r, c = np.nonzero(D)
        order = np.lexsort((r, c))
        r, c = r[order], c[order]
        v = D[r, c]
        shape = D.shape
    else:
        shape = A.shape
        r = np.repeat(np.arange(shape[0]), np.diff(A.indptr))
        c, v = A.indices, A.data
    path = str(path)
    op = gzip.open(path, "wt") if path.endswith(".gz") else open(path, "w")
    with op as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        for line in comment.splitlines():
            f.write(f"%{line}\n")
        f.write(f"{shape[0]} {shape[1]} {len(v)}\n")
        for i, j, x in zip(r, c, v):
            f.write(f"{i + 1} {j + 1} {float(x):.17g}\n")
