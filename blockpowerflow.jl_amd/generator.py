"""Synthetic polar Newton-Raphson power-flow Jacobians (the workload of the multi-RHS path).

The reference solves ``Gx * X = Gu`` where ``Gx`` is the power-flow Jacobian with
respect to the state, read from ``benchmarks/dump/Gx_<case>.mtx``
(benchmarks/benchmarks_nvidia.jl:29-35).  Those dumps cannot be shipped, so the
workloads are generated here from a seeded grid model:

* buses are points in the unit square; a k-nearest-neighbour graph (k = 6) is
  built and a random spanning tree of it is drawn (Kruskal on random weights);
  extra non-tree kNN edges are added until the mean degree is 2.7.  Drawing the
  extra edges among geographic neighbours keeps the grid near-planar, like real
  transmission grids (small separators, moderate LU fill);
* bus labels are a random permutation (real case files are not numbered
  geographically either);
* line r ~ U(0.005, 0.05), x ~ U(0.05, 0.3) p.u. (series impedance only);
* one slack bus, 20 % PV buses, the rest PQ;
* operating point |V| ~ U(0.95, 1.05), angle ~ N(0, 0.1) rad.

The Jacobian is the standard polar one: unknowns ``[theta(non-slack); |V|(PQ)]``
(in bus order), equations ``[P(non-slack); Q(PQ)]``.  CSR, int32 indices,
float64 values, column indices sorted within each row.

Random numbers come from ``numpy.random.default_rng(seed)`` (PCG64).
"""
from __future__ import annotations

import dataclasses

import numpy as np
import scipy.sparse as sp
from scipy.spatial import cKDTree


@dataclasses.dataclass
class Grid:
    nbus: int
    f: np.ndarray          # from-bus of each line
    t: np.ndarray          # to-bus of each line
    r: np.ndarray
    x: np.ndarray
    slack: int
    pv: np.ndarray         # sorted bus indices
    pq: np.ndarray         # sorted bus indices
    vm: np.ndarray
    va: np.ndarray

    @property
    def ybus(self) -> sp.csr_matrix:
        y = 1.0 / (self.r + 1j * self.x)
        n = self.nbus
        rows = np.concatenate([self.f, self.t, self.f, self.t])
        cols = np.concatenate([self.t, self.f, self.f, self.t])
        vals = np.concatenate([-y, -y, y, y])
        return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def _find(parent, i):
    root = i
    while parent[root] != root:
        root = parent[root]
    while parent[i] != root:
        parent[i], i = root, parent[i]
    return root


def make_grid(nbus: int, seed: int, mean_degree: float = 2.7, knn: int = 6) -> Grid:
    """Random near-planar grid with ``nbus`` buses (deterministic in ``seed``)."""
    if nbus < 3:
        raise ValueError("nbus must be >= 3")
    rng = np.random.default_rng(seed)
    pts = rng.random((nbus, 2))
    k = min(knn, nbus - 1)
    _, nbr = cKDTree(pts).query(pts, k=k + 1)
    a = np.repeat(np.arange(nbus), k)
    b = nbr[:, 1:].reshape(-1)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    cand = np.unique(lo.astype(np.int64) * nbus + hi)
    cu, cv = cand // nbus, cand % nbus
    order = rng.permutation(len(cand))
    parent = list(range(nbus))
    in_tree = np.zeros(len(cand), dtype=bool)
    ntree = 0
    for e in order:
        ru, rv = _find(parent, int(cu[e])), _find(parent, int(cv[e]))
        if ru != rv:
            parent[ru] = rv
            in_tree[e] = True
            ntree += 1
    # The kNN graph can be disconnected (rare); join components by nearest roots.
    if ntree < nbus - 1:
        extra_u, extra_v = [], []
        roots = {}
        for i in range(nbus):
            roots.setdefault(_find(parent, i), i)
        reps = list(roots.values())
        for j in range(1, len(reps)):
            extra_u.append(reps[j - 1])
            extra_v.append(reps[j])
            parent[_find(parent, reps[j - 1])] = _find(parent, reps[j])
        cu = np.concatenate([cu, extra_u])
        cv = np.concatenate([cv, extra_v])
        in_tree = np.concatenate([in_tree, np.ones(len(extra_u), dtype=bool)])
        order = np.concatenate([order, np.arange(len(order), len(cu))])
    nline = int(round(mean_degree * nbus / 2.0))
    tree_idx = np.nonzero(in_tree)[0]
    rest = np.array([e for e in order if not in_tree[e]], dtype=np.int64)
    nextra = max(0, min(len(rest), nline - len(tree_idx)))
    lines = np.concatenate([tree_idx, rest[:nextra]])
    # Random bus labels.
    label = rng.permutation(nbus)
    f = label[cu[lines]].astype(np.int64)
    t = label[cv[lines]].astype(np.int64)
    nl = len(lines)
    r = rng.uniform(0.005, 0.05, nl)
    x = rng.uniform(0.05, 0.3, nl)
    buses = rng.permutation(nbus)
    slack = int(buses[0])
    npv = int(round(0.2 * nbus))
    pv = np.sort(buses[1:1 + npv])
    pq = np.sort(buses[1 + npv:])
    vm = rng.uniform(0.95, 1.05, nbus)
    va = rng.normal(0.0, 0.1, nbus)
    return Grid(nbus, f, t, r, x, slack, pv, pq, vm, va)


def jacobian(grid: Grid, vm: np.ndarray | None = None, va: np.ndarray | None = None):
    """Polar power-flow Jacobian d[P(ns); Q(pq)] / d[theta(ns); |V|(pq)] as CSR.

    Returns ``(indptr int32, indices int32, data float64, n)`` with sorted
    column indices.  ``vm``/``va`` override the grid's operating point (the
    refactorization workload perturbs them between Newton steps).
    """
    vm = grid.vm if vm is None else vm
    va = grid.va if va is None else va
    Y = grid.ybus
    V = vm * np.exp(1j * va)
    I = Y @ V
    diagV = sp.diags(V)
    diagI = sp.diags(I)
    Vnorm = sp.diags(V / np.abs(V))
    dS_dVa = 1j * diagV @ np.conj(diagI - Y @ diagV)
    dS_dVm = diagV @ np.conj(Y @ Vnorm) + np.conj(diagI) @ Vnorm
    ns = np.setdiff1d(np.arange(grid.nbus), [grid.slack])
    pq = grid.pq
    J11 = dS_dVa[ns][:, ns].real
    J12 = dS_dVm[ns][:, pq].real
    J21 = dS_dVa[pq][:, ns].imag
    J22 = dS_dVm[pq][:, pq].imag
    J = sp.bmat([[J11, J12], [J21, J22]], format="csr")
    J.sum_duplicates()
    J.sort_indices()
    n = J.shape[0]
    return J.indptr.astype(np.int32), J.indices.astype(np.int32), J.data.astype(np.float64), n


def state_index(grid: Grid):
    """Row index of the P equation of each non-slack bus (or -1 for the slack)."""
    ns = np.setdiff1d(np.arange(grid.nbus), [grid.slack])
    pos = -np.ones(grid.nbus, dtype=np.int64)
    pos[ns] = np.arange(len(ns))
    return pos


@dataclasses.dataclass
class System:
    """One CSR system ``A X = B`` (B is n x p, column-major / Fortran order)."""
    indptr: np.ndarray
    indices: np.ndarray
    data: np.ndarray
    n: int
    B: np.ndarray
    seed: int

    @property
    def nnz(self) -> int:
        return int(self.indptr[-1])

    def scipy(self) -> sp.csr_matrix:
        return sp.csr_matrix((self.data, self.indices, self.indptr), shape=(self.n, self.n))


def rhs_normal(n: int, p: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.standard_normal((n, p)))


def rhs_injections(grid: Grid, n: int, nnormal: int, nunit: int, seed: int) -> np.ndarray:
    """``nnormal`` N(0,1) columns followed by ``nunit`` unit injections e_i at the
    P row of random non-slack buses."""
    rng = np.random.default_rng(seed)
    B = np.zeros((n, nnormal + nunit), order="F")
    B[:, :nnormal] = rng.standard_normal((n, nnormal))
    pos = state_index(grid)
    ns = np.setdiff1d(np.arange(grid.nbus), [grid.slack])
    buses = rng.choice(ns, size=nunit, replace=nunit > len(ns))
    for c, b in enumerate(buses):
        B[pos[b], nnormal + c] = 1.0
    return B


def make_system(nbus: int, p: int, seed: int, rhs_seed: int | None = None,
                rhs: str = "normal", nunit: int = 0) -> tuple[Grid, System]:
    grid = make_grid(nbus, seed)
    indptr, indices, data, n = jacobian(grid)
    rs = seed + 1000 if rhs_seed is None else rhs_seed
    if rhs == "normal":
        B = rhs_normal(n, p, rs)
    elif rhs == "injection":
        B = rhs_injections(grid, n, p - nunit, nunit, rs)
    else:
        raise ValueError(rhs)
    return grid, System(indptr, indices, data, n, B, seed)


def newton_value_sets(grid: Grid, nsets: int, seed: int, sigma: float = 0.01):
    """Successive Newton value sets: |V| and angle perturbed by N(0, sigma) per step.
    Yields CSR value arrays on the (fixed) pattern of the base Jacobian."""
    rng = np.random.default_rng(seed)
    vm = grid.vm.copy()
    va = grid.va.copy()
    base_ptr, base_idx, _, _ = jacobian(grid)
    out = []
    for _ in range(nsets):
        vm = vm + rng.normal(0.0, sigma, grid.nbus)
        va = va + rng.normal(0.0, sigma, grid.nbus)
        ptr, idx, val, _ = jacobian(grid, vm, va)
        if not (np.array_equal(ptr, base_ptr) and np.array_equal(idx, base_idx)):
            val = _project_values(base_ptr, base_idx, ptr, idx, val)
        out.append(val)
    return out


def _project_values(bptr, bidx, ptr, idx, val):
    """Place values of (ptr, idx) on the base pattern (entries absent are zero)."""
    n = len(bptr) - 1
    A = sp.csr_matrix((val, idx, ptr), shape=(n, n))
    B = sp.csr_matrix((np.ones(len(bidx)), bidx, bptr), shape=(n, n))
    M = A.multiply(B) + B * 0.0
    M = (A + B) - B   # keeps the union pattern
    M = sp.csr_matrix(M)
    M.sort_indices()
    out = np.zeros(len(bidx))
    for i in range(n):
        cols = bidx[bptr[i]:bptr[i + 1]]
        row = M.getrow(i)
        d = dict(zip(row.indices, row.data))
        out[bptr[i]:bptr[i + 1]] = [d.get(c, 0.0) for c in cols]
    return out


def batch_sizes(nsys: int, seed: int, nmin: int = 500, nmax: int = 20000):
    """Unknown counts log-uniform in [nmin, nmax] and RHS counts in {8,16,32,64}."""
    rng = np.random.default_rng(seed)
    n_target = np.exp(rng.uniform(np.log(nmin), np.log(nmax), nsys))
    p = rng.choice([8, 16, 32, 64], size=nsys)
    # n ~= 1.8 * nbus for this grid model (PQ buses carry two unknowns)
    nbus = np.maximum(4, np.round(n_target / 1.8)).astype(int)
    return nbus, p.astype(int)
