"""Batched solves of many independent systems against the FP64 direct oracle
and against solving each system alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _systems(gen, nsys, seed, nmin=60, nmax=3000, ps=None):
    nbus, p = gen.batch_sizes(nsys, seed, nmin=nmin, nmax=nmax)
    if ps is not None:
        p = np.array([ps[i % len(ps)] for i in range(nsys)])
    out = []
    for i in range(nsys):
        _, S = gen.make_system(int(nbus[i]), int(p[i]), seed=seed + 17 * i)
        out.append(S)
    return out


def _berr(oracle, S, X):
    return oracle.backward_error(S.indptr, S.indices, S.data, X, S.B)


@pytest.mark.parametrize("prec,deg", [("mixed", 0), ("fp64", 0), ("mixed", 4)])
def test_batch_backward_error(bpf, gen, oracle, prec, deg):
    systems = _systems(gen, 14, 71, ps=[4, 8, 16, 32])
    As = [bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data) for S in systems]
    Bs = [bpf.HipMatrix.from_numpy(S.B) for S in systems]
    print(f"batch: {len(systems)} systems, n = {[S.n for S in systems]}", flush=True)
    Xs, st, solved = bpf.block_gmres_batched(As, Bs, precision=prec, berr_tol=1e-10, memory=40,
                                             poly_degree=deg, itmax=600)
    assert solved.all(), (st, solved)
    assert st.solved and st.max_berr <= 1e-10
    for S, X in zip(systems, Xs):
        Xh = X.to_numpy()
        assert _berr(oracle, S, Xh).max() <= 1e-10
        Xd = oracle.direct_solve(S.indptr, S.indices, S.data, S.B)
        rel = np.abs(Xh - Xd).max(axis=0) / np.abs(Xd).max(axis=0)
        assert np.all(rel < 1e-5)


def test_batch_matches_single_system_solves(bpf, gen, oracle):
    # The reference stopping rule per system: ||Lambda||_F <= atol + rtol ||R0||_F.
    systems = _systems(gen, 6, 72, nmax=1500, ps=[8, 16])
    As = [bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data) for S in systems]
    Bs = [bpf.HipMatrix.from_numpy(S.B) for S in systems]
    Xs, st, solved = bpf.block_gmres_batched(As, Bs, precision="fp64", atol=0.0, rtol=1e-10, memory=60)
    assert solved.all()
    for S, A, B, X in zip(systems, As, Bs, Xs):
        P = bpf.ilu0(A)
        X1, st1 = bpf.block_gmres(A, B, N=P, ldiv=True, memory=60, atol=0.0, rtol=1e-10)
        assert st1.solved
        x1 = X1.to_numpy()
        np.testing.assert_allclose(X.to_numpy(), x1, rtol=0, atol=1e-7 * np.abs(x1).max())


def test_batch_uneven_with_restart_and_zero_rhs(bpf, gen, oracle):
    # Restarted block GMRES(30) with ILU(0) converges on all of these in the
    # CPU restatement (at most ~330 block steps); GMRES(15) stagnates on some.
    systems = _systems(gen, 9, 73, nmin=30, nmax=1500, ps=[8, 8, 16])
    systems[3].B[:] = 0.0
    As = [bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data) for S in systems]
    Bs = [bpf.HipMatrix.from_numpy(S.B) for S in systems]
    batch = bpf.BatchedSystems(As, [S.B.shape[1] for S in systems])
    try:
        for rep in range(2):   # the handle is reusable
            Xs, st, solved = batch.solve(Bs, precision="mixed", berr_tol=1e-10, memory=30, restart=True,
                                         itmax=2000)
            assert solved.all(), (rep, st, solved)
            assert np.all(Xs[3].to_numpy() == 0.0)   # zero RHS: X = 0 (0/0 backward error)
            for i, (S, X) in enumerate(zip(systems, Xs)):
                if i != 3:
                    assert _berr(oracle, S, X.to_numpy()).max() <= 1e-10
    finally:
        batch.destroy()


def test_batch_overlapped_hessenberg_update_is_exact(bpf, gen):
    # Groups stop at different iterations; with the Hessenberg updates on a
    # second stream (and one speculative iteration queued) the results are
    # bitwise those of the one-iteration-after-another order.
    systems = _systems(gen, 9, 83, ps=[8, 16, 32])
    As = [bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data) for S in systems]
    Bs = [bpf.HipMatrix.from_numpy(S.B) for S in systems]
    out = {}
    try:
        for ovl in (0, 1):
            bpf.set_option("overlap", ovl)
            Xs, st, solved = bpf.block_gmres_batched(As, Bs, precision="mixed", berr_tol=1e-10, memory=40,
                                                     itmax=600)
            out[ovl] = ([X.to_numpy() for X in Xs], st.niter, solved.copy())
    finally:
        bpf.set_option("overlap", 1)
    assert out[0][1] == out[1][1]
    np.testing.assert_array_equal(out[0][2], out[1][2])
    for a, b in zip(out[0][0], out[1][0]):
        np.testing.assert_array_equal(b, a)


def test_batch_bins_solved_concurrently_are_exact(bpf, gen, oracle):
    # Bins (systems grouped by RHS count) solved from several host threads at
    # once, each on its own streams, give bitwise the one-bin-at-a-time result.
    systems = _systems(gen, 12, 97, nmin=300, ps=[4, 8, 16, 32, 64])
    As = [bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data) for S in systems]
    Bs = [bpf.HipMatrix.from_numpy(S.B) for S in systems]
    out = {}
    try:
        for nthr in (1, 4, 2):
            bpf.set_option("batch_threads", nthr)
            assert bpf.get_option("batch_threads") == nthr
            Xs, st, solved = bpf.block_gmres_batched(As, Bs, precision="mixed", berr_tol=1e-10, memory=40,
                                                     itmax=600)
            out[nthr] = ([X.to_numpy() for X in Xs], st.niter, solved.copy())
    finally:
        bpf.set_option("batch_threads", 4)
    assert out[1][2].all()
    for nthr in (4, 2):
        assert out[nthr][1] == out[1][1]
        np.testing.assert_array_equal(out[nthr][2], out[1][2])
        for a, b in zip(out[1][0], out[nthr][0]):
            np.testing.assert_array_equal(b, a)
    for S, X in zip(systems, out[4][0]):
        be = oracle.backward_error(S.indptr, S.indices, S.data, X, S.B)
        assert be.max() <= 1e-10


def test_batch_handle_solves_again_identically(bpf, gen, oracle):
    # A BatchedSystems handle keeps one Krylov engine per RHS-count bin across
    # solves; bins of different widths share the context's buffers.  Solving
    # twice with the same handle gives bitwise the same answer, and a third
    # solve with other right-hand sides is still correct.
    systems = _systems(gen, 10, 91, ps=[4, 8, 16, 32, 64])
    As = [bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data) for S in systems]
    Bs = [bpf.HipMatrix.from_numpy(S.B) for S in systems]
    batch = bpf.BatchedSystems(As, [S.B.shape[1] for S in systems])
    try:
        runs = []
        for _ in range(2):
            Xs, st, solved = batch.solve(Bs, precision="mixed", berr_tol=1e-10, memory=40, itmax=600)
            assert solved.all()
            runs.append([X.to_numpy() for X in Xs])
        for a, b in zip(runs[0], runs[1]):
            np.testing.assert_array_equal(b, a)
        rng = np.random.default_rng(5)
        B2 = [rng.standard_normal(S.B.shape) for S in systems]
        Xs, st, solved = batch.solve([bpf.HipMatrix.from_numpy(b) for b in B2], precision="mixed",
                                     berr_tol=1e-10, memory=40, itmax=600)
        assert solved.all()
        for S, X, b in zip(systems, Xs, B2):
            assert oracle.backward_error(S.indptr, S.indices, S.data, X.to_numpy(), b).max() <= 1e-10
    finally:
        batch.destroy()


def test_batch_errors(bpf, gen):
    _, S = gen.make_system(30, 4, seed=74)
    A = bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data)
    with pytest.raises(bpf.BpfError):
        bpf.BatchedSystems([A], [200])
    batch = bpf.BatchedSystems([A], [4])
    with pytest.raises(ValueError):
        batch.solve([bpf.HipMatrix.from_numpy(np.zeros((S.n, 3)))])
    with pytest.raises(bpf.BpfError):
        batch.solve([bpf.HipMatrix.from_numpy(S.B)], precision="mixed32")



def test_batch_graph_replay_is_exact(bpf, gen, oracle):
    # Each bin's inner FP32 runs replay a captured HIP graph (bins on their
    # own threads and streams, captured concurrently): bitwise the result of
    # issuing every launch from the host.
    systems = _systems(gen, 12, 99, nmin=300, ps=[8, 16, 32, 64])
    As = [bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data) for S in systems]
    Bs = [bpf.HipMatrix.from_numpy(S.B) for S in systems]
    out = {}
    try:
        for graph in (0, 1):
            bpf.set_option("graph", graph)
            bpf.set_option("graph_replays", 0)
            Xs, st, solved = bpf.block_gmres_batched(As, Bs, precision="mixed", berr_tol=1e-10, memory=40,
                                                     itmax=600, inner_its=6, bmgs_blocks=16)
            out[graph] = ([X.to_numpy() for X in Xs], st.niter, solved.copy(), bpf.get_option("graph_replays"))
    finally:
        bpf.set_option("graph", 1)
    assert out[0][3] == 0 and out[1][3] > 0, (out[0][3], out[1][3])
    assert out[1][2].all()
    assert out[0][1] == out[1][1]
    np.testing.assert_array_equal(out[0][2], out[1][2])
    for a, b in zip(out[0][0], out[1][0]):
        np.testing.assert_array_equal(b, a)
    for S, X in zip(systems, out[1][0]):
        assert oracle.backward_error(S.indptr, S.indices, S.data, X, S.B).max() <= 1e-10


@pytest.mark.parametrize("width", [8, 16])
def test_batch_inner_groups_split(bpf, gen, oracle, width):
    # Each system's columns split into inner Krylov blocks of `width` columns
    # (independent FP32 block GMRES runs over the system's rows): still the
    # backward-error target and the direct solution for every system.
    systems = _systems(gen, 8, 101, nmin=300, ps=[8, 16, 32, 64])
    As = [bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data) for S in systems]
    Bs = [bpf.HipMatrix.from_numpy(S.B) for S in systems]
    batch = bpf.BatchedSystems(As, [S.B.shape[1] for S in systems])
    try:
        Xs, st, solved = batch.solve(Bs, precision="mixed", berr_tol=1e-10, memory=40, itmax=600, inner_its=6,
                                     inner_width=width, bmgs_blocks=16)
        assert solved.all(), (st, solved)
        np.testing.assert_array_equal(batch.inner_widths(),
                                      [min(width, S.B.shape[1]) for S in systems])
        for S, X in zip(systems, Xs):
            Xh = X.to_numpy()
            assert _berr(oracle, S, Xh).max() <= 1e-10
            Xd = oracle.direct_solve(S.indptr, S.indices, S.data, S.B)
            assert np.all(np.abs(Xh - Xd).max(axis=0) / np.abs(Xd).max(axis=0) < 1e-5)
    finally:
        batch.destroy()


def test_batch_inner_width_autotuned(bpf, gen, oracle):
    # inner_width="auto": each bin's first solve times the candidate widths
    # and keeps the fastest; the next solve uses it and equals, bit for bit,
    # a solve with that width given explicitly.
    systems = _systems(gen, 8, 103, nmin=300, ps=[8, 16, 32, 64])
    As = [bpf.HipSparseMatrixCSR(S.indptr, S.indices, S.data) for S in systems]
    Bs = [bpf.HipMatrix.from_numpy(S.B) for S in systems]
    kw = dict(precision="mixed", berr_tol=1e-10, memory=40, itmax=600, inner_its=6, bmgs_blocks=16)
    batch = bpf.BatchedSystems(As, [S.B.shape[1] for S in systems])
    try:
        Xs, st, solved = batch.solve(Bs, inner_width="auto", **kw)
        assert solved.all(), (st, solved)
        w = batch.inner_widths()
        for S, wi, X in zip(systems, w, Xs):
            p = S.B.shape[1]
            assert wi in [c for c in (8, 16, 32, 64) if c <= p], (p, wi)
            assert oracle.backward_error(S.indptr, S.indices, S.data, X.to_numpy(), S.B).max() <= 1e-10
        Xa, _, solved_a = batch.solve(Bs, inner_width="auto", **kw)
        np.testing.assert_array_equal(batch.inner_widths(), w)
        # bins of one RHS count share a width: solving with those widths
        # given explicitly, one count at a time, reproduces the tuned solve
        for p in sorted({S.B.shape[1] for S in systems}):
            idx = [i for i, S in enumerate(systems) if S.B.shape[1] == p]
            sub = bpf.BatchedSystems([As[i] for i in idx], [p] * len(idx))
            try:
                Xe, _, _ = sub.solve([Bs[i] for i in idx], inner_width=int(w[idx[0]]), **kw)
                for j, i in enumerate(idx):
                    np.testing.assert_array_equal(Xe[j].to_numpy(), Xa[i].to_numpy())
            finally:
                sub.destroy()
    finally:
        batch.destroy()
