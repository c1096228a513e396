"""Matrix Market reader / writer (the reference benchmarks' on-disk format,
benchmarks/benchmarks_nvidia.jl:29-35).  Pinned against scipy.io.mmread, an
independent reader, on every case file the reference ships, and against
hand-written files for the semantics MatrixMarket.mmread + sparse() fix:
duplicates summed, symmetric parts mirrored, pattern entries 1."""
import glob
import importlib.util
import os
import sys

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

from conftest import PKG_DIR, ROOT

MTX = os.path.join(ROOT, "tests", "golden", "mtx")


def _mm():
    name = "bpf_matrixmarket"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG_DIR, "matrixmarket.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


mm = _mm()
CASE_FILES = sorted(glob.glob(os.path.join(MTX, "*.mtx.gz")))


def test_fixtures_present():
    # nine cases, a Jacobian and a right-hand-side file each
    assert len(CASE_FILES) == 18
    for f in CASE_FILES:
        if os.path.basename(f).startswith("Gx_"):
            assert os.path.exists(f.replace("Gx_", "Gu_"))


@pytest.mark.parametrize("path", CASE_FILES, ids=[os.path.basename(f) for f in CASE_FILES])
def test_reader_matches_scipy(path):
    A = mm.mmread(path)
    R = sp.csr_matrix(scipy.io.mmread(path))
    R.sum_duplicates()
    R.sort_indices()
    assert A.shape == R.shape
    assert A.indptr.dtype == np.int32 and A.indices.dtype == np.int32
    np.testing.assert_array_equal(A.indptr, R.indptr)
    np.testing.assert_array_equal(A.indices, R.indices)
    np.testing.assert_array_equal(A.data, R.data)   # same decimal strings, same doubles


def test_case9_by_hand():
    # dump/Gx_case9.mtx: 14 x 14 with 82 stored entries, first entry (1, 1)
    A = mm.mmread(os.path.join(MTX, "Gx_case9.mtx.gz"))
    assert A.shape == (14, 14) and A.nnz == 82
    assert A.indices[0] == 0 and A.data[0] == 16.743463396815724
    B = mm.mmread_dense(os.path.join(MTX, "Gu_case9.mtx.gz"))
    assert B.shape == (14, 5) and B.flags.f_contiguous
    assert np.count_nonzero(B) == 10
    assert B[2, 0] == -0.6888559757161761


def test_semantics(tmp_path):
    f = tmp_path / "dup.mtx"
    f.write_text("%%MatrixMarket matrix coordinate real general\n% comment\n\n3 4 4\n"
                 "1 2 1.5\n3 4 2\n1 2 0.25\n2 1 0\n")
    A = mm.mmread(f)
    np.testing.assert_array_equal(A.to_dense(), [[0, 1.75, 0, 0], [0, 0, 0, 0], [0, 0, 0, 2]])
    assert A.nnz == 3                     # the explicit zero stays a stored entry
    f = tmp_path / "sym.mtx"
    f.write_text("%%MatrixMarket matrix coordinate real symmetric\n3 3 3\n1 1 4\n3 1 -1\n2 2 5\n")
    np.testing.assert_array_equal(mm.mmread_dense(f), [[4, 0, -1], [0, 5, 0], [-1, 0, 0]])
    f = tmp_path / "skew.mtx"
    f.write_text("%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n2 1 3\n")
    np.testing.assert_array_equal(mm.mmread_dense(f), [[0, -3], [3, 0]])
    f = tmp_path / "pat.mtx"
    f.write_text("%%MatrixMarket matrix coordinate pattern general\n2 3 2\n1 3\n2 1\n")
    np.testing.assert_array_equal(mm.mmread_dense(f), [[0, 0, 1], [1, 0, 0]])
    f = tmp_path / "arr.mtx"
    f.write_text("%%MatrixMarket matrix array real general\n2 3\n1\n2\n3\n4\n5\n6\n")
    np.testing.assert_array_equal(mm.mmread_dense(f), [[1, 3, 5], [2, 4, 6]])
    f = tmp_path / "arrsym.mtx"
    f.write_text("%%MatrixMarket matrix array real symmetric\n2 2\n1\n2\n3\n")
    np.testing.assert_array_equal(mm.mmread_dense(f), [[1, 2], [2, 3]])


def test_errors(tmp_path):
    cases = {
        "hdr.mtx": "%%NotMatrixMarket matrix coordinate real general\n1 1 0\n",
        "cplx.mtx": "%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n",
        "short.mtx": "%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1\n",
        "range.mtx": "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n",
        "nosize.mtx": "%%MatrixMarket matrix coordinate real general\n% only a comment\n",
    }
    for name, text in cases.items():
        f = tmp_path / name
        f.write_text(text)
        with pytest.raises(ValueError):
            mm.mmread(f)


def test_write_read_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    D = np.where(rng.random((7, 5)) < 0.4, rng.standard_normal((7, 5)), 0.0)
    for name in ("d.mtx", "d.mtx.gz"):
        mm.mmwrite(tmp_path / name, D, comment="round trip")
        np.testing.assert_array_equal(mm.mmread_dense(tmp_path / name), D)
    A = mm.mmread(os.path.join(MTX, "Gx_case57.mtx.gz"))
    mm.mmwrite(tmp_path / "g.mtx", A)
    A2 = mm.mmread(tmp_path / "g.mtx")
    np.testing.assert_array_equal(A2.indptr, A.indptr)
    np.testing.assert_array_equal(A2.indices, A.indices)
    np.testing.assert_array_equal(A2.data, A.data)


def test_empty_and_crlf(tmp_path):
    f = tmp_path / "empty.mtx"
    f.write_text("%%MatrixMarket matrix coordinate real general\n3 2 0\n")
    A = mm.mmread(f)
    assert A.shape == (3, 2) and A.nnz == 0
    np.testing.assert_array_equal(A.indptr, [0, 0, 0, 0])
    np.testing.assert_array_equal(mm.mmread_dense(f), np.zeros((3, 2)))
    f = tmp_path / "crlf.mtx"
    f.write_bytes(b"%%MatrixMarket matrix coordinate real general\r\n% c\r\n2 2 2\r\n1 1 3.5\r\n2 2 -1e-3\r\n")
    np.testing.assert_array_equal(mm.mmread_dense(f), [[3.5, 0], [0, -1e-3]])
