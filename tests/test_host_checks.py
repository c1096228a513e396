"""CPU checks of the host mirror's argument handling: element types and
shapes are rejected before any native call (a mismatched block would make
the library read past a device allocation)."""
import numpy as np
import pytest


class _Blk:
    def __init__(self, shape, dtype):
        self.shape = shape
        self.dtype = np.dtype(dtype)
        self.nrows = shape[0]
        self.ncols = shape[1]


class _Csr:
    def __init__(self, n):
        self.n = n
        self.shape = (n, n)
        self.ctx = None


class _NoNative:
    def __getattr__(self, name):
        raise AssertionError(f"native call {name} made before the argument checks")


def test_check_system_element_types(bpf):
    A = _Csr(10)
    assert bpf._check_system(A, _Blk((10, 3), np.float64)) == np.float64
    assert bpf._check_system(A, _Blk((10, 3), np.float32)) == np.float32
    for bad in (np.float16, np.int32, np.complex128):
        with pytest.raises(TypeError):
            bpf._check_system(A, _Blk((10, 3), bad))
    with pytest.raises(TypeError):
        bpf._check_system(A, _Blk((10, 3), np.float32), _Blk((10, 3), np.float64))
    with pytest.raises(ValueError):
        bpf._check_system(A, _Blk((10, 3), np.float64), _Blk((10, 2), np.float64))
    with pytest.raises(ValueError):
        bpf._check_system(A, _Blk((9, 3), np.float64))


def test_default_tolerances_follow_element_type(bpf):
    assert bpf._default_tol(np.float64) == pytest.approx(np.sqrt(np.finfo(np.float64).eps))
    assert bpf._default_tol(np.float32) == pytest.approx(np.sqrt(np.finfo(np.float32).eps))


@pytest.mark.parametrize("fn", ["block_gmres", "block_bicgstab"])
def test_bad_element_type_rejected_before_native_calls(bpf, monkeypatch, fn):
    monkeypatch.setattr(bpf, "L", _NoNative())
    with pytest.raises(TypeError):
        getattr(bpf, fn)(_Csr(10), _Blk((10, 2), np.float16))


def test_batched_rejects_non_float64(bpf, monkeypatch):
    monkeypatch.setattr(bpf, "L", _NoNative())
    with pytest.raises((TypeError, ValueError)):
        bpf.block_gmres_batched([_Csr(10)], [_Blk((10, 2), np.float32)])
