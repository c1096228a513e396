"""CUSOLVERRF: fixed-pattern LU refactorization (src/cusolverRF.jl)."""
from __future__ import annotations


class CUSOLVERRF:
    class _RfBase:
        pass

    class CusolverRfLU(_RfBase):
        pass
