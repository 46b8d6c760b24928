"""Hilbert-space bases (reference bindings/python/qkrylov/basis.py)."""
from qkrylov_amd import Basis, SpinHalfBasis, FermionBasis, HubbardBasis, TJBasis  # noqa: F401
