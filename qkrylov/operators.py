"""Operator names and the OpSum expression algebra (reference
bindings/python/qkrylov/operators.py)."""
from qkrylov_amd import (Op, OpSum, Sz, Sp, Sm, Sx, Sy, CdagUp, CUp, CdagDn, CDn, Nup, Ndn,  # noqa: F401
                         Nupdn, Bdag, B, N)
from qkrylov_amd import LocalOpExpr, TermExpr, OpSumExpr  # noqa: F401
