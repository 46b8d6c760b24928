"""Local site physics (reference bindings/python/qkrylov/site.py)."""
from qkrylov_amd import Site, SpinHalfSite, FermionSite, HubbardSite, TJSite  # noqa: F401
