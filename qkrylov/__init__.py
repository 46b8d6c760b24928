"""qkrylov — drop-in import name for this library.

Scripts written for the reference package (``import qkrylov``;
bindings/python/qkrylov/__init__.py) run unchanged: the same names, module
layout (``qkrylov.operators``, ``.site``, ``.basis``, ``.hamiltonian``,
``.solvers``, ``._qkrylov_cpp``) and default arguments
(``MatrixFreeHamiltonian(..., device="cpu", dtype=np.float32)``).  Every
numerical call runs the HIP kernels of ``libqkrylov.so`` through
``qkrylov_amd``; the ``device`` argument is accepted for its signature and
every value selects the GPU.
"""
from .operators import (Op, OpSum, Sz, Sp, Sm, Sx, Sy, CdagUp, CUp, CdagDn, CDn, Nup, Ndn,
                        Nupdn, Bdag, B, N)
from .site import Site, SpinHalfSite, FermionSite, HubbardSite, TJSite
from .basis import Basis, SpinHalfBasis, FermionBasis, HubbardBasis, TJBasis
from .hamiltonian import MatrixFreeHamiltonian
from .solvers import (LanczosResult, lanczos_ground_state, DavidsonResult, davidson_lowest,
                      DynamicsResult, continued_fraction_coeffs, evaluate_spectral_function,
                      FTLMResult, ftlm)
from ._qkrylov_cpp import Device_FP32 as Device


def find_gpu():
    """'hip' when a GPU is visible, else None (reference __init__.py:30-34)."""
    if Device.is_gpu_build() and Device.gpu_count() > 0:
        return Device.backend_name()
    return None


def gpu_count():
    """Number of visible GPUs (reference __init__.py:36-38)."""
    return Device.gpu_count()


__version__ = "0.0.0"

__all__ = [
    "Op", "OpSum", "Sz", "Sp", "Sm", "Sx", "Sy", "CdagUp", "CUp", "CdagDn", "CDn",
    "Nup", "Ndn", "Nupdn", "Bdag", "B", "N",
    "Site", "SpinHalfSite", "FermionSite", "HubbardSite", "TJSite",
    "Basis", "SpinHalfBasis", "FermionBasis", "HubbardBasis", "TJBasis",
    "MatrixFreeHamiltonian",
    "LanczosResult", "lanczos_ground_state", "DavidsonResult", "davidson_lowest",
    "DynamicsResult", "continued_fraction_coeffs", "evaluate_spectral_function",
    "FTLMResult", "ftlm", "find_gpu", "gpu_count",
]
