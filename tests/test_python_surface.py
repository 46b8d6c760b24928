"""The Python mirror of the reference package (bindings/python/qkrylov):
names, constructor keywords and host-only behaviour, checked the way the
reference's own tests (bindings/python/tests/test_basic.py:4-49) check them.
GPU-free parts only; the Hamiltonian side is in test_gpu_python_surface.py."""
import numpy as np
import pytest


def test_reference_names_exist(qk):
    for name in ("Op", "OpSum", "Sz", "Sp", "Sm", "Sx", "Sy", "CdagUp", "CUp", "CdagDn", "CDn",
                 "Nup", "Ndn", "Nupdn", "Bdag", "B", "N", "Site", "SpinHalfSite", "FermionSite",
                 "HubbardSite", "TJSite", "Basis", "SpinHalfBasis", "FermionBasis", "HubbardBasis",
                 "TJBasis", "MatrixFreeHamiltonian", "LanczosResult", "lanczos_ground_state",
                 "DavidsonResult", "davidson_lowest", "DynamicsResult", "continued_fraction_coeffs",
                 "evaluate_spectral_function", "FTLMResult", "ftlm", "find_gpu", "gpu_count"):
        assert hasattr(qk, name), name


def test_bases_as_reference_tests(qk):
    # test_basic.py:4-12, 27-29
    b = qk.SpinHalfBasis(4)
    assert b.size == 16 and b.nsites == 4
    assert qk.FermionBasis(4, conserve_n=True, n=2).size == 6
    assert qk.TJBasis(2).size == 9
    # dtype keyword of the reference constructors is accepted
    assert qk.SpinHalfBasis(4, conserve_sz=True, sz=0, dtype=np.float32).size == 6
    assert qk.HubbardBasis(2, conserve_nup=True, nup=1, conserve_ndn=True, ndn=1, dtype=np.float64).size == 4


def test_opsum_tuple_validation(qk):
    # test_basic.py:31-44: a malformed tuple is a ValueError
    os_ = qk.OpSum(dtype=np.float32)
    os_ += (1.0, "Sz", 0, "Sz", 1)
    with pytest.raises(ValueError):
        os_ += (1.0, "Sz")
    os_ += 0.5 * (qk.Bdag(0) * qk.B(1))
    assert os_.size == 2
    os_.add_term(0.25, qk.Op.Sz, 0, qk.Op.Sz, 1)
    assert os_.terms[-1][1] == [("Sz", 0), ("Sz", 1)]


def test_no_gpu_here(qk):
    # this container has no GPU: the product path refuses instead of falling back
    assert qk.gpu_count() == 0 and qk.find_gpu() is None
    with pytest.raises(ValueError):
        qk.MatrixFreeHamiltonian(qk.SpinHalfBasis(2), qk.SpinHalfSite(), qk.OpSum(), device="cpu")
