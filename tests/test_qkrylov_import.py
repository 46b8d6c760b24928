"""`import qkrylov` resolves to this library with the reference package's
names, module layout and defaults (bindings/python/qkrylov/__init__.py,
hamiltonian.py:25, src/binding_impl.hpp:72-277).  GPU-free parts."""
import inspect

import numpy as np
import pytest


def test_names_and_modules():
    import qkrylov
    import qkrylov._qkrylov_cpp as cpp
    from qkrylov import basis, hamiltonian, operators, site, solvers  # noqa: F401
    for name in qkrylov.__all__:
        assert hasattr(qkrylov, name), name
    sig = inspect.signature(qkrylov.MatrixFreeHamiltonian.__init__)
    assert sig.parameters["device"].default == "cpu"
    assert sig.parameters["dtype"].default is np.float32
    for sfx in ("_FP32", "_FP64"):
        for name in ("Device", "Sector", "SpinHalfBasis", "FermionBasis", "HubbardBasis", "TJBasis",
                     "SpinHalfSite", "FermionSite", "HubbardSite", "TJSite", "OpSum",
                     "MatrixFreeHamiltonianHIP", "lanczos_ground_state_HIP", "davidson_lowest_HIP",
                     "continued_fraction_coeffs_HIP", "ftlm_HIP"):
            assert hasattr(cpp, name + sfx), name + sfx
    assert cpp.Device_FP32.is_gpu_build() and cpp.Device_FP32.backend_name() == "hip"


def test_bases_and_opsum_as_reference_tests():
    import qkrylov
    import qkrylov._qkrylov_cpp as cpp
    b = qkrylov.SpinHalfBasis(4)
    assert b.size == 16 and b.nsites == 4
    assert qkrylov.FermionBasis(4, conserve_n=True, n=2).size == 6
    assert qkrylov.TJBasis(2).size == 9
    assert qkrylov.SpinHalfBasis(N=4, sz=0).size == 6
    os_ = qkrylov.OpSum()
    for i in range(3):
        os_ += 1.0 * qkrylov.Sz(i) * qkrylov.Sz(i + 1) + 0.5 * (
            qkrylov.Sp(i) * qkrylov.Sm(i + 1) + qkrylov.Sm(i) * qkrylov.Sp(i + 1))
    assert os_.size == 9
    with pytest.raises(ValueError):
        os_ += (1.0, "Sz")
    s = cpp.Sector_FP64()
    s.use_sz, s.sz2 = True, 0
    lb = cpp.SpinHalfBasis_FP64(6, s)
    assert lb.size() == 20 and lb.nsites() == 6 and lb.index(lb.state(7)) == 7
    h = cpp.Sector_FP32()
    h.use_nup, h.nup, h.use_ndn, h.ndn = True, 1, True, 1
    assert cpp.HubbardBasis_FP32(2, h).size() == 4


def test_no_gpu_here_fails_loudly():
    import qkrylov
    assert qkrylov.find_gpu() is None and qkrylov.gpu_count() == 0
    with pytest.raises(Exception):
        qkrylov.MatrixFreeHamiltonian(qkrylov.SpinHalfBasis(2), qkrylov.SpinHalfSite(),
                                      qkrylov.OpSum())
