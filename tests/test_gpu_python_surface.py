"""Hamiltonian-side Python surface of the reference package on the GPU
(bindings/python/qkrylov/hamiltonian.py:8-177, tests/test_basic.py:14-49,
tests/test_workflow.py): dimension, diagonal, dtype narrowing, H @ x,
LinearOperator, explicit CSR, wrong-size input, unknown operators."""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

pytestmark = pytest.mark.gpu


def test_hamiltonian_as_reference_tests(qk):
    # test_basic.py:14-25
    os_ = qk.OpSum()
    os_.add_term(1.0, "Sz", 0, "Sz", 1)
    H = qk.MatrixFreeHamiltonian(qk.SpinHalfBasis(2), qk.SpinHalfSite(), os_)
    assert H.dimension == 4
    assert len(H.diagonal()) == 4
    # test_basic.py:45-49
    with pytest.raises(ValueError):
        H.apply(np.zeros(10, dtype=np.complex128))
    assert qk.gpu_count() >= 1 and qk.find_gpu() == "hip"


def test_dtype_to_matmul_linear_operator_sparse(qk, orc, lat):
    n = 10
    ops = lat.heisenberg_chain(n)
    H = qk.MatrixFreeHamiltonian(qk.SpinHalfBasis(n, conserve_sz=True, sz=0), qk.SpinHalfSite(), ops)
    A = orc.Ham(orc.Basis("spin", n, {"use_sz": True, "sz2": 0}), "spin", ops).to_scipy().toarray()
    rng = np.random.default_rng(0)
    x = rng.standard_normal(H.dimension) + 1j * rng.standard_normal(H.dimension)
    y = H @ x
    assert y.dtype == np.complex128
    np.testing.assert_allclose(y, A @ x, rtol=0, atol=1e-12 * np.abs(A @ x).max())
    H32 = H.to(dtype=np.float32)
    y32 = H32.apply(x)
    assert y32.dtype == np.complex64 and H32.diagonal().dtype == np.complex64
    np.testing.assert_allclose(y32, (A @ x).astype(np.complex64), rtol=1e-6, atol=1e-5)
    S = H.to_sparse(block=37)
    assert np.array_equal(S.toarray(), np.where(A != 0, A, 0))
    w = sla.eigsh(H.aslinearoperator(), k=1, which="SA", tol=1e-12)[0][0]
    e = qk.lanczos_ground_state(H).energy
    assert abs(w - e) < 1e-9


def test_unknown_operator_is_rejected(qk):
    os_ = qk.OpSum()
    os_ += 1.0 * (qk.Bdag(0) * qk.B(1))
    with pytest.raises(RuntimeError):
        qk.MatrixFreeHamiltonian(qk.SpinHalfBasis(2), qk.SpinHalfSite(), os_)
