"""Reference-style scripts through `import qkrylov` on the GPU: the
assertions of the reference package's tests (bindings/python/tests/
test_basic.py, test_workflow.py) and of its README workflow, written here,
with the reference defaults (device="cpu", dtype=np.float32)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _heis_chain(qkrylov, n, periodic):
    os_ = qkrylov.OpSum()
    bonds = [(i, i + 1) for i in range(n - 1)] + ([(n - 1, 0)] if periodic else [])
    for i, j in bonds:
        os_ += 1.0 * qkrylov.Sz(i) * qkrylov.Sz(j) + 0.5 * (
            qkrylov.Sp(i) * qkrylov.Sm(j) + qkrylov.Sm(i) * qkrylov.Sp(j))
    return os_


def test_basic_and_workflow():
    import qkrylov
    print("[dropin] basic + workflow", flush=True)
    os_ = qkrylov.OpSum()
    os_.add_term(1.0, "Sz", 0, "Sz", 1)
    H = qkrylov.MatrixFreeHamiltonian(qkrylov.SpinHalfBasis(2), qkrylov.SpinHalfSite(), os_)
    assert H.dimension == 4 and len(H.diagonal()) == 4
    assert H.diagonal().dtype == np.complex64          # FP32 default narrows
    with pytest.raises(ValueError):
        H.apply(np.zeros(10, dtype=np.complex128))
    # 4-site OBC chain, Sz = 0 (test_workflow.py)
    basis = qkrylov.SpinHalfBasis(N=4, sz=0)
    H = qkrylov.MatrixFreeHamiltonian(basis, qkrylov.SpinHalfSite(), _heis_chain(qkrylov, 4, False))
    assert H.dimension == 6
    res = qkrylov.lanczos_ground_state(H, 200, 1e-12)
    assert math.isclose(res.energy, -1.6160254038, abs_tol=1e-9)
    assert len(res.eigenvector) == 6 and res.eigenvector.dtype == np.complex64
    assert qkrylov.find_gpu() == "hip" and qkrylov.gpu_count() >= 1


def test_solvers_and_low_level_names(orc, lat):
    import qkrylov
    import qkrylov._qkrylov_cpp as cpp
    print("[dropin] solvers", flush=True)
    n = 10
    ops = lat.heisenberg_chain(n)
    basis = qkrylov.SpinHalfBasis(n, conserve_sz=True, sz=0)
    H = qkrylov.MatrixFreeHamiltonian(basis, qkrylov.SpinHalfSite(), ops, device="cpu",
                                      dtype=np.float64)
    oh = orc.Ham(orc.Basis("spin", n, {"use_sz": True, "sz2": 0}), "spin", ops)
    e_ref = oh.lanczos(200, 1e-12)[0]
    assert abs(qkrylov.lanczos_ground_state(H).energy - e_ref) < 1e-10
    dv = qkrylov.davidson_lowest(H, n_eig=1, max_subspace=20, tol=1e-8)
    dense = np.linalg.eigvalsh(oh.to_scipy().toarray())
    assert abs(dv.eigenvalues[0] - dense[0]) < 1e-8 and len(dv.eigenvectors) == 1
    # dynamics through the reference's signature evaluate_spectral_function(res, w, E0, eta)
    rng = np.random.default_rng(3)
    phi = rng.standard_normal(H.dimension) + 0j
    cf = qkrylov.continued_fraction_coeffs(H, phi, 40)
    a_ref, b_ref, nrm_ref = oh.cf(phi, 40)
    assert abs(cf.norm_phi0 - nrm_ref) <= 1e-12 * nrm_ref
    assert np.allclose(cf.alphas[:10], a_ref[:10], rtol=0, atol=1e-8)
    s = qkrylov.evaluate_spectral_function(cf, 0.7, e_ref, 0.05)
    assert abs(s - orc.spectral(a_ref, b_ref, nrm_ref, 0.7, e_ref, 0.05)) <= 1e-8 * max(1.0, abs(s))
    ft = qkrylov.ftlm(H, 0.5, 4, 30)
    ft_ref = qkrylov.ftlm(qkrylov.MatrixFreeHamiltonian(basis, qkrylov.SpinHalfSite(), ops,
                                                        dtype=np.float64), 0.5, 4, 30)
    assert ft.partition_function == ft_ref.partition_function and ft.beta == 0.5
    # the nanobind-level names of the reference build, HIP backend
    s_ = cpp.Sector_FP64()
    s_.use_sz, s_.sz2 = True, 0
    Hc = cpp.MatrixFreeHamiltonianHIP_FP64(cpp.SpinHalfBasis_FP64(n, s_), cpp.SpinHalfSite_FP64(), ops)
    assert Hc.dimension() == H.dimension
    e, v = cpp.lanczos_ground_state_HIP_FP64(Hc, 200, 1e-12)
    assert abs(e - e_ref) < 1e-10 and v.dtype == np.complex128
    e32, v32 = cpp.lanczos_ground_state_HIP_FP32(
        cpp.MatrixFreeHamiltonianHIP_FP32(cpp.SpinHalfBasis_FP32(n, s_), cpp.SpinHalfSite_FP32(), ops),
        200, 1e-12)
    assert abs(e32 - e_ref) < 1e-10 and v32.dtype == np.complex64
    al, be, nr = cpp.continued_fraction_coeffs_HIP(Hc, phi, 40)
    assert np.array_equal(al, cf.alphas) and nr == cf.norm_phi0
