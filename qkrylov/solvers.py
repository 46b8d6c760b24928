"""Solvers with the reference wrapper's signatures and result types
(bindings/python/qkrylov/solvers.py:6-146).  An FP32 Hamiltonian
(dtype=np.float32) narrows returned vectors to complex64, as the
reference's FP32 build returns them; energies stay Python floats."""
import numpy as np

import qkrylov_amd as _amd


def _narrow(H, v):
    if v is not None and getattr(H, "dtype", np.float64) == np.float32:
        return np.asarray(v).astype(np.complex64)
    return v


class LanczosResult:
    """energy and eigenvector (reference solvers.py:6-21); iterations and
    converged are extra."""

    def __init__(self, energy, eigenvector, iterations=None, converged=None):
        self.energy, self.eigenvector = energy, eigenvector
        self.iterations, self.converged = iterations, converged

    def __repr__(self):
        return f"LanczosResult(energy={self.energy:.10f})"


def lanczos_ground_state(H, maxiter: int = 200, tol: float = 1e-12) -> LanczosResult:
    r = _amd.lanczos_ground_state(H, maxiter, tol)
    return LanczosResult(float(r.energy), _narrow(H, r.eigenvector), r.iterations, r.converged)


class DavidsonResult:
    def __init__(self, eigenvalues, eigenvectors, iterations=None, converged=None):
        self.eigenvalues, self.eigenvectors = eigenvalues, eigenvectors
        self.iterations, self.converged = iterations, converged

    def __repr__(self):
        return f"DavidsonResult(energies={self.eigenvalues})"


def davidson_lowest(H, n_eig: int = 1, max_subspace: int = 20, tol: float = 1e-8) -> DavidsonResult:
    r = _amd.davidson_lowest(H, n_eig, max_subspace, tol)
    vecs = [_narrow(H, v) for v in (r.eigenvectors or [])]
    return DavidsonResult(np.array(r.eigenvalues, dtype=float), vecs, r.iterations, r.converged)


class DynamicsResult:
    def __init__(self, alphas, betas, norm_phi0):
        self.alphas = np.asarray(alphas)
        self.betas = np.asarray(betas)
        self.norm_phi0 = norm_phi0


def continued_fraction_coeffs(H, phi0: np.ndarray, n_iter: int = 100) -> DynamicsResult:
    r = _amd.continued_fraction_coeffs(H, np.asarray(phi0, dtype=np.complex128), n_iter)
    return DynamicsResult(r.alphas, r.betas, r.norm_phi0)


def evaluate_spectral_function(res: DynamicsResult, omega: float, E0: float, eta: float = 0.1) -> float:
    return _amd.evaluate_spectral_function(np.ascontiguousarray(res.alphas, dtype=np.float64),
                                           np.ascontiguousarray(res.betas, dtype=np.float64),
                                           res.norm_phi0, omega, E0, eta)


class FTLMResult:
    def __init__(self, beta, partition_function, internal_energy, specific_heat):
        self.beta = beta
        self.partition_function = partition_function
        self.internal_energy = internal_energy
        self.specific_heat = specific_heat


def ftlm(H, beta: float, n_random: int = 50, n_steps: int = 100) -> FTLMResult:
    r = _amd.ftlm(H, beta, n_random, n_steps)
    return FTLMResult(r.beta, r.partition_function, r.internal_energy, r.specific_heat)
