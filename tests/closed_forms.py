"""Closed forms used to pin results at sizes no CPU reference run reaches.

bethe_ground_energy(N): ground-state energy of the periodic spin-1/2
Heisenberg chain H = J sum_i S_i . S_{i+1} (J = 1, N even, Sz = 0) from the
Bethe ansatz (Bethe 1931; e.g. Karbach & Mueller, Computers in Physics 11,
36 (1997)).  With M = N/2 magnon rapidities l_j and Bethe quantum numbers
I_j = -(M-1)/2, ..., (M-1)/2 the logarithmic Bethe equations read

    2 N atan(2 l_j) = 2 pi I_j + sum_{k != j} 2 atan(l_j - l_k)

and E = N/4 - sum_j 2 / (1 + 4 l_j^2).  Solved by Newton's method from the
free-magnon start l_j = tan(pi I_j / N) / 2.  N = 2 gives -1.5 (the one bond
counted twice), N = 4 gives -2.
"""
import numpy as np


def bethe_ground_energy(n: int, tol: float = 1e-15, max_iter: int = 100) -> float:
    if n % 2:
        raise ValueError("even N only")
    m = n // 2
    I = np.arange(m) - (m - 1) / 2.0
    lam = 0.5 * np.tan(np.pi * I / n)
    for _ in range(max_iter):
        diff = lam[:, None] - lam[None, :]
        F = 2 * n * np.arctan(2 * lam) - 2 * np.pi * I - 2 * np.arctan(diff).sum(1)
        K = 2.0 / (1.0 + diff ** 2)
        J = K.copy()
        np.fill_diagonal(J, 0.0)
        J[np.diag_indices(m)] = 4 * n / (1 + 4 * lam ** 2) - (K.sum(1) - 2.0)
        step = np.linalg.solve(J, F)
        lam = lam - step
        if np.abs(step).max() < tol:
            break
    return float(n / 4.0 - np.sum(2.0 / (1.0 + 4.0 * lam ** 2)))
