"""Writes the oracle fixtures of the two benchmark workloads whose CPU
reference takes longer than a GPU test should wait (about a minute each on
eight host cores), so that the GPU tests compare against stored oracle output:

tj4x4_lowest.json     t-J on the 4x4 periodic square lattice (t=1, J=0.4),
                      7 up and 7 down (d = 411,840): the 10 lowest
                      eigenvalues of the oracle's CSR (the reference's
                      matrix_free_hamiltonian.cpp:45-91 compile) by scipy
                      eigsh, with each eigenvector's residual.  The basis is
                      built here from its definition (7 up electrons on even
                      bits, 7 down on odd bits, no site doubly occupied,
                      ascending order), not by the library.  k = 10: the
                      second level is five-fold, and with k = 8 the solver
                      returned only three of its five states, so the fixture
                      keeps two values past the lowest eight.
ftlm_xxz20.json       FTLM (reference ftlm.cpp:122-170) on the periodic XXZ
                      chain N=20 (Jxy=1, Jz=1.5), every Sz sector, R=8
                      random vectors per sector (seed 42, re-seeded per
                      sector), M=100 steps, 200 temperatures
                      T_i = 0.05 * 200^(i/199): Z, E, Cv per sector and
                      temperature, from the oracle.

    python tests/golden/make_config_fixtures.py
"""
import itertools
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from oracle import oracle as orc  # noqa: E402
from qkrylov_amd import lattices as lat  # noqa: E402


def tj_states(nsites: int, nup: int, ndn: int) -> np.ndarray:
    """Ascending t-J states: up electron of site i at bit 2i, down at 2i+1."""
    out = []
    for up in itertools.combinations(range(nsites), nup):
        u = sum(1 << (2 * i) for i in up)
        rest = [s for s in range(nsites) if s not in up]
        for dn in itertools.combinations(rest, ndn):
            out.append(u | sum(1 << (2 * i + 1) for i in dn))
    return np.sort(np.array(out, dtype=np.uint64))


def tj_fixture():
    import scipy.sparse.linalg as sla
    states = tj_states(16, 7, 7)
    ob = orc.Basis.from_states("tj", 16, states)
    oh = orc.Ham(ob, "tj", lat.tj_model(lat.square_bonds(4, 4), 1.0, 0.4))
    A = oh.to_scipy()
    t = time.time()
    w, v = sla.eigsh(A, k=10, which="SA", tol=1e-14, ncv=60)
    order = np.argsort(w)
    w, v = w[order], v[:, order]
    res = [float(np.linalg.norm(A @ v[:, k] - w[k] * v[:, k])) for k in range(len(w))]
    return {"workload": "t-J 4x4 periodic, t=1, J=0.4, nup=ndn=7", "dim": int(states.size),
            "nnz": int(A.nnz), "eigenvalues": [float(x) for x in w], "residuals": res,
            "solver": f"scipy.sparse.linalg.eigsh(k=10, which='SA', tol=1e-14, ncv=60), {time.time() - t:.0f} s"}


def ftlm_fixture():
    n, R, M = 20, 8, 100
    temps = 0.05 * 200.0 ** (np.arange(200) / 199.0)
    betas = 1.0 / temps
    os_ = lat.xxz_chain(n, 1.0, 1.5)
    out = {"workload": "XXZ chain N=20, Jxy=1, Jz=1.5, all 21 Sz sectors", "R": R, "M": M,
           "betas": [float(b) for b in betas], "sectors": []}
    for k in range(n + 1):
        sz2 = 2 * k - n
        oh = orc.Ham(orc.Basis("spin", n, {"use_sz": True, "sz2": sz2}), "spin", os_)
        Z, E, Cv = oh.ftlm(betas, R, M)
        out["sectors"].append({"sz2": sz2, "dim": int(oh.dim), "Z": [float(x) for x in Z],
                               "E": [float(x) for x in E], "Cv": [float(x) for x in Cv]})
    return out


def main():
    which = sys.argv[1:] or ["tj", "ftlm"]
    if "tj" in which:
        (HERE / "tj4x4_lowest.json").write_text(json.dumps(tj_fixture(), indent=1) + "\n")
    if "ftlm" in which:
        (HERE / "ftlm_xxz20.json").write_text(json.dumps(ftlm_fixture()) + "\n")


if __name__ == "__main__":
    main()
