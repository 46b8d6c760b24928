"""OpSum builders for the standard lattice models.

Conventions (fixed so that every caller builds the same operator):
site index i = x + Lx*y; bonds are listed bond-major; factors act on the ket
in list order (the reference applies the first factor first,
src/hamiltonian/matrix_free_hamiltonian.cpp:54-68), so c^dag_i c_j is the
factor list [C_j, Cdag_i].
"""
from __future__ import annotations

from typing import List, Tuple

from . import OpSum

Bond = Tuple[int, int]


def chain_bonds(n: int, periodic: bool = True) -> List[Bond]:
    bonds = [(i, i + 1) for i in range(n - 1)]
    if periodic and n > 2:
        bonds.append((n - 1, 0))
    return bonds


def square_bonds(lx: int, ly: int) -> List[Bond]:
    """Nearest-neighbour bonds of a periodic lx*ly square lattice: for each
    site, the bond to the right and the bond upwards."""
    bonds = []
    for y in range(ly):
        for x in range(lx):
            i = x + lx * y
            bonds.append((i, (x + 1) % lx + lx * y))
            bonds.append((i, x + lx * ((y + 1) % ly)))
    return bonds


def square_diagonal_bonds(lx: int, ly: int) -> List[Bond]:
    """Both next-nearest (diagonal) bonds of every site, periodic."""
    bonds = []
    for y in range(ly):
        for x in range(lx):
            i = x + lx * y
            bonds.append((i, (x + 1) % lx + lx * ((y + 1) % ly)))
            bonds.append((i, (x + 1) % lx + lx * ((y - 1) % ly)))
    return bonds


def add_xxz(os: OpSum, bonds: List[Bond], jz: float, jxy: float) -> OpSum:
    """Per bond: Jz Sz_i Sz_j + Jxy/2 (Sp_i Sm_j + Sm_i Sp_j)."""
    for i, j in bonds:
        os.add_term(jz, "Sz", i, "Sz", j)
        os.add_term(0.5 * jxy, "Sp", i, "Sm", j)
        os.add_term(0.5 * jxy, "Sm", i, "Sp", j)
    return os


def heisenberg_chain(n: int, j: float = 1.0, periodic: bool = True) -> OpSum:
    return add_xxz(OpSum(), chain_bonds(n, periodic), j, j)


def xxz_chain(n: int, jxy: float, jz: float, periodic: bool = True) -> OpSum:
    return add_xxz(OpSum(), chain_bonds(n, periodic), jz, jxy)


def j1j2_square(lx: int, ly: int, j1: float = 1.0, j2: float = 0.5) -> OpSum:
    os = add_xxz(OpSum(), square_bonds(lx, ly), j1, j1)
    return add_xxz(os, square_diagonal_bonds(lx, ly), j2, j2)


def tj_model(bonds: List[Bond], t: float, jx: float) -> OpSum:
    """t-J: -t sum_sigma (c+_i c_j + c+_j c_i) + J (S_i.S_j - n_i n_j / 4).
    Per bond: 4 hops, 2 four-factor spin flips (J/2), and the diagonal part
    -J/2 (Nup_i Ndn_j + Ndn_i Nup_j)."""
    os = OpSum()
    for i, j in bonds:
        for cd, c in (("CdagUp", "CUp"), ("CdagDn", "CDn")):
            os.add_term(-t, c, j, cd, i)   # c+_i c_j
            os.add_term(-t, c, i, cd, j)   # c+_j c_i
        # S+_i S-_j = c+_i,up c_i,dn c+_j,dn c_j,up
        os.add_term(0.5 * jx, "CUp", j, "CdagDn", j, "CDn", i, "CdagUp", i)
        # S-_i S+_j = c+_i,dn c_i,up c+_j,up c_j,dn
        os.add_term(0.5 * jx, "CDn", j, "CdagUp", j, "CUp", i, "CdagDn", i)
        os.add_term(-0.5 * jx, "Nup", i, "Ndn", j)
        os.add_term(-0.5 * jx, "Ndn", i, "Nup", j)
    return os


def hubbard_model(bonds: List[Bond], nsites: int, t: float, u: float) -> OpSum:
    os = OpSum()
    for i, j in bonds:
        for cd, c in (("CdagUp", "CUp"), ("CdagDn", "CDn")):
            os.add_term(-t, c, j, cd, i)
            os.add_term(-t, c, i, cd, j)
    for i in range(nsites):
        os.add_term(u, "Nupdn", i)
    return os


def spinless_fermion_chain(n: int, t: float, v: float, periodic: bool = False) -> OpSum:
    os = OpSum()
    for i, j in chain_bonds(n, periodic):
        os.add_term(-t, "C", j, "Cdag", i)
        os.add_term(-t, "C", i, "Cdag", j)
        os.add_term(v, "N", i, "N", j)
    return os
