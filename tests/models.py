"""Model zoo shared by the parity tests: every entry builds the same OpSum
for the HIP library and for the oracle."""
from __future__ import annotations

import qkrylov_amd as qk
from qkrylov_amd import lattices as lat

SITE = {"spin": qk.SpinHalfSite, "fermion": qk.FermionSite, "hubbard": qk.HubbardSite,
        "tj": qk.TJSite}
BASIS = {"spin": qk.SpinHalfBasis, "fermion": qk.FermionBasis, "hubbard": qk.HubbardBasis,
         "tj": qk.TJBasis}


def make_sector(sec: dict) -> qk.Sector:
    s = qk.Sector()
    if sec.get("use_sz"):
        s.set_sz(sec["sz2"])
    if sec.get("use_n"):
        s.set_n(sec["n"])
    if sec.get("use_nup") or sec.get("use_ndn"):
        s.set_hubbard_particles(sec.get("nup", 0), sec.get("ndn", 0))
    return s


def sz(sz2):
    return {"use_sz": True, "sz2": sz2}


def nn(n):
    return {"use_n": True, "n": n}


def ud(nup, ndn):
    return {"use_nup": True, "nup": nup, "use_ndn": True, "ndn": ndn}


def _fields(n):
    os = lat.xxz_chain(n, 1.0, 0.7)
    for i in range(n):
        os.add_term(0.3, "Sx", i)
        os.add_term(-0.2, "Sy", i)
        os.add_term(0.1 * (i + 1), "Sz", i)
    return os


def _three_body(n):
    os = lat.heisenberg_chain(n)
    for i in range(n - 2):
        os.add_term(0.25, "Sz", i, "Sz", i + 1, "Sz", i + 2)
        os.add_term(0.15, "Sp", i, "Sz", i + 1, "Sm", i + 2)
        os.add_term(0.15, "Sm", i, "Sz", i + 1, "Sp", i + 2)
    return os


def _complex_coeffs(n):
    os = qk.OpSum()
    for i, j in lat.chain_bonds(n):
        os.add_term(0.8, "Sz", i, "Sz", j)
        os.add_term(0.3 + 0.7j, "Sp", i, "Sm", j)
        os.add_term(0.3 - 0.7j, "Sm", i, "Sp", j)
        os.add_term(0.2 + 0.1j, "Sp", i, "Sm", j)   # non-Hermitian on purpose
    return os


def _fermion_complex(n):
    os = lat.spinless_fermion_chain(n, 1.0, 0.5, periodic=True)
    for i in range(n):
        os.add_term(0.3 * i - 0.4, "N", i)
    os.add_term(0.25j, "C", 0, "Cdag", n - 1)
    os.add_term(-0.25j, "C", n - 1, "Cdag", 0)
    return os


def _hubbard(n, u):
    os = lat.hubbard_model(lat.chain_bonds(n), n, 1.0, u)
    for i in range(n):
        os.add_term(0.1 * i, "Nup", i)
        os.add_term(-0.05 * i, "Ndn", i)
    return os


def _fermion_ops_on_spin(n):
    os = qk.OpSum()
    for i in range(n - 1):
        os.add_term(-1.0, "C", i + 1, "Cdag", i)
        os.add_term(-1.0, "C", i, "Cdag", i + 1)
        os.add_term(0.5, "N", i)
    os.add_term(1.0, "Id", 0)
    return os


def _outside_bits(n):
    os = lat.heisenberg_chain(n)
    os.add_term(0.7, "Sz", n + 3)             # reads a bit above nsites
    os.add_term(0.4, "Sp", n + 1, "Sm", 0)    # leaves the basis
    os.add_term(0.9, "Sx", 2)                 # leaves the Sz sector
    return os


# name -> (basis kind, nsites, sector, site kind, opsum factory)
MODELS = {
    "heis_chain10_sz0": ("spin", 10, sz(0), "spin", lambda: lat.heisenberg_chain(10)),
    "heis_obc4_sz0": ("spin", 4, sz(0), "spin", lambda: lat.heisenberg_chain(4, periodic=False)),
    "xxz_fields_full9": ("spin", 9, {}, "spin", lambda: _fields(9)),
    "j1j2_4x3_sz0": ("spin", 12, sz(0), "spin", lambda: lat.j1j2_square(4, 3, 1.0, 0.5)),
    "three_body_12_sz2": ("spin", 12, sz(2), "spin", lambda: _three_body(12)),
    "complex_coeff_10_sz0": ("spin", 10, sz(0), "spin", lambda: _complex_coeffs(10)),
    "outside_bits_8_sz0": ("spin", 8, sz(0), "spin", lambda: _outside_bits(8)),
    "fermion_complex_11_n4": ("fermion", 11, nn(4), "fermion", lambda: _fermion_complex(11)),
    "fermion_full8": ("fermion", 8, {}, "fermion", lambda: _fermion_complex(8)),
    "hubbard_chain6_u4": ("hubbard", 6, ud(3, 2), "hubbard", lambda: _hubbard(6, 4.0)),
    "hubbard_full5": ("hubbard", 5, {}, "hubbard", lambda: _hubbard(5, 2.0)),
    "hubbard_chain8_half": ("hubbard", 8, ud(4, 4), "hubbard", lambda: _hubbard(8, 8.0)),
    "tj_3x2": ("tj", 6, ud(2, 2), "tj", lambda: lat.tj_model(lat.square_bonds(3, 2), 1.0, 0.4)),
    "tj_4x3": ("tj", 12, ud(5, 5), "tj", lambda: lat.tj_model(lat.square_bonds(4, 3), 1.0, 0.4)),
    "tj_full4": ("tj", 4, {}, "tj", lambda: lat.tj_model(lat.chain_bonds(4), 1.0, 0.4)),
    "fermion_ops_on_spin8": ("spin", 8, {}, "fermion", lambda: _fermion_ops_on_spin(8)),
    "heis_chain20_sz0": ("spin", 20, sz(0), "spin", lambda: lat.heisenberg_chain(20)),
}


def build_pair(orc, name):
    """(HIP Hamiltonian, oracle Ham, opsum) for a model of MODELS."""
    kind, n, sec, site, f = MODELS[name]
    os = f()
    basis = BASIS[kind](n, sector=make_sector(sec))
    H = qk.MatrixFreeHamiltonian(basis, SITE[site](), os)
    ob = orc.Basis(kind, n, sec)
    oh = orc.Ham(ob, site, os)
    return H, oh, os
