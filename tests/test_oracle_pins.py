"""Pins for the oracle (CPU).  The oracle is only trusted because of these:
the reference's own basis/site/OpSum sources compiled into oracle/_ref,
the reference tests' known answers, and closed forms / brute force on small
inputs, each chosen so that a plausible mistake (sign, index, conjugation,
factor order, missing factor 2) fails one of them."""
import itertools
import math

import numpy as np
import pytest

from models import MODELS

REF = pytest.mark.skipif(
    not __import__("oracle.oracle", fromlist=["x"]).ref_available(),
    reason="oracle/_ref not built (reference tree absent)")

BASIS_CASES = [
    ("spin", 1, {}), ("spin", 6, {}), ("spin", 10, {"use_sz": True, "sz2": 0}),
    ("spin", 11, {"use_sz": True, "sz2": -3}), ("spin", 12, {"use_sz": True, "sz2": 12}),
    ("spin", 9, {"use_sz": True, "sz2": 2}),     # parity mismatch -> empty
    ("fermion", 8, {}), ("fermion", 13, {"use_n": True, "n": 5}),
    ("hubbard", 3, {}), ("hubbard", 6, {"use_nup": True, "nup": 3, "use_ndn": True, "ndn": 2}),
    ("hubbard", 5, {"use_nup": True, "nup": 2}),  # only one flag (C++ API)
    ("tj", 4, {}), ("tj", 6, {"use_nup": True, "nup": 2, "use_ndn": True, "ndn": 2}),
    ("tj", 5, {"use_ndn": True, "ndn": 1}),
]


@REF
@pytest.mark.parametrize("kind,n,sec", BASIS_CASES)
def test_basis_matches_reference_build(orc, kind, n, sec):
    o = orc.Basis(kind, n, sec)
    r = orc.Basis(kind, n, sec, ref=True)
    assert o.size == r.size
    assert np.array_equal(o.states(), r.states())
    st = o.states()
    for s in list(st[:50]) + list(st[-50:]) + [st.max() + 1 if st.size else 1, 3, 12345]:
        assert o.index(int(s)) == r.index(int(s))


def test_basis_closed_forms(orc):
    assert orc.Basis("spin", 4, {"use_sz": True, "sz2": 0}).size == 6       # test_c_api.cpp:23
    assert orc.Basis("fermion", 4).size == 16                                # test_fermion_basis.cpp:12
    assert orc.Basis("fermion", 4, {"use_n": True, "n": 2}).size == 6        # :20
    assert orc.Basis("tj", 2).size == 9                                      # test_tj_basis.cpp:14
    assert orc.Basis("hubbard", 2).size == 16                                # runtests.jl:74
    assert orc.Basis("spin", 16, {"use_sz": True, "sz2": 0}).size == 12870
    assert orc.Basis("spin", 20, {"use_sz": True, "sz2": 0}).size == 184756
    assert orc.Basis("tj", 12, {"use_nup": True, "nup": 5, "use_ndn": True, "ndn": 5}).size == 16632
    assert sum(orc.Basis("spin", 12, {"use_sz": True, "sz2": 2 * k - 12}).size
               for k in range(13)) == 2 ** 12
    # multinomial 12!/(5!5!2!) and product of binomials for Hubbard
    assert math.factorial(12) // (math.factorial(5) ** 2 * 2) == 16632
    assert orc.Basis("hubbard", 6, {"use_nup": True, "nup": 3, "use_ndn": True,
                                    "ndn": 2}).size == math.comb(6, 3) * math.comb(6, 2)


def test_colex_rank_is_combinatorial_number_system(orc):
    """Ascending order at fixed popcount: index = sum_j C(p_j, j) (j from 1)."""
    n, k = 12, 5
    o = orc.Basis("spin", n, {"use_sz": True, "sz2": 2 * k - n})
    for idx, s in enumerate(o.states()):
        bits = [p for p in range(n) if (int(s) >> p) & 1]
        assert sum(math.comb(p, j + 1) for j, p in enumerate(bits)) == idx


SITE_OPS = {"spin": ["Sz", "Sp", "Sm", "Sx", "Sy"], "fermion": ["N", "Id", "C", "Cdag"],
            "hubbard": ["Nup", "Ndn", "Nupdn", "CUp", "CdagUp", "CDn", "CdagDn"],
            "tj": ["Nup", "Ndn", "CUp", "CdagUp", "CDn", "CdagDn"]}


@REF
@pytest.mark.parametrize("kind", sorted(SITE_OPS))
def test_site_actions_match_reference_build(orc, kind):
    nbits = 8
    nsites = nbits // 2 if kind in ("hubbard", "tj") else nbits
    for op in SITE_OPS[kind]:
        for site in range(nsites):
            for s in range(1 << nbits):
                assert orc.site_apply(kind, op, site, s) == orc.site_apply(kind, op, site, s, ref=True)
    with pytest.raises(ValueError):
        orc.site_apply(kind, "Bogus", 0, 0)
    with pytest.raises(ValueError):
        orc.site_apply(kind, "Bogus", 0, 0, ref=True)


def test_site_fixtures(orc):
    # tests/test_fermion_site.cpp:13-26
    assert orc.site_apply("fermion", "N", 1, 0b1010) == (True, 0b1010, 1.0)
    assert orc.site_apply("fermion", "N", 0, 0b1010) == (True, 0b1010, 0.0)
    assert orc.site_apply("fermion", "C", 1, 0b1010) == (True, 0b1000, 1.0)
    assert orc.site_apply("fermion", "C", 3, 0b1010) == (True, 0b0010, -1.0)
    # tests/test_tj_site.cpp:13-18
    assert orc.site_apply("tj", "CUp", 0, 0b0001) == (True, 0, 1.0)
    assert orc.site_apply("tj", "CdagDn", 0, 0b0001)[0] is False
    # spinhalf_site.cpp:97-116: Sy = -0.5i on up, +0.5i on down
    assert orc.site_apply("spin", "Sy", 2, 0b100) == (True, 0, -0.5j)
    assert orc.site_apply("spin", "Sy", 2, 0) == (True, 0b100, 0.5j)
    # Hubbard: the down phase counts the up bit of the same site (hubbard_site.cpp:43-49)
    assert orc.site_apply("hubbard", "CDn", 1, 0b1110) == (True, 0b0110, 1.0)
    assert orc.site_apply("hubbard", "CDn", 1, 0b1010) == (True, 0b0010, -1.0)


@REF
@pytest.mark.parametrize("name", sorted(n for n in MODELS if n != "heis_chain20_sz0"))
def test_csr_matches_reference_build(orc, name):
    kind, n, sec, site, f = MODELS[name]
    os_ = f()
    oh = orc.Ham(orc.Basis(kind, n, sec), site, os_)
    rows, cols, vals, diag = orc.RefCsr(orc.Basis(kind, n, sec, ref=True), site, os_).get()
    r2, c2, v2 = oh.csr()
    assert np.array_equal(rows, r2) and np.array_equal(cols, c2)
    assert np.array_equal(vals.view(np.float64), v2.view(np.float64))
    assert np.array_equal(diag.view(np.float64), oh.diag().view(np.float64))


@REF
def test_mutations_are_caught(orc):
    """A planted sign or conjugation mistake in the oracle must break the pins."""
    lib = orc.lib()
    try:
        for mut, name in ((1, "fermion_complex_11_n4"), (1, "tj_3x2"), (2, "complex_coeff_10_sz0")):
            lib.orc_set_mutation(mut)
            kind, n, sec, site, f = MODELS[name]
            os_ = f()
            _, _, v2 = orc.Ham(orc.Basis(kind, n, sec), site, os_).csr()
            _, _, vals, _ = orc.RefCsr(orc.Basis(kind, n, sec, ref=True), site, os_).get()
            assert not np.array_equal(vals, v2), (mut, name)
    finally:
        lib.orc_set_mutation(0)


def test_julia_diagonal_fixture(orc, qk):
    """bindings/julia/test/runtests.jl:102-106: Sz0Sz1 + 0.5 Sz0Sz1Sz2 on 3
    sites, <111|H|111> = 0.3125."""
    os_ = qk.OpSum()
    os_.add_term(1.0, "Sz", 0, "Sz", 1)
    os_.add_term(0.5, "Sz", 0, "Sz", 1, "Sz", 2)
    oh = orc.Ham(orc.Basis("spin", 3), "spin", os_)
    assert oh.diag()[7] == 0.3125
    x = np.zeros(8, complex)
    x[7] = 1
    assert oh.apply(x)[7] == 0.3125


def _dense(oh):
    return oh.to_scipy().toarray()


def test_known_energies(orc, lat):
    cases = [(("spin", 2, {}), lat.heisenberg_chain(2, periodic=False), -0.75),
             (("spin", 4, {"use_sz": True, "sz2": 0}), lat.heisenberg_chain(4, periodic=False),
              -(3 + 2 * math.sqrt(3)) / 4),
             (("spin", 4, {}), lat.heisenberg_chain(4), -2.0)]
    for (kind, n, sec), os_, e in cases:
        oh = orc.Ham(orc.Basis(kind, n, sec), "spin", os_)
        assert abs(oh.lanczos(200, 1e-12)[0] - e) < 1e-10
        assert abs(np.linalg.eigvalsh(_dense(oh))[0] - e) < 1e-12
    assert abs(-(3 + 2 * math.sqrt(3)) / 4 + 1.6160254038) < 1e-10   # test_workflow.cpp:46


def test_lanczos_matches_dense(orc):
    for name in ("heis_chain10_sz0", "j1j2_4x3_sz0", "hubbard_chain6_u4", "tj_3x2",
                 "fermion_complex_11_n4"):
        kind, n, sec, site, f = MODELS[name]
        oh = orc.Ham(orc.Basis(kind, n, sec), site, f())
        e, it, conv, vec = oh.lanczos(300, 1e-13, want_vec=True)
        ev = np.linalg.eigvalsh(_dense(oh))
        assert abs(e - ev[0]) < 1e-10, name
        M = _dense(oh)
        assert np.linalg.norm(M @ vec - e * vec) < 1e-5


def test_davidson_matches_dense(orc):
    for name, k in (("heis_chain10_sz0", 3), ("tj_3x2", 2), ("hubbard_chain6_u4", 3)):
        kind, n, sec, site, f = MODELS[name]
        oh = orc.Ham(orc.Basis(kind, n, sec), site, f())
        ev, _, it, conv = oh.davidson(k, 20, 1e-8)
        assert conv
        assert np.abs(ev - np.linalg.eigvalsh(_dense(oh))[:k]).max() < 1e-10


def test_ftlm_exact_krylov_pin(orc, lat):
    """With M >= d the Lanczos run from r spans r's whole Krylov space, so the
    FTLM estimate equals (1/R) sum_r <r|exp(-beta H)|r> exactly; r are the
    reference's normal(0,1) vectors (imaginary part drawn first)."""
    n = 6
    os_ = lat.heisenberg_chain(n)
    oh = orc.Ham(orc.Basis("spin", n), "spin", os_)
    d = oh.dim
    M = _dense(oh)
    w, U = np.linalg.eigh(M)
    R = 5
    stream = orc.normal_stream(R * d, 42)
    betas = np.array([0.3, 1.0, 3.0])
    Z, E, Cv = oh.ftlm(betas, R, d + 5)
    for b, beta in enumerate(betas):
        zs = es = e2s = 0.0
        for r in range(R):
            v = stream[r * d:(r + 1) * d]
            c = np.abs(U.conj().T @ v) ** 2
            wt = c * np.exp(-beta * w)
            zs += wt.sum()
            es += (w * wt).sum()
            e2s += (w * w * wt).sum()
        z, e, e2 = zs / R, es / R, e2s / R
        assert abs(Z[b] - z) <= 1e-10 * z
        assert abs(E[b] - e / z) <= 1e-10 * abs(e / z) + 1e-12
        assert abs(Cv[b] - beta ** 2 * (e2 / z - (e / z) ** 2)) <= 1e-9 * max(1.0, abs(Cv[b]))


def test_ftlm_two_level_factor_two(orc, qk):
    """tests/test_ftlm.cpp:30-33: H = Sz0 Sz1 on 2 sites, Tr exp(-H) =
    4 cosh(1/4).  Complex normal vectors have E|r|^2 = 2d, so the estimate
    converges to 2 Tr exp(-beta H)."""
    os_ = qk.OpSum()
    os_.add_term(1.0, "Sz", 0, "Sz", 1)
    oh = orc.Ham(orc.Basis("spin", 2), "spin", os_)
    Z, E, _ = oh.ftlm(np.array([1.0]), 4000, 10)
    exact = 4 * math.cosh(0.25)
    assert abs(Z[0] / (2 * exact) - 1) < 0.05
    assert abs(E[0] - (-0.25 * math.tanh(0.25))) < 0.02


def test_continued_fraction_pins(orc, lat):
    eta = 0.05
    # exact eigenvector start (fully polarised chain, H phi = N/4 phi with no
    # rounding): one level, beta_0 = 0 < 1e-15, a Lorentzian of weight |phi|^2
    n = 8
    full = orc.Ham(orc.Basis("spin", n), "spin", lat.heisenberg_chain(n))
    phi = np.zeros(full.dim, complex)
    phi[-1] = 2.0
    a, b, nrm = full.cf(phi, 20)
    assert len(a) == 1 and a[0] == n / 4 and len(b) == 0 and nrm == 2.0
    for om in (0.0, 0.3, 1.7):
        s = orc.spectral(a, b, nrm, om, -1.0, eta)
        lor = 4.0 * eta / math.pi / ((om - 1.0 - n / 4) ** 2 + eta ** 2)
        assert abs(s - lor) <= 1e-12 * lor
    kind, n, sec, site, f = MODELS["heis_chain10_sz0"]
    oh = orc.Ham(orc.Basis(kind, n, sec), site, f())
    M = _dense(oh)
    w, U = np.linalg.eigh(M)
    e0 = w[0]
    # complete continued fraction = resolvent, and the sum rule
    rng = np.random.default_rng(3)
    phi = rng.standard_normal(oh.dim) + 1j * rng.standard_normal(oh.dim)
    a, b, nrm = oh.cf(phi, oh.dim + 10)
    wts = np.abs(U.conj().T @ phi) ** 2
    for om in np.linspace(-1, 6, 15):
        s = orc.spectral(a, b, nrm, om, e0, eta)
        s_exact = -1 / math.pi * np.imag(np.sum(wts / (om + e0 + 1j * eta - w)))
        assert abs(s - s_exact) <= 1e-8 * max(abs(s_exact), 1e-6)
    grid = np.linspace(w[0] - e0 - 60, w[-1] - e0 + 60, 120001)
    vals = np.array([orc.spectral(a, b, nrm, om, e0, eta) for om in grid[::40]])
    integral = np.trapezoid(vals, grid[::40])
    assert abs(integral / nrm ** 2 - 1) < 2e-3


def test_q_zero_structure_factor_vanishes(orc, lat):
    """q = 0 in the Sz = 0 sector: sum_j Sz_j psi = 0 exactly, so the
    reference returns no coefficients (dynamics.cpp:26)."""
    n = 8
    ob = orc.Basis("spin", n, {"use_sz": True, "sz2": 0})
    oh = orc.Ham(ob, "spin", lat.heisenberg_chain(n))
    st = ob.states()
    sz_tot = np.array([sum(((int(s) >> j) & 1) - 0.5 for j in range(n)) for s in st])
    psi = np.ones(ob.size, complex)
    a, b, nrm = oh.cf(sz_tot * psi, 10)
    assert len(a) == 0 and nrm == 0.0


def test_basis_from_states_equals_enumeration(orc, lat):
    """The explicit-list constructor (for states wider than 32 bits) gives the
    same basis and the same CSR as the enumerating one on sectors both reach."""
    import itertools
    n = 8
    b_enum = orc.Basis("hubbard", n, {"use_nup": True, "nup": 3, "use_ndn": True, "ndn": 2})
    states = sorted(
        sum(1 << (2 * i) for i in up) | sum(1 << (2 * i + 1) for i in dn)
        for up in itertools.combinations(range(n), 3) for dn in itertools.combinations(range(n), 2))
    b_list = orc.Basis.from_states("hubbard", n, states)
    assert np.array_equal(b_enum.states(), b_list.states())
    ops = lat.hubbard_model(lat.chain_bonds(n), n, 1.0, 4.0)
    a = orc.Ham(b_enum, "hubbard", ops).to_scipy()
    c = orc.Ham(b_list, "hubbard", ops).to_scipy()
    assert (a != c).nnz == 0
    with pytest.raises(ValueError):
        orc.Basis.from_states("hubbard", n, [5, 3])
