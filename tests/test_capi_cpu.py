"""C ABI checks that need no GPU: every declared symbol is exported, handle
and error conventions of the reference (src/c_api.cpp), host-side basis
ranking against the oracle, random streams, and the host spectral function."""
import ctypes as ct
import re
from pathlib import Path

import numpy as np
import pytest

from models import make_sector

HEADER = Path(__file__).resolve().parent.parent / "include" / "qkrylov" / "c_api.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    names = re.findall(r"\b(qkrylov_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_every_declared_symbol_is_exported(qk):
    lib = qk.load()
    names = declared_functions()
    assert len(names) >= 43 + 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the reference ABI (include/qkrylov/c_api.h:30-168 of the reference)
    ref_names = ["qkrylov_sector_create", "qkrylov_basis_index", "qkrylov_opsum_add_term_nbody",
                 "qkrylov_hamiltonian_apply_complex", "qkrylov_lanczos_ground_state_complex",
                 "qkrylov_davidson_lowest_complex", "qkrylov_continued_fraction_coeffs_complex",
                 "qkrylov_evaluate_spectral_function", "qkrylov_ftlm"]
    assert all(n in names for n in ref_names)
    assert set(qk._lib.SIGNATURES) == set(names)


def test_null_handles_and_errors(qk):
    lib = qk.load()
    assert lib.qkrylov_sector_set_sz(None, 0) == -1
    assert lib.qkrylov_basis_dimension(None) == 0
    assert lib.qkrylov_basis_nsites(None) == 0
    assert lib.qkrylov_basis_state(None, 0) == 0
    assert lib.qkrylov_basis_index(None, 0) == -1
    assert lib.qkrylov_basis_contains(None, 0) == 0
    assert lib.qkrylov_opsum_clear(None) == -1
    assert lib.qkrylov_hamiltonian_dimension(None) == 0
    assert lib.qkrylov_hamiltonian_create(None, None, None) is None
    lib.qkrylov_sector_destroy(None)
    lib.qkrylov_basis_destroy(None)
    lib.qkrylov_site_destroy(None)
    lib.qkrylov_opsum_destroy(None)
    lib.qkrylov_hamiltonian_destroy(None)
    r = qk._lib.LanczosResultC()
    assert lib.qkrylov_lanczos_ground_state(None, 10, 1e-6, ct.byref(r)) == -1
    assert lib.qkrylov_ftlm(None, 1.0, 1, 1, None) == -1
    assert lib.qkrylov_davidson_lowest_complex(None, 1, 10, 1e-6, None, None, None) == -1
    # bad N -> NULL handle (spinhalf_basis.cpp:19-24)
    assert lib.qkrylov_spinhalf_basis_create(0, None) is None
    assert "1 <= N <= 63" in qk.last_error()
    assert lib.qkrylov_hubbard_basis_create(32, None) is None
    # opsum argument checks (c_api.cpp:358-366)
    o = lib.qkrylov_opsum_create()
    assert lib.qkrylov_opsum_add_term_nbody(o, 1.0, 0.0, 0, None, None) == -1
    ops = (ct.c_char_p * 2)(b"Sz", None)
    sites = (ct.c_int * 2)(0, 1)
    assert lib.qkrylov_opsum_add_term_nbody(o, 1.0, 0.0, 2, ops, sites) == -1
    assert lib.qkrylov_opsum_add_term_2body(o, 1.0, 0.0, b"Sz", 0, None, 1) == -1
    assert lib.qkrylov_opsum_add_term_1body(o, 1.0, 0.0, b"Sz", 0) == 0
    assert lib.qkrylov_opsum_size(o) == 1
    assert lib.qkrylov_opsum_clear(o) == 0 and lib.qkrylov_opsum_size(o) == 0
    lib.qkrylov_opsum_destroy(o)
    # spectral function guard (c_api.cpp:586)
    assert lib.qkrylov_evaluate_spectral_function(None, None, 0, 1.0, 0.5, 0.0, 0.1) == 0.0


def test_unknown_operator_rejected(qk):
    """An op the site does not define throws in the reference (e.g.
    spinhalf_site.cpp:118-120) and qkrylov_hamiltonian_create returns NULL."""
    os_ = qk.OpSum()
    os_.add_term(1.0, "Sz", 0, "Bogus", 1)
    with pytest.raises(RuntimeError, match="Unknown spin operator: Bogus"):
        qk.MatrixFreeHamiltonian(qk.SpinHalfBasis(4), qk.SpinHalfSite(), os_)
    os_ = qk.OpSum()
    os_.add_term(1.0, "Nupdn", 0)
    with pytest.raises(RuntimeError, match="Unknown TJ operator"):
        qk.MatrixFreeHamiltonian(qk.TJBasis(3), qk.TJSite(), os_)


def test_sector_getters(qk):
    s = qk.Sector()
    assert s.get_sz() is None and s.get_n() is None and s.get_hubbard_particles() is None
    s.set_sz(0)
    s.set_hubbard_particles(1, 2)
    s.set_n(2)
    s.set_nb(1)
    assert s.get_sz() == 0 and s.get_hubbard_particles() == (1, 2)
    assert s.get_n() == 2 and s.get_nb() == 1


BASES = [
    ("spin", 4, {"use_sz": True, "sz2": 0}), ("spin", 16, {}), ("spin", 20, {"use_sz": True, "sz2": 0}),
    ("spin", 21, {"use_sz": True, "sz2": -5}), ("spin", 18, {"use_sz": True, "sz2": 18}),
    ("spin", 7, {"use_sz": True, "sz2": 2}), ("fermion", 18, {"use_n": True, "n": 7}),
    ("fermion", 9, {}), ("hubbard", 8, {"use_nup": True, "nup": 4, "use_ndn": True, "ndn": 3}),
    ("hubbard", 5, {}), ("tj", 12, {"use_nup": True, "nup": 5, "use_ndn": True, "ndn": 5}),
    ("tj", 7, {}), ("tj", 9, {"use_nup": True, "nup": 0, "use_ndn": True, "ndn": 9}),
]


@pytest.mark.parametrize("kind,n,sec", BASES)
def test_basis_ranking_matches_oracle(qk, orc, kind, n, sec):
    from models import BASIS
    b = BASIS[kind](n, sector=make_sector(sec))
    o = orc.Basis(kind, n, sec)
    assert b.size == o.size and b.nsites == n
    st = o.states()
    assert np.array_equal(b.states(), st)
    assert np.array_equal(b.indices(st), np.arange(st.size))
    rng = np.random.default_rng(n)
    nbits = 2 * n if kind in ("hubbard", "tj") else n
    probe = rng.integers(0, 1 << nbits, size=2000, dtype=np.uint64)
    got = b.indices(probe)
    want = np.array([o.index(int(s)) for s in probe])
    assert np.array_equal(got, want)
    if st.size:
        assert b.state(st.size - 1) == int(st[-1])
        with pytest.raises(IndexError):
            b.state(st.size)


def test_basis_n28_spot_checks(qk):
    """N = 28, Sz = 0 (d = 40,116,600) without enumerating: rank of a state is
    sum_j C(p_j, j); state(index) inverts it."""
    import math
    b = qk.SpinHalfBasis(28, sz=0)
    assert b.size == math.comb(28, 14)
    rng = np.random.default_rng(5)
    for _ in range(200):
        bits = sorted(rng.choice(28, size=14, replace=False))
        s = sum(1 << int(p) for p in bits)
        idx = sum(math.comb(int(p), j + 1) for j, p in enumerate(bits))
        assert b.index(s) == idx and b.state(idx) == s
    assert b.state(0) == (1 << 14) - 1 and b.state(b.size - 1) == ((1 << 14) - 1) << 14


def test_reference_random_streams(qk, orc):
    lib = qk.load()
    n = 20000
    a = np.empty(n, np.complex128)
    assert lib.qkrylov_reference_start_f64(n, 1234, a.ctypes.data) == 0
    assert np.array_equal(a.view(np.float64), orc.uniform_start(n, 1234).view(np.float64))
    assert lib.qkrylov_reference_normals_f64(n, 42, a.ctypes.data) == 0
    assert np.array_equal(a.view(np.float64), orc.normal_stream(n, 42).view(np.float64))
    # the start vector is converted in chunks of 2^20 elements on several
    # threads: a length across a chunk boundary, another seed
    m = (1 << 20) + 77
    c = np.empty(m, np.complex128)
    assert lib.qkrylov_reference_start_f64(m, 5, c.ctypes.data) == 0
    assert np.array_equal(c.view(np.float64), orc.uniform_start(m, 5).view(np.float64))
    # imaginary part first: element 0's imaginary part is the first draw
    import random
    assert abs(a[0].imag) < 10 and abs(a[0].real) < 10


def test_spectral_function_host(qk, orc):
    rng = np.random.default_rng(1)
    a = rng.standard_normal(30)
    b = np.abs(rng.standard_normal(29)) + 0.1
    for om in np.linspace(-2, 3, 11):
        s = qk.evaluate_spectral_function(a, b, 1.3, om, -0.7, 0.05)
        s_ref = orc.spectral(a, b, 1.3, om, -0.7, 0.05)
        assert abs(s - s_ref) <= 1e-13 * max(1.0, abs(s_ref))
    lib = qk.load()
    af = a.astype(np.float32)
    bf = b.astype(np.float32)
    s32 = lib.qkrylov_evaluate_spectral_function(af.ctypes.data, bf.ctypes.data, a.size, 1.3, 0.5,
                                                 -0.7, 0.05)
    s64 = orc.spectral(af.astype(np.float64), bf.astype(np.float64), float(np.float32(1.3)), 0.5,
                       float(np.float32(-0.7)), float(np.float32(0.05)))
    assert abs(s32 - s64) <= 1e-5 * abs(s64)


def test_hermitian_eigh_host(qk):
    """The host eigensolver of the solvers' projected matrices against
    LAPACK (numpy.linalg.eigh): random, degenerate, diagonal and already
    tridiagonal Hermitian matrices."""
    lib = qk.load()
    rng = np.random.default_rng(21)
    cases = []
    for n in (1, 2, 3, 17, 48, 96):
        a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        cases.append(a + a.conj().T)
    q, _ = np.linalg.qr(rng.standard_normal((30, 30)) + 1j * rng.standard_normal((30, 30)))
    cases.append(q @ np.diag(np.repeat([-2.0, 0.5, 0.5001, 3.0, 3.0, 3.0], 5)) @ q.conj().T)
    cases.append(np.diag(rng.standard_normal(12)).astype(complex))
    t = np.diag(rng.standard_normal(20)) + np.diag(rng.standard_normal(19) * 1j, 1)
    cases.append(t + np.triu(t, 1).conj().T)
    for A in cases:
        n = A.shape[0]
        A = np.ascontiguousarray(A, dtype=np.complex128)
        ev = np.empty(n)
        vecs = np.empty((n, n), dtype=np.complex128)
        assert lib.qkrylov_hermitian_eigh_f64(n, A.ctypes.data, ev.ctypes.data, vecs.ctypes.data) == 0
        ref = np.linalg.eigvalsh(A)
        scale = max(1.0, np.max(np.abs(ref)))
        assert np.max(np.abs(ev - ref)) < 1e-13 * scale * max(1, n / 10), n
        res = np.linalg.norm(A @ vecs.T - vecs.T * ev[None, :], axis=0)
        assert np.max(res) < 1e-12 * scale * max(1, n / 10), n
        assert np.max(np.abs(vecs.conj() @ vecs.T - np.eye(n))) < 1e-12 * max(1, n / 10), n
