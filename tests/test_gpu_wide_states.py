"""States wider than 32 bits (64-bit state words): bases, H·x on every
kernel path and a Lanczos ground state against the oracle.  The oracle
enumerates bases by looping over all 2^nbits states, as the reference does,
which is out of reach here, so the sector's states are listed by
combinations (ascending) and the oracle builds its CSR on that list
(orc.Basis.from_states, pinned against the enumerating constructor on the
CPU).  Bar: basis states bit-exact, y within 1e-12 relative, energies
within 1e-10."""
import itertools

import numpy as np
import pytest

import qkrylov_amd as qkm
from models import make_sector

pytestmark = pytest.mark.gpu


def _spin_states(n, ndown):
    full = (1 << n) - 1
    return sorted(full ^ sum(1 << i for i in c) for c in itertools.combinations(range(n), ndown))


def _fermion_states(n, k):
    return sorted(sum(1 << i for i in c) for c in itertools.combinations(range(n), k))


def _two_species_states(n, nup, ndn, tj):
    out = []
    for up in itertools.combinations(range(n), nup):
        for dn in itertools.combinations(range(n), ndn):
            if tj and set(up) & set(dn):
                continue
            out.append(sum(1 << (2 * i) for i in up) | sum(1 << (2 * i + 1) for i in dn))
    return sorted(out)


def _models(lat):
    def fermion_ops(n):
        os = lat.spinless_fermion_chain(n, 1.0, 0.5, periodic=True)
        for i in range(n):
            os.add_term(0.3 * i - 0.4, "N", i)
        os.add_term(0.25j, "C", 0, "Cdag", n - 1)
        os.add_term(-0.25j, "C", n - 1, "Cdag", 0)
        return os

    def hubbard_ops(n):
        os = lat.hubbard_model(lat.chain_bonds(n), n, 1.0, 4.0)
        for i in range(n):
            os.add_term(0.1 * i, "Nup", i)
        return os

    return {
        # name: (basis kind, nsites, sector, states, opsum)
        "heis_chain36_2down": ("spin", 36, {"use_sz": True, "sz2": 32}, _spin_states(36, 2),
                               lat.heisenberg_chain(36)),
        "fermion_ring40_n2": ("fermion", 40, {"use_n": True, "n": 2}, _fermion_states(40, 2),
                              fermion_ops(40)),
        "hubbard_chain17_1_2": ("hubbard", 17, {"use_nup": True, "nup": 1, "use_ndn": True, "ndn": 2},
                                _two_species_states(17, 1, 2, False), hubbard_ops(17)),
        "tj_chain18_2_1": ("tj", 18, {"use_nup": True, "nup": 2, "use_ndn": True, "ndn": 1},
                           _two_species_states(18, 2, 1, True),
                           lat.tj_model(lat.chain_bonds(18), 1.0, 0.4)),
    }


BASIS = {"spin": "SpinHalfBasis", "fermion": "FermionBasis", "hubbard": "HubbardBasis", "tj": "TJBasis"}
SITE = {"spin": "SpinHalfSite", "fermion": "FermionSite", "hubbard": "HubbardSite", "tj": "TJSite"}


def _build(qk, orc, lat, name):
    kind, n, sec, states, ops = _models(lat)[name]
    b = getattr(qk, BASIS[kind])(n, sector=make_sector(sec))
    H = qk.MatrixFreeHamiltonian(b, getattr(qk, SITE[kind])(), ops)
    ob = orc.Basis.from_states(kind, n, states)
    oh = orc.Ham(ob, kind, ops)
    return b, H, ob, oh, np.asarray(states, dtype=np.uint64)


NAMES = ["heis_chain36_2down", "fermion_ring40_n2", "hubbard_chain17_1_2", "tj_chain18_2_1"]


@pytest.mark.parametrize("name", NAMES)
def test_wide_basis_states(qk, orc, lat, name):
    b, H, ob, oh, states = _build(qk, orc, lat, name)
    assert b.size == states.size
    got = np.array([b.state(i) for i in range(b.size)], dtype=np.uint64)
    assert np.array_equal(got, states)
    assert H.info()["state32"] == 0


@pytest.mark.parametrize("path", ["auto", "generic", "rows"])
@pytest.mark.parametrize("name", NAMES)
def test_wide_apply_matches_oracle(qk, orc, lat, name, path, monkeypatch):
    print(f"[wide apply] {name} path={path}", flush=True)
    if path != "auto":
        monkeypatch.setenv("QKRYLOV_APPLY", path)
    b, H, ob, oh, states = _build(qk, orc, lat, name)
    info = H.info()
    print(f"[wide apply] {name}: {info}", flush=True)
    if path == "generic":
        assert info["exchange_path"] == 0 and info["row_path"] == 0
    elif path == "rows" or not name.startswith("heis"):
        assert info["exchange_path"] == 0 and info["row_path"] == 1
    else:   # a real two-bit exchange model: the exchange kernels, 64-bit states
        assert info["exchange_path"] == 1
    rng = np.random.default_rng(17)
    X = rng.standard_normal((3, b.size)) + 1j * rng.standard_normal((3, b.size))
    Y = H.apply(X)
    for j in range(3):
        y_ref = oh.apply(X[j])
        scale = np.abs(y_ref).max()
        assert np.abs(Y[j] - y_ref).max() <= 1e-12 * scale
        np.testing.assert_array_equal(H.apply(X[j]), Y[j])
    # the diagonal, term by term in OpSum order, bit-exact
    assert np.array_equal(H.diagonal().view(np.float64), oh.diag().view(np.float64))


@pytest.mark.parametrize("name", ["heis_chain36_2down", "tj_chain18_2_1"])
def test_wide_lanczos_matches_dense(qk, orc, lat, name):
    """Also the block recurrence on 64-bit states (CF with the apply's fused
    alpha on the exchange path), against the oracle's CF."""
    b, H, ob, oh, states = _build(qk, orc, lat, name)
    e_ref = np.linalg.eigvalsh(oh.to_scipy().toarray())[0]
    r = qk.lanczos_ground_state(H, 300, 1e-12)
    print(f"[wide lanczos] {name}: E0={r.energy:.12f} dense={e_ref:.12f} it={r.iterations}", flush=True)
    assert abs(r.energy - e_ref) <= 1e-10 * max(1.0, abs(e_ref))
    rng = np.random.default_rng(8)
    phi = rng.standard_normal((2, b.size)) + 1j * rng.standard_normal((2, b.size))
    cf = qk.continued_fraction_block(H, phi, 20)
    for j in range(2):
        a_ref, b_ref, n_ref = oh.cf(phi[j], 20)
        assert np.abs(np.asarray(cf[j].alphas[:8]) - a_ref[:8]).max() < 1e-10
        assert abs(cf[j].norm_phi0 - n_ref) <= 1e-13 * n_ref


def test_wide_operator_between_sectors(qk, orc, lat):
    """S+(q) from two down spins to one on the N=36 chain (64-bit states,
    the operator kernel's cross-basis ranking), against the oracle's gather
    matrix on the union of the two sectors (the oracle's Hamiltonian is
    square, so both sectors share one listed basis)."""
    n = 36
    q = 2 * np.pi * 5 / n
    sp = qk.OpSum()
    for j in range(n):
        sp.add_term(np.exp(1j * q * j) / np.sqrt(n), "Sp", j)
    s_in, s_out = _spin_states(n, 2), _spin_states(n, 1)
    b_in = qk.SpinHalfBasis(n, sector=make_sector({"use_sz": True, "sz2": 32}))
    b_out = qk.SpinHalfBasis(n, sector=make_sector({"use_sz": True, "sz2": 34}))
    op = qk.Operator(b_in, b_out, qk.SpinHalfSite(), sp)
    union = sorted(set(s_in) | set(s_out))
    ob = orc.Basis.from_states("spin", n, union)
    G = orc.Ham(ob, "spin", sp).to_scipy().tocsr()
    O = G.conj().T.tocsr()
    pos = {s: i for i, s in enumerate(union)}
    rows = np.array([pos[s] for s in s_out])
    cols = np.array([pos[s] for s in s_in])
    ref = O[rows][:, cols].toarray()
    rng = np.random.default_rng(23)
    x = rng.standard_normal((2, len(s_in))) + 1j * rng.standard_normal((2, len(s_in)))
    y = op.apply(x)
    yr = x @ ref.T
    assert np.abs(y - yr).max() <= 1e-12 * np.abs(yr).max()
