"""Pins the closed forms of tests/closed_forms.py against the oracle: the
Bethe-ansatz ground-state energy of the periodic Heisenberg chain equals
the oracle's Lanczos (reference lanczos.cpp:101-202) and a dense
diagonalisation of the oracle's CSR for N = 4..14; and the exact small
cases -1.5 (N = 2) and -2 (N = 4, reference bindings/julia/test/runtests.jl:122)."""
import numpy as np
import pytest

from closed_forms import bethe_ground_energy


def test_exact_small_chains():
    assert bethe_ground_energy(2) == pytest.approx(-1.5, abs=1e-14)
    assert bethe_ground_energy(4) == pytest.approx(-2.0, abs=1e-14)
    # N = 6: the exact ring energy -(2 + sqrt(13)) / 2
    assert bethe_ground_energy(6) == pytest.approx(-(2 + np.sqrt(13)) / 2, abs=1e-13)


@pytest.mark.parametrize("n", [8, 10, 12, 14])
def test_bethe_matches_oracle(orc, lat, n):
    oh = orc.Ham(orc.Basis("spin", n, {"use_sz": True, "sz2": 0}), "spin", lat.heisenberg_chain(n))
    e_dense = np.linalg.eigvalsh(oh.to_scipy().toarray())[0] if n <= 12 else None
    e_lanczos = oh.lanczos(300, 1e-13)[0]
    e_ba = bethe_ground_energy(n)
    assert abs(e_ba - e_lanczos) < 1e-10
    if e_dense is not None:
        assert abs(e_ba - e_dense) < 1e-11
