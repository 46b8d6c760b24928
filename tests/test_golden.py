"""The oracle against data the reference produced or its tests hold
(tests/golden/): known answers of the reference's tests, and the CSR and
basis states of four small models built by the reference's own basis, site
and OpSum classes (tests/golden/make_golden.py).  CPU only; these pins hold
where oracle/_ref is not built."""
import json
from pathlib import Path

import numpy as np
import pytest

from models import MODELS

GOLD = Path(__file__).resolve().parent / "golden"
FIX = json.loads((GOLD / "reference_fixtures.json").read_text())
CSR = json.loads((GOLD / "ref_csr.json").read_text())


@pytest.mark.parametrize("case", FIX["basis_dims"], ids=lambda c: c["ref"])
def test_basis_dims(orc, case):
    assert orc.Basis(case["kind"], case["nsites"], case["sector"]).size == case["dim"]


@pytest.mark.parametrize("case", FIX["site_actions"], ids=lambda c: c["ref"])
def test_site_actions(orc, case):
    got = orc.site_apply(case["kind"], case["op"], case["site"], case["state"])
    assert got[0] is case["valid"]
    if case["valid"]:
        assert got[1] == case["out"]
        assert got[2] == complex(*case["me"])


@pytest.mark.parametrize("case", FIX["ground_energies"], ids=lambda c: c["ref"])
def test_ground_energies(orc, lat, case):
    n = case["nsites"]
    os_ = lat.heisenberg_chain(n, periodic=case["model"] == "heisenberg_pbc")
    oh = orc.Ham(orc.Basis("spin", n, case["sector"]), "spin", os_)
    assert abs(oh.lanczos(200, 1e-12)[0] - case["energy"]) < case["tol"]


def test_diagonal_fixture(orc, qk):
    case = FIX["diagonal"][0]
    os_ = qk.OpSum()
    for t in case["terms"]:
        os_.add_term(t[0], *t[1:])
    oh = orc.Ham(orc.Basis("spin", case["nsites"]), "spin", os_)
    assert oh.diag()[case["state_index"]] == case["value"]


@pytest.mark.parametrize("name", sorted(CSR))
def test_oracle_csr_equals_reference_output(orc, name):
    g = CSR[name]
    kind, n, sec, site, f = MODELS[name]
    b = orc.Basis(kind, n, sec)
    assert [int(s) for s in b.states()] == g["states"]
    oh = orc.Ham(b, site, f())
    rows, cols, vals = oh.csr()
    assert list(map(int, rows)) == g["row_offsets"]
    assert list(map(int, cols)) == g["cols"]
    gv = np.array([complex(*v) for v in g["vals"]])
    assert np.array_equal(vals, gv)
    gd = np.array([complex(*v) for v in g["diag"]])
    assert np.array_equal(oh.diag(), gd)
