"""The GPU apply (C ABI) against the reference's own output for four small
models (tests/golden/ref_csr.json, written by tests/golden/make_golden.py
from the reference's basis, site and OpSum classes): basis states and
indices bit-exact, the diagonal bit-exact, y = H x within 1e-12."""
import json
from pathlib import Path

import numpy as np
import pytest

from models import MODELS, make_sector

pytestmark = pytest.mark.gpu

CSR = json.loads((Path(__file__).resolve().parent / "golden" / "ref_csr.json").read_text())


@pytest.mark.parametrize("name", sorted(CSR))
def test_apply_equals_reference_output(qk, name):
    import scipy.sparse as sp
    g = CSR[name]
    kind, n, sec, site, f = MODELS[name]
    mk = {"spin": qk.SpinHalfBasis, "fermion": qk.FermionBasis, "hubbard": qk.HubbardBasis,
          "tj": qk.TJBasis}[kind]
    st = {"spin": qk.SpinHalfSite, "fermion": qk.FermionSite, "hubbard": qk.HubbardSite,
          "tj": qk.TJSite}[site]()
    b = mk(n, sector=make_sector(sec))
    assert [int(s) for s in b.states()] == g["states"]
    assert list(b.indices(g["states"])) == list(range(len(g["states"])))
    H = qk.MatrixFreeHamiltonian(b, st, f())
    d = len(g["states"])
    A = sp.csr_matrix((np.array([complex(*v) for v in g["vals"]]), np.array(g["cols"]),
                       np.array(g["row_offsets"])), shape=(d, d))
    assert np.array_equal(H.diagonal(), np.array([complex(*v) for v in g["diag"]]))
    rng = np.random.default_rng(d)
    x = rng.standard_normal((3, d)) + 1j * rng.standard_normal((3, d))
    y = H.apply(x)
    ref = (A @ x.T).T
    err = np.max(np.abs(y - ref)) / np.max(np.abs(ref))
    print(f"[golden] {name}: d={d} nnz={len(g['cols'])} max rel err {err:.1e}", flush=True)
    assert err < 1e-12
