"""Writes the golden vectors of tests/golden/ (run here, where the reference
tree is present and oracle/_ref is built from its basis, site and OpSum
sources; see oracle/Makefile).

ref_csr.json: for a few small models, the basis states and the CSR
(row offsets, columns, conj(amp) values, diagonal) that the reference's own
classes produce (reference matrix_free_hamiltonian.cpp:45-91 semantics,
driven by oracle/ref_shim.cpp), so that the oracle and the GPU apply can be
checked against reference output where oracle/_ref is not built.

    python tests/golden/make_golden.py
"""
import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

from models import MODELS  # noqa: E402
from oracle import oracle as orc  # noqa: E402

NAMES = ["heis_obc4_sz0", "tj_3x2", "fermion_complex_11_n4", "hubbard_chain6_u4"]


def main():
    if not orc.ref_available():
        sys.exit("oracle/_ref is not built: the golden vectors come from the reference build")
    out = {}
    for name in NAMES:
        kind, n, sec, site, f = MODELS[name]
        b = orc.Basis(kind, n, sec, ref=True)
        rows, cols, vals, diag = orc.RefCsr(b, site, f()).get()
        out[name] = {
            "kind": kind, "nsites": n, "sector": sec, "site": site,
            "states": [int(s) for s in b.states()],
            "row_offsets": [int(v) for v in rows],
            "cols": [int(v) for v in cols],
            "vals": [[float(v.real), float(v.imag)] for v in vals],
            "diag": [[float(v.real), float(v.imag)] for v in diag],
        }
    (HERE / "ref_csr.json").write_text(json.dumps(out, separators=(",", ":")))
    print({k: (len(v["states"]), len(v["cols"])) for k, v in out.items()})


if __name__ == "__main__":
    main()
