"""The reference's own C-API test program (tests/test_c_api.cpp of the
reference), compiled unchanged against include/qkrylov/c_api.h and linked to
qkrylov_amd/libqkrylov.so by oracle/Makefile, run on the GPU as a child
process: its asserts (dimension 6, apply and complex apply agreeing, Lanczos
energy -1.6160254038 within 1e-5 with a unit eigenvector, two Davidson
eigenpairs with |H psi - E psi| < 1e-3, continued fraction of norm 1, a
positive spectral function, FTLM Z > 0) must all hold.  Skipped where the
binary was not built (no reference tree when the oracle was built)."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

BIN = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "test_c_api"


@pytest.mark.skipif(not BIN.exists(), reason="oracle/_ref/test_c_api not built")
def test_reference_c_api_program():
    print(f"[reference c api] running {BIN.name}", flush=True)
    r = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.returncode
    assert "C API tests passed successfully!" in r.stdout
