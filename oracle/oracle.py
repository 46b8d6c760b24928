"""ORACLE loader — test infrastructure only (tests/, smoke(), bench.py's CPU
baseline).  ctypes access to oracle/liboracle.so (the CPU restatement of the
reference algorithms, qk_oracle.cpp) and, where it was built, to
oracle/_ref/libqkref.so (the reference's own basis/site/OpSum sources).
"""
from __future__ import annotations

import ctypes as ct
import subprocess
from pathlib import Path
from typing import Optional

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "liboracle.so"
REF_SO = HERE / "_ref" / "libqkref.so"

KIND = {"spin": 0, "fermion": 1, "hubbard": 2, "tj": 3}

_orc = None
_ref = None


def build() -> None:
    subprocess.run(["make", "-C", str(HERE)], check=True, capture_output=True)


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def lib():
    global _orc
    if _orc is None:
        if not ORACLE_SO.exists():
            build()
        L = ct.CDLL(str(ORACLE_SO))
        vp, i, d, u64, i64 = ct.c_void_p, ct.c_int, ct.c_double, ct.c_uint64, ct.c_int64
        sig = {
            "orc_last_error": (ct.c_char_p, []),
            "orc_set_mutation": (None, [i]),
            "orc_threads": (i, []),
            "orc_basis_create": (vp, [i] * 10),
            "orc_basis_destroy": (None, [vp]),
            "orc_basis_from_states": (vp, [i, i, vp, u64]),
            "orc_basis_dim": (u64, [vp]),
            "orc_basis_states": (None, [vp, vp]),
            "orc_basis_index": (i64, [vp, u64]),
            "orc_site_apply": (i, [i, ct.c_char_p, i, u64, vp, vp, vp]),
            "orc_ham_create": (vp, [vp, i, i, vp, vp, vp, vp, vp, i64, i64]),
            "orc_ham_destroy": (None, [vp]),
            "orc_ham_nnz": (u64, [vp]),
            "orc_ham_nrows": (u64, [vp]),
            "orc_ham_csr": (None, [vp, vp, vp, vp]),
            "orc_ham_diag": (None, [vp, vp]),
            "orc_ham_apply": (None, [vp, vp, vp]),
            "orc_lanczos": (i, [vp, i, d, vp, vp, vp, vp, vp]),
            "orc_davidson": (i, [vp, i, i, d, vp, vp, vp, vp]),
            "orc_cf": (i, [vp, vp, i, vp, vp, vp]),
            "orc_spectral": (d, [vp, vp, u64, d, d, d, d]),
            "orc_ftlm": (None, [vp, vp, i, i, i, vp, vp, vp]),
            "orc_uniform_start": (None, [u64, ct.c_uint, vp]),
            "orc_normal_stream": (None, [u64, ct.c_uint, vp]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype, f.argtypes = r, a
        _orc = L
    return _orc


def ref_available() -> bool:
    return REF_SO.exists()


def ref_lib():
    global _ref
    if _ref is None:
        if not REF_SO.exists():
            raise RuntimeError("oracle/_ref/libqkref.so was not built (reference tree absent)")
        L = ct.CDLL(str(REF_SO))
        vp, i, u64, i64 = ct.c_void_p, ct.c_int, ct.c_uint64, ct.c_int64
        sig = {
            "ref_basis_create": (vp, [i] * 10),
            "ref_basis_destroy": (None, [vp]),
            "ref_basis_dim": (u64, [vp]),
            "ref_basis_states": (None, [vp, vp]),
            "ref_basis_index": (i64, [vp, u64]),
            "ref_site_apply": (i, [i, ct.c_char_p, i, u64, vp, vp, vp]),
            "ref_csr_create": (vp, [vp, i, i, vp, vp, vp, vp, vp]),
            "ref_csr_destroy": (None, [vp]),
            "ref_csr_nnz": (u64, [vp]),
            "ref_csr_get": (None, [vp, vp, vp, vp, vp]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype, f.argtypes = r, a
        _ref = L
    return _ref


def _sector_args(sector: dict):
    s = sector or {}
    return [int(s.get("use_sz", False)), int(s.get("sz2", 0)), int(s.get("use_nup", False)),
            int(s.get("nup", 0)), int(s.get("use_ndn", False)), int(s.get("ndn", 0)),
            int(s.get("use_n", False)), int(s.get("n", 0))]


def _terms(opsum):
    """(cre, cim, nfac, ops char**, sites) ctypes-ready from a qkrylov_amd.OpSum."""
    cre, cim, nfac, ops, sites = opsum.flat()
    arr = (ct.c_char_p * max(len(ops), 1))(*[o.encode() for o in ops])
    if sites.size == 0:
        sites = np.zeros(1, dtype=np.int32)
    return cre, cim, nfac, arr, sites


class Basis:
    def __init__(self, kind: str, n: int, sector: Optional[dict] = None, ref: bool = False):
        self.L = ref_lib() if ref else lib()
        self.prefix = "ref" if ref else "orc"
        self.kind = KIND[kind]
        self.h = getattr(self.L, f"{self.prefix}_basis_create")(self.kind, n, *_sector_args(sector))
        if not self.h:
            raise ValueError("basis create failed")

    @classmethod
    def from_states(cls, kind: str, n: int, states) -> "Basis":
        """Basis of an explicit ascending state list (states wider than 32
        bits, where enumerating all 2^nbits states is out of reach)."""
        b = cls.__new__(cls)
        b.L, b.prefix, b.kind = lib(), "orc", KIND[kind]
        arr = np.ascontiguousarray(np.asarray(states, dtype=np.uint64))
        b.h = b.L.orc_basis_from_states(b.kind, n, _p(arr), arr.size)
        if not b.h:
            raise ValueError(b.L.orc_last_error().decode())
        return b

    def __del__(self):
        if getattr(self, "h", None):
            getattr(self.L, f"{self.prefix}_basis_destroy")(self.h)
            self.h = None

    @property
    def size(self) -> int:
        return int(getattr(self.L, f"{self.prefix}_basis_dim")(self.h))

    def states(self) -> np.ndarray:
        out = np.empty(self.size, dtype=np.uint64)
        getattr(self.L, f"{self.prefix}_basis_states")(self.h, _p(out))
        return out

    def index(self, s: int) -> int:
        return int(getattr(self.L, f"{self.prefix}_basis_index")(self.h, int(s)))


def site_apply(kind: str, op: str, site: int, state: int, ref: bool = False):
    L = ref_lib() if ref else lib()
    ns, re, im = ct.c_uint64(), ct.c_double(), ct.c_double()
    v = getattr(L, ("ref" if ref else "orc") + "_site_apply")(
        KIND[kind], op.encode(), site, state, ct.byref(ns), ct.byref(re), ct.byref(im))
    if v < 0:
        raise ValueError("unknown operator")
    return bool(v), int(ns.value), complex(re.value, im.value)


class Ham:
    """Host CSR Hamiltonian (reference matrix_free_hamiltonian.cpp:45-91);
    rows [row0, row1) only when given (bounded samples of large bases)."""

    def __init__(self, basis: Basis, site_kind: str, opsum, row0: int = 0, row1: int = -1):
        self.L = lib()
        cre, cim, nfac, ops, sites = _terms(opsum)
        self._keep = (cre, cim, nfac, ops, sites)
        self.h = self.L.orc_ham_create(basis.h, KIND[site_kind], len(cre), _p(cre), _p(cim),
                                       _p(nfac), ops, _p(sites), row0, row1)
        if not self.h:
            raise RuntimeError(self.L.orc_last_error().decode())
        self.dim = basis.size
        self.nrows = int(self.L.orc_ham_nrows(self.h))

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_ham_destroy(self.h)
            self.h = None

    def csr(self):
        nnz = int(self.L.orc_ham_nnz(self.h))
        rows = np.empty(self.nrows + 1, dtype=np.int64)
        cols = np.empty(max(nnz, 1), dtype=np.int64)
        vals = np.empty(max(nnz, 1), dtype=np.complex128)
        self.L.orc_ham_csr(self.h, _p(rows), _p(cols), _p(vals))
        return rows, cols[:nnz], vals[:nnz]

    def diag(self) -> np.ndarray:
        out = np.empty(self.nrows, dtype=np.complex128)
        self.L.orc_ham_diag(self.h, _p(out))
        return out

    def apply(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.empty(self.nrows, dtype=np.complex128)
        self.L.orc_ham_apply(self.h, _p(x), _p(y))
        return y

    def to_scipy(self):
        import scipy.sparse as sp
        rows, cols, vals = self.csr()
        return sp.csr_matrix((vals, cols, rows), shape=(self.nrows, self.dim))

    def lanczos(self, maxiter=200, tol=1e-12, start=None, want_vec=False):
        e, it, cv = ct.c_double(), ct.c_int(), ct.c_int()
        st = None if start is None else np.ascontiguousarray(start, dtype=np.complex128)
        vec = np.empty(self.dim, dtype=np.complex128) if want_vec else None
        rc = self.L.orc_lanczos(self.h, maxiter, tol, None if st is None else _p(st),
                                ct.byref(e), ct.byref(it), ct.byref(cv),
                                None if vec is None else _p(vec))
        if rc != 0:
            raise RuntimeError(self.L.orc_last_error().decode())
        return e.value, it.value, bool(cv.value), vec

    def davidson(self, n_eig=1, max_subspace=20, tol=1e-8, want_vecs=False):
        ev = np.zeros(max(n_eig, 1))
        vecs = np.zeros((max(n_eig, 1), self.dim), dtype=np.complex128) if want_vecs else None
        it, cv = ct.c_int(), ct.c_int()
        n = self.L.orc_davidson(self.h, n_eig, max_subspace, tol, _p(ev),
                                None if vecs is None else _p(vecs), ct.byref(it), ct.byref(cv))
        if n < 0:
            raise RuntimeError(self.L.orc_last_error().decode())
        return ev[:n], (None if vecs is None else vecs[:n]), it.value, bool(cv.value)

    def cf(self, phi0, n_iter):
        p = np.ascontiguousarray(phi0, dtype=np.complex128)
        a = np.zeros(n_iter)
        b = np.zeros(n_iter)
        nrm = ct.c_double()
        n = self.L.orc_cf(self.h, _p(p), n_iter, _p(a), _p(b), ct.byref(nrm))
        nb = n if (n == n_iter and n > 0 and b[n - 1] >= 1e-15) else max(n - 1, 0)
        return a[:n].copy(), b[:nb].copy(), nrm.value

    def ftlm(self, betas, n_random, n_steps):
        bt = np.ascontiguousarray(np.atleast_1d(betas), dtype=np.float64)
        Z, E, Cv = (np.zeros(bt.size) for _ in range(3))
        self.L.orc_ftlm(self.h, _p(bt), bt.size, n_random, n_steps, _p(Z), _p(E), _p(Cv))
        return Z, E, Cv


def spectral(alphas, betas, norm_phi0, omega, E0, eta):
    a = np.ascontiguousarray(alphas, dtype=np.float64)
    b = np.ascontiguousarray(betas if len(betas) else [0.0], dtype=np.float64)
    return lib().orc_spectral(_p(a), _p(b), a.size, norm_phi0, omega, E0, eta)


def uniform_start(n: int, seed: int = 1234) -> np.ndarray:
    out = np.empty(n, dtype=np.complex128)
    lib().orc_uniform_start(n, seed, _p(out))
    return out


def normal_stream(n: int, seed: int = 42) -> np.ndarray:
    out = np.empty(n, dtype=np.complex128)
    lib().orc_normal_stream(n, seed, _p(out))
    return out


class RefCsr:
    """CSR rows built by driving the reference's own Basis/Site/OpSum classes."""

    def __init__(self, basis: Basis, site_kind: str, opsum):
        self.L = ref_lib()
        cre, cim, nfac, ops, sites = _terms(opsum)
        self.h = self.L.ref_csr_create(basis.h, KIND[site_kind], len(cre), _p(cre), _p(cim),
                                       _p(nfac), ops, _p(sites))
        if not self.h:
            raise RuntimeError("reference CSR build failed")
        self.dim = basis.size

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ref_csr_destroy(self.h)
            self.h = None

    def get(self):
        nnz = int(self.L.ref_csr_nnz(self.h))
        rows = np.empty(self.dim + 1, dtype=np.int64)
        cols = np.empty(max(nnz, 1), dtype=np.int64)
        vals = np.empty(max(nnz, 1), dtype=np.complex128)
        diag = np.empty(max(self.dim, 1), dtype=np.complex128)
        self.L.ref_csr_get(self.h, _p(rows), _p(cols), _p(vals), _p(diag))
        return rows, cols[:nnz], vals[:nnz], diag[:self.dim]
