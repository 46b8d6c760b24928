"""qkrylov_amd — qkrylov's matrix-free H·x and Krylov solvers on one MI355X.

Host-side mirror of the reference Python package (bindings/python/qkrylov:
basis.py, site.py, operators.py, hamiltonian.py, solvers.py) over the C ABI
of ``libqkrylov.so`` (include/qkrylov/c_api.h).  Every numerical call runs
the HIP kernels; vectors cross the boundary as complex128 numpy arrays.
"""
from __future__ import annotations

import ctypes as ct
import os
import enum
from typing import List, Optional, Sequence, Tuple, Union

import numpy as np

from . import _lib
from ._lib import check, last_error, load

__all__ = [
    "Sector", "Op", "Basis", "SpinHalfBasis", "FermionBasis", "HubbardBasis", "TJBasis",
    "Site", "SpinHalfSite", "FermionSite", "HubbardSite", "TJSite",
    "OpSum", "Sz", "Sp", "Sm", "Sx", "Sy", "CdagUp", "CUp", "CdagDn", "CDn",
    "Nup", "Ndn", "Nupdn", "N", "C", "Cdag", "Id",
    "MatrixFreeHamiltonian", "LanczosResult", "lanczos_ground_state", "lanczos_checkpointed",
    "save_eigenpairs", "load_eigenpairs", "save_result",
    "DavidsonResult", "davidson_lowest", "DynamicsResult", "continued_fraction_coeffs",
    "continued_fraction_block", "evaluate_spectral_function", "FTLMResult", "ftlm",
    "ftlm_sweep", "device_info", "gpu_count", "find_gpu", "Bdag", "B",
    "MLPPreconditioner", "davidson_lowest_ex", "jd_correction", "mlp_features",
    "continued_fraction_block_device", "block_lanczos", "Operator", "mlp_correction",
    "mlp_pair_loss",
]


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _release(obj, destroy: str) -> None:
    """Free obj's library handle; at interpreter exit the module globals may
    already be gone, and then the process releases everything anyway."""
    h = getattr(obj, "_h", None)
    if not h:
        return
    obj._h = None
    try:
        getattr(load(), destroy)(h)
    except Exception:  # noqa: BLE001
        pass


def device_info() -> str:
    buf = ct.create_string_buffer(256)
    check(load().qkrylov_device_info(buf, 256), "device_info")
    return buf.value.decode()


def gpu_count() -> int:
    """Number of visible GPUs (reference __init__.py:37-39)."""
    return max(int(load().qkrylov_device_count()), 0)


def find_gpu() -> Optional[str]:
    """'hip' when a GPU is visible, else None (reference __init__.py:31-35)."""
    return "hip" if gpu_count() > 0 else None


# --------------------------------------------------------------------------- sector
class Sector:
    """Quantum-number selection (reference include/qkrylov/symmetry/sector.hpp)."""

    def __init__(self):
        self._h = load().qkrylov_sector_create()
        if not self._h:
            raise RuntimeError(last_error())

    def __del__(self):
        try:
            _release(self, "qkrylov_sector_destroy")
        except Exception:  # noqa: BLE001  (interpreter exit)
            pass

    def set_sz(self, sz2: int):
        check(load().qkrylov_sector_set_sz(self._h, int(sz2)), "set_sz")
        return self

    def set_hubbard_particles(self, nup: int, ndn: int):
        check(load().qkrylov_sector_set_hubbard_particles(self._h, int(nup), int(ndn)), "set_particles")
        return self

    def set_n(self, n: int):
        check(load().qkrylov_sector_set_n(self._h, int(n)), "set_n")
        return self

    def set_nb(self, nb: int):
        check(load().qkrylov_sector_set_nb(self._h, int(nb)), "set_nb")
        return self

    def _get(self, fn, n):
        vals = [ct.c_int() for _ in range(n)]
        check(fn(self._h, *[ct.byref(v) for v in vals]), "sector get")
        return [v.value for v in vals]

    def get_sz(self):
        sz2, act = self._get(load().qkrylov_sector_get_sz, 2)
        return sz2 if act else None

    def get_n(self):
        n, act = self._get(load().qkrylov_sector_get_n, 2)
        return n if act else None

    def get_nb(self):
        nb, act = self._get(load().qkrylov_sector_get_nb, 2)
        return nb if act else None

    def get_hubbard_particles(self):
        nup, ndn, act = self._get(load().qkrylov_sector_get_hubbard_particles, 3)
        return (nup, ndn) if act else None


# --------------------------------------------------------------------------- bases
class Basis:
    """Ascending-integer basis with closed-form ranking (reference src/basis/*)."""

    _create = None

    def __init__(self, N: int, sector: Optional[Sector] = None):
        sec = sector if sector is not None else Sector()
        self._sector = sec
        self._h = getattr(load(), self._create)(int(N), sec._h)
        if not self._h:
            raise ValueError(last_error())

    def __del__(self):
        try:
            _release(self, "qkrylov_basis_destroy")
        except Exception:  # noqa: BLE001  (interpreter exit)
            pass

    @property
    def size(self) -> int:
        return int(load().qkrylov_basis_dimension(self._h))

    @property
    def nsites(self) -> int:
        return int(load().qkrylov_basis_nsites(self._h))

    def state(self, i: int) -> int:
        if not 0 <= i < self.size:
            raise IndexError("basis state index out of range")
        return int(load().qkrylov_basis_state(self._h, int(i)))

    def index(self, s: int) -> int:
        idx = int(load().qkrylov_basis_index(self._h, int(s)))
        if idx < 0:
            raise KeyError("State not present in basis")
        return idx

    def contains(self, s: int) -> bool:
        return bool(load().qkrylov_basis_contains(self._h, int(s)))

    def states(self, first: int = 0, count: Optional[int] = None) -> np.ndarray:
        n = self.size - first if count is None else count
        out = np.empty(n, dtype=np.uint64)
        check(load().qkrylov_basis_states(self._h, first, n, _ptr(out)), "states")
        return out

    def indices(self, states: Sequence[int]) -> np.ndarray:
        st = np.ascontiguousarray(states, dtype=np.uint64)
        out = np.empty(st.size, dtype=np.int64)
        check(load().qkrylov_basis_indices(self._h, _ptr(st), st.size, _ptr(out)), "indices")
        return out

    def __len__(self):
        return self.size


class SpinHalfBasis(Basis):
    _create = "qkrylov_spinhalf_basis_create"
    site_kind = 0

    def __init__(self, N: int, conserve_sz: bool = False, sz: Optional[float] = None,
                 sector: Optional[Sector] = None, dtype=None):
        if sector is None:
            sector = Sector()
            if sz is not None or conserve_sz:
                sector.set_sz(int(round(2 * (sz or 0))))
        super().__init__(N, sector)


class FermionBasis(Basis):
    _create = "qkrylov_fermion_basis_create"
    site_kind = 1

    def __init__(self, N: int, conserve_n: bool = False, n: int = 0,
                 sector: Optional[Sector] = None, dtype=None):
        if sector is None:
            sector = Sector()
            if conserve_n or n != 0:
                sector.set_n(n)
        super().__init__(N, sector)


class _TwoSpecies(Basis):
    def __init__(self, N: int, conserve_nup: bool = False, nup: int = 0,
                 conserve_ndn: bool = False, ndn: int = 0, sector: Optional[Sector] = None,
                 dtype=None):
        if sector is None:
            sector = Sector()
            lib = load()
            # the reference python wrapper sets each flag on its own
            if conserve_nup or nup != 0 or conserve_ndn or ndn != 0:
                sector.set_hubbard_particles(nup, ndn)
                if not (conserve_nup or nup != 0) or not (conserve_ndn or ndn != 0):
                    raise ValueError("conserve both nup and ndn through this wrapper, or pass a Sector")
        super().__init__(N, sector)


class HubbardBasis(_TwoSpecies):
    _create = "qkrylov_hubbard_basis_create"
    site_kind = 2


class TJBasis(_TwoSpecies):
    _create = "qkrylov_tj_basis_create"
    site_kind = 3


# --------------------------------------------------------------------------- sites
class Site:
    _create = None
    kind = -1

    def __init__(self, dtype=None):
        self._h = getattr(load(), self._create)()

    def __del__(self):
        try:
            _release(self, "qkrylov_site_destroy")
        except Exception:  # noqa: BLE001  (interpreter exit)
            pass


class SpinHalfSite(Site):
    _create, kind = "qkrylov_spinhalf_site_create", 0


class FermionSite(Site):
    _create, kind = "qkrylov_fermion_site_create", 1


class HubbardSite(Site):
    _create, kind = "qkrylov_hubbard_site_create", 2


class TJSite(Site):
    _create, kind = "qkrylov_tj_site_create", 3


def site_for(basis: Basis) -> Site:
    return [SpinHalfSite, FermionSite, HubbardSite, TJSite][basis.site_kind]()


# --------------------------------------------------------------------------- operators
class LocalOpExpr:
    """One operator on one site, e.g. Sz(3) (reference operators.py:25-42)."""

    def __init__(self, name: str, site: int):
        self.name, self.site = name, int(site)

    def __mul__(self, other):
        return TermExpr(1.0, [self]) * other

    def __rmul__(self, other):
        return TermExpr(complex(other), [self])

    def __add__(self, other):
        return TermExpr(1.0, [self]) + other

    def __sub__(self, other):
        return TermExpr(1.0, [self]) - other


class TermExpr:
    def __init__(self, coeff: complex, ops: List[LocalOpExpr]):
        self.coeff, self.ops = complex(coeff), list(ops)

    def __mul__(self, other):
        if isinstance(other, LocalOpExpr):
            return TermExpr(self.coeff, self.ops + [other])
        if isinstance(other, TermExpr):
            return TermExpr(self.coeff * other.coeff, self.ops + other.ops)
        if isinstance(other, OpSumExpr):
            return OpSumExpr([self * t for t in other.terms])
        return TermExpr(self.coeff * complex(other), self.ops)

    def __rmul__(self, other):
        return TermExpr(self.coeff * complex(other), self.ops)

    def __add__(self, other):
        return OpSumExpr([self]) + other

    def __sub__(self, other):
        return OpSumExpr([self]) - other

    def __neg__(self):
        return TermExpr(-self.coeff, self.ops)


class OpSumExpr:
    def __init__(self, terms: List[TermExpr]):
        self.terms = list(terms)

    def _as_terms(self, other):
        if isinstance(other, OpSumExpr):
            return other.terms
        if isinstance(other, TermExpr):
            return [other]
        if isinstance(other, LocalOpExpr):
            return [TermExpr(1.0, [other])]
        raise TypeError(f"cannot combine OpSumExpr with {type(other)}")

    def __add__(self, other):
        return OpSumExpr(self.terms + self._as_terms(other))

    def __sub__(self, other):
        return OpSumExpr(self.terms + [-t for t in self._as_terms(other)])

    def __mul__(self, other):
        return OpSumExpr([t * other for t in self.terms])

    def __rmul__(self, other):
        return OpSumExpr([complex(other) * t for t in self.terms])


class Op(str, enum.Enum):
    """Operator names (reference operators.py:6-22), plus the fermion-site
    names C, Cdag, Id of the reference sites."""
    Sz = "Sz"
    Sp = "Sp"
    Sm = "Sm"
    Sx = "Sx"
    Sy = "Sy"
    CdagUp = "CdagUp"
    CUp = "CUp"
    CdagDn = "CdagDn"
    CDn = "CDn"
    Nup = "Nup"
    Ndn = "Ndn"
    Nupdn = "Nupdn"
    Bdag = "Bdag"
    B = "B"
    N = "N"
    C = "C"
    Cdag = "Cdag"
    Id = "Id"


def _op(name):
    return lambda i: LocalOpExpr(name, i)


Sz, Sp, Sm, Sx, Sy = (_op(n) for n in ("Sz", "Sp", "Sm", "Sx", "Sy"))
CdagUp, CUp, CdagDn, CDn = (_op(n) for n in ("CdagUp", "CUp", "CdagDn", "CDn"))
Nup, Ndn, Nupdn, N, C_, Cdag, Id = (_op(n) for n in ("Nup", "Ndn", "Nupdn", "N", "C", "Cdag", "Id"))
globals()["C"] = C_  # fermion annihilator, named as the reference site op "C"
# boson names of the reference's operator module (operators.py:126-127); no
# site defines them, so a Hamiltonian that uses them is rejected at creation
Bdag, B = _op("Bdag"), _op("B")


class OpSum:
    """Append-only list of (coeff, [(op, site), ...]) (reference opsum.hpp)."""

    def __init__(self, dtype=None):
        self.terms: List[Tuple[complex, List[Tuple[str, int]]]] = []

    def add_term(self, coeff: complex, *ops):
        if len(ops) % 2:
            raise ValueError("add_term(coeff, op1, site1, [op2, site2, ...])")
        facs = [(ops[i].value if isinstance(ops[i], Op) else str(ops[i]), int(ops[i + 1]))
                for i in range(0, len(ops), 2)]
        self.terms.append((complex(coeff), facs))
        return self

    def __iadd__(self, item):
        if isinstance(item, tuple):
            if len(item) < 3 or len(item) % 2 == 0:
                raise ValueError("OpSum += requires (coeff, op1, site1, [op2, site2, ...]) with odd "
                                 "tuple length >= 3")
            return self.add_term(*item)
        if isinstance(item, LocalOpExpr):
            item = TermExpr(1.0, [item])
        if isinstance(item, TermExpr):
            item = OpSumExpr([item])
        if isinstance(item, OpSumExpr):
            for t in item.terms:
                self.terms.append((t.coeff, [(o.name, o.site) for o in t.ops]))
            return self
        raise TypeError(f"cannot add {type(item)} to OpSum")

    def clear(self):
        self.terms.clear()

    @property
    def size(self) -> int:
        return len(self.terms)

    def __len__(self):
        return len(self.terms)

    def _handle(self):
        lib = load()
        h = lib.qkrylov_opsum_create()
        for coeff, facs in self.terms:
            n = len(facs)
            ops = (ct.c_char_p * max(n, 1))(*[f[0].encode() for f in facs])
            sites = (ct.c_int * max(n, 1))(*[f[1] for f in facs])
            code = lib.qkrylov_opsum_add_term_f64(h, coeff.real, coeff.imag, n, ops, sites)
            if code != 0:
                lib.qkrylov_opsum_destroy(h)
                check(code, "add_term")
        return h

    def flat(self):
        """(coeff_re, coeff_im, nfac, ops, sites) arrays in term order."""
        cre = np.array([t[0].real for t in self.terms], dtype=np.float64)
        cim = np.array([t[0].imag for t in self.terms], dtype=np.float64)
        nfac = np.array([len(t[1]) for t in self.terms], dtype=np.int32)
        ops = [f[0] for t in self.terms for f in t[1]]
        sites = np.array([f[1] for t in self.terms for f in t[1]], dtype=np.int32)
        return cre, cim, nfac, ops, sites


# --------------------------------------------------------------------------- Hamiltonian
class Operator:
    """y = O x for an OpSum O taking states of ``basis_in`` to ``basis_out``
    (same site kind, size and sector type; any two sectors), on the GPU in
    FP64: y[b] = sum_a <b|O|a> x[a], factors acting on the ket in list order.
    O need not be Hermitian nor conserve the sector (S+(q), c_k).  No
    reference counterpart: the reference applies Hamiltonians within one
    basis only (its dynamics tests ask for this, tests/test_dynamics.cpp:43-46)."""

    def __init__(self, basis_in: Basis, basis_out: Basis, site: Optional[Site], ops: OpSum):
        if site is None:
            site = site_for(basis_in)
        self.basis_in, self.basis_out, self.site, self.ops = basis_in, basis_out, site, ops
        lib = load()
        oh = ops._handle()
        try:
            self._h = lib.qkrylov_operator_create(basis_in._h, basis_out._h, site._h, oh)
        finally:
            lib.qkrylov_opsum_destroy(oh)
        if not self._h:
            raise RuntimeError(f"Operator create failed: {last_error()}")

    def __del__(self):
        try:
            _release(self, "qkrylov_operator_destroy")
        except Exception:  # noqa: BLE001  (interpreter exit)
            pass

    @property
    def shape(self):
        lib = load()
        return (int(lib.qkrylov_operator_dimension_out(self._h)),
                int(lib.qkrylov_operator_dimension_in(self._h)))

    def apply(self, x: np.ndarray) -> np.ndarray:
        dout, din = self.shape
        x = np.ascontiguousarray(x, dtype=np.complex128)
        if x.shape[-1] != din or x.ndim > 2:
            raise ValueError("input length must equal the input basis dimension")
        nvec = 1 if x.ndim == 1 else x.shape[0]
        y = np.empty(x.shape[:-1] + (dout,), dtype=np.complex128)
        if dout == 0 or nvec == 0:
            return y
        check(load().qkrylov_operator_apply_f64(self._h, _ptr(x) if din else None, _ptr(y), nvec),
              "operator apply")
        return y

    __matmul__ = apply

    def apply_device(self, x_ptr: int, y_ptr: int, nvec: int = 1, ld_in: int = 0, ld_out: int = 0,
                     stream: int = 0) -> None:
        """Device buffers (complex FP64), asynchronous on ``stream``."""
        check(load().qkrylov_operator_apply_device_f64(self._h, x_ptr or None, y_ptr, int(nvec),
                                                       int(ld_in), int(ld_out), stream or None),
              "operator apply")


class MatrixFreeHamiltonian:
    """Matrix-free H on the GPU (reference hamiltonian.py / matrix_free_hamiltonian.hpp).

    ``apply(x)`` computes y[a] = sum_terms conj(amp(a->b)) x[b] in FP64.
    """

    def __init__(self, basis: Basis, site: Optional[Site] = None, ops: Optional[OpSum] = None,
                 device: str = "hip", dtype=np.float64):
        """device: accepted for the reference's signature (hamiltonian.py:25);
        this library always runs on the GPU ("cpu" is refused).  dtype:
        np.float64 (default here) or np.float32, which narrows the outputs of
        apply/diagonal to complex64 as the reference's FP32 build returns;
        the arithmetic is FP64 either way."""
        if ops is None and isinstance(site, OpSum):
            site, ops = None, site
        if site is None:
            site = site_for(basis)
        if str(device).lower() not in ("hip", "gpu", "cuda", "rocm"):
            raise ValueError(f"device {device!r}: this library runs on the GPU only (use 'hip' or 'gpu')")
        if dtype not in (np.float32, np.float64):
            raise ValueError("dtype must be np.float32 or np.float64")
        self.device, self.dtype = str(device), dtype
        self.basis, self.site, self.ops = basis, site, ops
        lib = load()
        oh = ops._handle()
        try:
            self._h = lib.qkrylov_hamiltonian_create(basis._h, site._h, oh)
        finally:
            lib.qkrylov_opsum_destroy(oh)
        if not self._h:
            raise RuntimeError(f"Hamiltonian create failed: {last_error()}")

    def __del__(self):
        try:
            _release(self, "qkrylov_hamiltonian_destroy")
        except Exception:  # noqa: BLE001  (interpreter exit)
            pass

    @property
    def dimension(self) -> int:
        return int(load().qkrylov_hamiltonian_dimension(self._h))

    def info(self) -> dict:
        import json
        buf = ct.create_string_buffer(4096)
        check(load().qkrylov_hamiltonian_info(self._h, buf, 4096), "info")
        return json.loads(buf.value.decode())

    def autotune(self, nvec: int = 1) -> dict:
        """Time the apply-kernel variants for blocks of nvec vectors, keep the
        fastest, and return the timings (ms per apply)."""
        import json
        buf = ct.create_string_buffer(4096)
        check(load().qkrylov_hamiltonian_autotune(self._h, int(nvec), buf, 4096), "autotune")
        return json.loads(buf.value.decode())

    def apply(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.complex128)
        d = self.dimension
        if x.shape[-1] != d or x.ndim > 2:
            raise ValueError("Input array size does not match Hamiltonian dimension")
        nvec = 1 if x.ndim == 1 else x.shape[0]
        y = np.empty_like(x)
        if d == 0 or nvec == 0:
            return y
        check(load().qkrylov_hamiltonian_apply_f64(self._h, _ptr(x), _ptr(y), nvec), "apply")
        return y.astype(np.complex64) if self.dtype == np.float32 else y

    __matmul__ = apply

    def to(self, device: Optional[str] = None, dtype=None) -> "MatrixFreeHamiltonian":
        """A new Hamiltonian on the same basis, site and terms (reference
        hamiltonian.py:111-115)."""
        return MatrixFreeHamiltonian(self.basis, self.site, self.ops,
                                     device=device if device is not None else self.device,
                                     dtype=dtype if dtype is not None else self.dtype)

    def aslinearoperator(self):
        """scipy LinearOperator whose matvec is apply() (reference hamiltonian.py:120-144)."""
        import scipy.sparse.linalg as sla
        cdt = np.complex64 if self.dtype == np.float32 else np.complex128
        return sla.LinearOperator(shape=(self.dimension, self.dimension),
                                  matvec=lambda v: self.apply(np.ravel(v)),
                                  matmat=lambda m: self.apply(np.ascontiguousarray(np.asarray(m).T)).T,
                                  dtype=cdt)

    def to_sparse(self, block: int = 256):
        """Explicit CSR matrix for small bases (reference hamiltonian.py:146-177),
        built from applies to blocks of unit vectors; entries that are exactly
        zero are dropped, as the reference's np.nonzero does."""
        import scipy.sparse as sp
        d = self.dimension
        rows, cols, vals = [], [], []
        for c0 in range(0, d, block):
            nb = min(block, d - c0)
            E = np.zeros((nb, d), dtype=np.complex128)
            E[np.arange(nb), c0 + np.arange(nb)] = 1.0
            Y = np.atleast_2d(self.apply(E))
            r, c = np.nonzero(Y.T)
            rows.append(r)
            cols.append(c + c0)
            vals.append(Y.T[r, c])
        cdt = np.complex64 if self.dtype == np.float32 else np.complex128
        if not rows:
            return sp.csr_matrix((d, d), dtype=cdt)
        return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                             shape=(d, d), dtype=cdt)

    def apply_device(self, x_ptr: int, y_ptr: int, nvec: int = 1, ld: int = 0, stream: int = 0):
        check(load().qkrylov_hamiltonian_apply_device_f64(self._h, x_ptr, y_ptr, nvec, ld,
                                                          stream or None), "apply_device")

    def sz_q_device(self, psi_ptr: int, phi_ptr: int, nq: int, ld: int, stream: int = 0):
        check(load().qkrylov_sz_q_device_f64(self._h, psi_ptr, phi_ptr, nq, ld, stream or None),
              "sz_q")

    def sync(self):
        check(load().qkrylov_hamiltonian_sync(self._h), "sync")

    def diagonal(self) -> np.ndarray:
        out = np.empty(self.dimension, dtype=np.complex128)
        check(load().qkrylov_hamiltonian_diagonal_f64(self._h, _ptr(out)), "diagonal")
        return out.astype(np.complex64) if self.dtype == np.float32 else out

    def __repr__(self):
        return f"MatrixFreeHamiltonian(dim={self.dimension})"


# --------------------------------------------------------------------------- solvers
class LanczosResult:
    def __init__(self, energy, eigenvector, iterations, converged, second_passes=0):
        self.energy, self.eigenvector = energy, eigenvector
        self.iterations, self.converged = iterations, converged
        self.second_passes = second_passes   # steps that needed a second Gram-Schmidt pass

    def __iter__(self):
        yield self.energy
        yield self.eigenvector

    def __repr__(self):
        return (f"LanczosResult(energy={self.energy}, iterations={self.iterations}, "
                f"converged={self.converged})")


def lanczos_ground_state(H: MatrixFreeHamiltonian, maxiter: int = 200, tol: float = 1e-12,
                         start: Optional[np.ndarray] = None, return_state: bool = True):
    d = H.dimension
    res = _lib.LanczosResultF64()
    vec = np.empty(d, dtype=np.complex128) if return_state else None
    st = None if start is None else np.ascontiguousarray(start, dtype=np.complex128)
    check(load().qkrylov_lanczos_ground_state_f64(
        H._h, int(maxiter), float(tol), None if st is None else _ptr(st), ct.byref(res),
        None if vec is None else _ptr(vec)), "lanczos")
    return LanczosResult(res.energy, vec, res.iterations, bool(res.converged), res.second_passes)


def lanczos_checkpointed(H: MatrixFreeHamiltonian, maxiter: int = 200, tol: float = 1e-12,
                         save_path: Optional[str] = None, save_every: int = 0, stop_after: int = 0,
                         resume_path: Optional[str] = None, start: Optional[np.ndarray] = None,
                         return_state: bool = True):
    """lanczos_ground_state with checkpoints: the run's state goes to
    save_path every save_every steps and when it stops at stop_after (> 0:
    end after that many steps, converged False); resume_path continues a
    saved run.  A resumed run gives the same energy and iteration count as
    an uninterrupted one.  (No reference counterpart: the reference keeps
    its Lanczos basis in memory, src/solvers/lanczos.cpp:123-126.)"""
    d = H.dimension
    res = _lib.LanczosResultF64()
    vec = np.empty(d, dtype=np.complex128) if return_state and not stop_after else None
    st = None if start is None else np.ascontiguousarray(start, dtype=np.complex128)
    enc = (lambda p: None if p is None else os.fsencode(os.fspath(p)))
    check(load().qkrylov_lanczos_checkpointed_f64(
        H._h, int(maxiter), float(tol), None if st is None else _ptr(st), enc(save_path),
        int(save_every), int(stop_after), enc(resume_path), ct.byref(res),
        None if vec is None else _ptr(vec)), "lanczos")
    return LanczosResult(res.energy, vec, res.iterations, bool(res.converged), res.second_passes)


def save_eigenpairs(path, values, vectors=None, H: Optional[MatrixFreeHamiltonian] = None):
    """Eigenvalues and (optionally) eigenvectors to a raw-binary file; with H
    the file records H's fingerprint."""
    vals = np.ascontiguousarray(np.atleast_1d(values), dtype=np.float64)
    n = vals.size
    vecs = None
    dim = H.dimension if H is not None else 0
    if vectors is not None:
        vecs = np.ascontiguousarray(np.asarray(vectors, dtype=np.complex128).reshape(n, -1))
        dim = vecs.shape[1]
        if H is not None and dim != H.dimension:
            raise ValueError("eigenvectors do not match the Hamiltonian's dimension")
    fp = int(load().qkrylov_hamiltonian_fingerprint(H._h)) if H is not None else 0
    check(load().qkrylov_eigenpairs_save_f64(os.fsencode(os.fspath(path)), dim, n, _ptr(vals),
                                             None if vecs is None else _ptr(vecs), fp), "eigenpairs")


def load_eigenpairs(path, vectors: bool = True, H: Optional[MatrixFreeHamiltonian] = None):
    """(values, vectors or None) from save_eigenpairs' file; with H, checks
    that the file was written for it."""
    p = os.fsencode(os.fspath(path))
    dim, n, hv, fp = ct.c_uint64(), ct.c_int(), ct.c_int(), ct.c_uint64()
    check(load().qkrylov_eigenpairs_info(p, ct.byref(dim), ct.byref(n), ct.byref(hv), ct.byref(fp)),
          "eigenpairs")
    if H is not None and (fp.value != int(load().qkrylov_hamiltonian_fingerprint(H._h))
                          or dim.value != H.dimension):
        raise RuntimeError("eigenpairs: the file was written for a different Hamiltonian")
    vals = np.empty(max(n.value, 1), dtype=np.float64)
    vecs = np.empty((n.value, dim.value), dtype=np.complex128) if vectors and hv.value else None
    check(load().qkrylov_eigenpairs_load_f64(p, _ptr(vals), None if vecs is None else _ptr(vecs)),
          "eigenpairs")
    return vals[:n.value], vecs


def save_result(path, result, H: Optional[MatrixFreeHamiltonian] = None):
    """A LanczosResult (energy, eigenvector) or DavidsonResult (eigenvalues,
    eigenvectors) to a file, as the reference's examples/hdf5_storage.py
    stores them in HDF5."""
    if isinstance(result, LanczosResult):
        vec = None if result.eigenvector is None else [result.eigenvector]
        save_eigenpairs(path, [result.energy], vec, H)
    elif isinstance(result, DavidsonResult):
        vecs = result.eigenvectors
        save_eigenpairs(path, result.eigenvalues,
                        None if vecs is None or len(vecs) == 0 else np.asarray(vecs), H)
    else:
        raise TypeError("save_result takes a LanczosResult or a DavidsonResult")


def block_lanczos(H: MatrixFreeHamiltonian, block: int = 4, nev: int = 1, maxiter: int = 100,
                  tol: float = 1e-12, return_vectors: bool = False):
    """Lowest nev eigenpairs by block Lanczos with full reorthogonalisation
    (blocks of `block` vectors, projections on FP64 MFMA)."""
    d = H.dimension
    nev = max(1, min(nev, block))
    ev = np.zeros(nev)
    vecs = np.zeros((nev, d), dtype=np.complex128) if return_vectors else None
    it, cv = ct.c_int(), ct.c_int()
    check(load().qkrylov_block_lanczos_f64(H._h, int(block), int(nev), int(maxiter), float(tol),
                                           _ptr(ev), None if vecs is None else _ptr(vecs),
                                           ct.byref(it), ct.byref(cv)), "block lanczos")
    return DavidsonResult(ev, None if vecs is None else list(vecs), it.value, bool(cv.value))


class DavidsonResult:
    def __init__(self, eigenvalues, eigenvectors, iterations, converged):
        self.eigenvalues, self.eigenvectors = eigenvalues, eigenvectors
        self.iterations, self.converged = iterations, converged


def davidson_lowest(H: MatrixFreeHamiltonian, n_eig: int = 1, max_subspace: int = 20,
                    tol: float = 1e-8, return_vectors: bool = True):
    d = H.dimension
    k = min(n_eig, d)
    ev = np.zeros(max(n_eig, 1), dtype=np.float64)
    vecs = np.zeros((max(n_eig, 1), d), dtype=np.complex128) if return_vectors else None
    info = _lib.DavidsonResultC()
    check(load().qkrylov_davidson_lowest_f64(H._h, int(n_eig), int(max_subspace), float(tol),
                                             _ptr(ev), None if vecs is None else _ptr(vecs),
                                             ct.byref(info)), "davidson")
    return DavidsonResult(ev[:k], None if vecs is None else list(vecs[:k]), info.iterations,
                          bool(info.converged))


class MLPPreconditioner:
    """Learned correction for Davidson: t_i = z_i + s1_i w_i + s2_i w2_i.

    z = J r, w = J (H z - diag(H) z) and w2 = J (H w - diag(H) w) with
    J_i = 1/(H_ii - theta); the network (32 -> 128 -> 128 -> 2, ReLU, BF16
    weights, FP32 accumulation on MFMA) gives (a_i, b_i) and
    s1 = -0.5 + 0.5 tanh(a), s2 = 0.5 tanh(b); (s1, s2) = (-1, 1) would be the
    second-order Neumann expansion of (H - theta)^-1 r.  It starts with a zero
    output layer, i.e. as the damped first-order correction z - w/2.  There is
    no counterpart in the reference, whose Davidson uses the Jacobi diagonal
    (src/solvers/davidson.cpp:166-177)."""

    def __init__(self, seed: int = 7):
        self._h = load().qkrylov_mlp_create(int(seed))
        if not self._h:
            raise RuntimeError(f"MLP create failed: {last_error()}")
        self.last_train = None

    def __del__(self):
        try:
            _release(self, "qkrylov_mlp_destroy")
        except Exception:  # noqa: BLE001  (interpreter exit)
            pass

    @staticmethod
    def num_params() -> int:
        return int(load().qkrylov_mlp_num_params())

    def get_params(self) -> np.ndarray:
        out = np.empty(self.num_params(), dtype=np.float32)
        check(load().qkrylov_mlp_get_params(self._h, _ptr(out)), "mlp params")
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.num_params():
            raise ValueError(f"expected {self.num_params()} parameters")
        check(load().qkrylov_mlp_set_params(self._h, _ptr(p)), "mlp params")

    def eval(self, features: np.ndarray, grad_out: Optional[np.ndarray] = None):
        """Network output (n, 2) for rows of 32 features; with grad_out (n, 2)
        also the parameter gradient of sum(grad_out * o)."""
        X = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, 32)
        n = X.shape[0]
        o = np.empty((max(n, 1), 2), dtype=np.float32)
        g = None
        go = None
        if grad_out is not None:
            go = np.ascontiguousarray(grad_out, dtype=np.float32).reshape(-1)
            if go.size != 2 * n:
                raise ValueError("grad_out must have shape (n, 2)")
            g = np.empty(self.num_params(), dtype=np.float32)
        check(load().qkrylov_mlp_eval(self._h, _ptr(X), n, None if go is None else _ptr(go),
                                      _ptr(o), None if g is None else _ptr(g)), "mlp eval")
        return (o[:n], g) if g is not None else o[:n]

    def train(self, H_small: "MatrixFreeHamiltonian", n_eig: int = 8, max_subspace: int = 48,
              max_iter: int = 40, max_pairs: int = 256, epochs: int = 30, batch_pairs: int = 8,
              lr: float = 1e-3):
        """Generate (residual, accurate correction) pairs with a Jacobi Davidson
        run on H_small and train; returns the loss history (before, per epoch)."""
        hist = np.zeros(epochs + 1)
        info = _lib.MlpTrainInfo()
        check(load().qkrylov_mlp_train_f64(self._h, H_small._h, int(n_eig), int(max_subspace),
                                           int(max_iter), int(max_pairs), int(epochs),
                                           int(batch_pairs), float(lr), _ptr(hist),
                                           ct.byref(info)), "mlp train")
        self.last_train = {"pairs": info.pairs, "epochs": info.epochs,
                           "loss_first": info.loss_first, "loss_last": info.loss_last,
                           "solve_rel_res": info.solve_rel_res}
        return hist


def mlp_features(H: MatrixFreeHamiltonian, theta: float) -> np.ndarray:
    """The 32 network inputs of every basis state of H for Ritz value theta."""
    out = np.empty((max(H.dimension, 1), 32), dtype=np.float32)
    check(load().qkrylov_mlp_features(H._h, float(theta), _ptr(out)), "mlp features")
    return out[:H.dimension]


def mlp_correction(net: MLPPreconditioner, H: MatrixFreeHamiltonian, r: np.ndarray, theta: float):
    """The learned correction (t, z, w, w2) of residual r at Ritz value theta."""
    r = np.ascontiguousarray(r, dtype=np.complex128)
    if r.size != H.dimension:
        raise ValueError("residual length must equal the dimension")
    t, z, w, w2 = (np.empty_like(r) for _ in range(4))
    check(load().qkrylov_mlp_correction_f64(net._h, H._h, _ptr(r), float(theta), _ptr(t), _ptr(z),
                                            _ptr(w), _ptr(w2)), "mlp correction")
    return t, z, w, w2


def mlp_pair_loss(H: MatrixFreeHamiltonian, o: np.ndarray, z: np.ndarray, w: np.ndarray,
                  w2: np.ndarray, b: np.ndarray, u: np.ndarray, theta, grad: bool = True):
    """Summed training loss of pairs on H's basis (rows of z, w, w2, b, u; o is
    (pairs, n, 2); theta one value per pair) and, with grad, its gradient
    d(loss / pairs)/d(o)."""
    n = H.dimension
    z = np.ascontiguousarray(np.reshape(z, (-1, n)), dtype=np.complex128)
    P = z.shape[0]
    w, w2, b, u = (np.ascontiguousarray(np.reshape(a, (P, n)), dtype=np.complex128)
                   for a in (w, w2, b, u))
    o = np.ascontiguousarray(np.reshape(o, (P, n, 2)), dtype=np.float32)
    th = np.ascontiguousarray(np.reshape(theta, (P,)), dtype=np.float64)
    g = np.empty_like(o) if grad else None
    L = ct.c_double()
    check(load().qkrylov_mlp_pair_loss_f64(H._h, P, _ptr(o), _ptr(z), _ptr(w), _ptr(w2), _ptr(b),
                                           _ptr(u), _ptr(th), None if g is None else _ptr(g),
                                           ct.byref(L)),
          "mlp pair loss")
    return (L.value, g) if grad else L.value


def jd_correction(H: MatrixFreeHamiltonian, r: np.ndarray, u: np.ndarray, theta: float,
                  max_iter: int = 300, rtol: float = 1e-10):
    """MINRES solution of (I-uu^H)(H-theta)(I-uu^H) t = -r with t orthogonal to u."""
    r = np.ascontiguousarray(r, dtype=np.complex128)
    u = np.ascontiguousarray(u, dtype=np.complex128)
    t = np.empty_like(r)
    rr = ct.c_double()
    check(load().qkrylov_jd_correction_f64(H._h, _ptr(r), _ptr(u), float(theta), int(max_iter),
                                           float(rtol), _ptr(t), ct.byref(rr)), "jd correction")
    return t, rr.value


def davidson_lowest_ex(H: MatrixFreeHamiltonian, n_eig: int = 1, max_subspace: int = 20,
                       tol: float = 1e-8, max_iter: int = 100,
                       preconditioner: Optional[MLPPreconditioner] = None,
                       return_vectors: bool = False):
    """Davidson with an iteration limit and an optional learned preconditioner
    (None: the Jacobi diagonal)."""
    d = H.dimension
    k = min(n_eig, d)
    ev = np.zeros(max(n_eig, 1), dtype=np.float64)
    vecs = np.zeros((max(n_eig, 1), d), dtype=np.complex128) if return_vectors else None
    info = _lib.DavidsonResultC()
    check(load().qkrylov_davidson_lowest_ex_f64(
        H._h, int(n_eig), int(max_subspace), float(tol), int(max_iter),
        None if preconditioner is None else preconditioner._h, _ptr(ev),
        None if vecs is None else _ptr(vecs), ct.byref(info)), "davidson")
    return DavidsonResult(ev[:k], None if vecs is None else list(vecs[:k]), info.iterations,
                          bool(info.converged))


class DynamicsResult:
    def __init__(self, alphas, betas, norm_phi0):
        self.alphas, self.betas, self.norm_phi0 = alphas, betas, norm_phi0


def continued_fraction_block(H: MatrixFreeHamiltonian, phi0: np.ndarray, n_iter: int = 100):
    p = np.ascontiguousarray(np.atleast_2d(phi0), dtype=np.complex128)
    nb = p.shape[0]
    a = np.zeros((nb, n_iter))
    b = np.zeros((nb, n_iter))
    nrm = np.zeros(nb)
    cnt = np.zeros(nb, dtype=np.int32)
    check(load().qkrylov_continued_fraction_block_f64(H._h, _ptr(p), nb, int(n_iter), _ptr(a),
                                                      _ptr(b), _ptr(nrm), _ptr(cnt)), "cf")
    out = []
    for j in range(nb):
        n = int(cnt[j])
        nbeta = n if n == n_iter and b[j, n - 1] >= 1e-15 else max(n - 1, 0)
        out.append(DynamicsResult(a[j, :n].copy(), b[j, :nbeta].copy(), float(nrm[j])))
    return out


def continued_fraction_block_device(H: MatrixFreeHamiltonian, phi0_ptr: int, nblock: int,
                                    n_iter: int = 100):
    """As continued_fraction_block with the start vectors in device memory
    (vector b at phi0_ptr + 16*b*dim bytes)."""
    a = np.zeros((nblock, n_iter))
    b = np.zeros((nblock, n_iter))
    nrm = np.zeros(nblock)
    cnt = np.zeros(nblock, dtype=np.int32)
    check(load().qkrylov_continued_fraction_block_device_f64(
        H._h, int(phi0_ptr), int(nblock), int(n_iter), _ptr(a), _ptr(b), _ptr(nrm), _ptr(cnt)), "cf")
    out = []
    for j in range(nblock):
        n = int(cnt[j])
        nbeta = n if n == n_iter and b[j, n - 1] >= 1e-15 else max(n - 1, 0)
        out.append(DynamicsResult(a[j, :n].copy(), b[j, :nbeta].copy(), float(nrm[j])))
    return out


def continued_fraction_coeffs(H: MatrixFreeHamiltonian, phi0: np.ndarray, n_iter: int = 100):
    return continued_fraction_block(H, phi0, n_iter)[0]


def evaluate_spectral_function(alphas, betas, norm_phi0: float, omega: float, E0: float,
                               eta: float = 0.1) -> float:
    if isinstance(alphas, DynamicsResult):
        # evaluate_spectral_function(cfr, omega, E0, eta) as in the Julia wrapper
        cfr = alphas
        return evaluate_spectral_function(cfr.alphas, cfr.betas, cfr.norm_phi0, betas, norm_phi0,
                                          omega)
    a = np.ascontiguousarray(alphas, dtype=np.float64)
    b = np.ascontiguousarray(betas, dtype=np.float64)
    if a.size == 0:
        return 0.0
    if b.size < a.size - 1:
        raise ValueError("alphas array length must match betas array length or betas array length + 1")
    return float(load().qkrylov_evaluate_spectral_function_f64(_ptr(a), _ptr(b), a.size,
                                                               float(norm_phi0), float(omega),
                                                               float(E0), float(eta)))


class FTLMResult:
    def __init__(self, beta, Z, E, Cv):
        self.beta, self.partition_function = beta, Z
        self.internal_energy, self.specific_heat = E, Cv


def ftlm(H: MatrixFreeHamiltonian, beta: float, n_random: int = 50, n_steps: int = 100):
    r = _lib.FTLMResultF64()
    check(load().qkrylov_ftlm_f64(H._h, float(beta), int(n_random), int(n_steps), ct.byref(r)),
          "ftlm")
    return FTLMResult(r.beta, r.partition_function, r.internal_energy, r.specific_heat)


def ftlm_sweep(Hs: Sequence[MatrixFreeHamiltonian], betas, n_random: int = 50,
               n_steps: int = 100, batched: bool = True, plain_alpha: str = "fused"):
    """Z, E, Cv arrays of shape [len(Hs), len(betas)]; row s equals ftlm(Hs[s], beta).
    batched=False runs one vector at a time: plain_alpha="fused" through the
    batched sweep's apply kernel, alpha from its partial sums per tile (the
    batched sweep's arithmetic per vector); "dot" with alpha from a separate
    pass (the reference loop's structure); "workgroup" through the
    single-vector kernel, alpha from its partial sums per workgroup."""
    modes = {"fused": 1, "dot": 2, "workgroup": 3}
    if plain_alpha not in modes:
        raise ValueError("plain_alpha is 'fused', 'dot' or 'workgroup'")
    b = np.ascontiguousarray(betas, dtype=np.float64)
    ns, nb = len(Hs), b.size
    hs = (ct.c_void_p * ns)(*[h._h for h in Hs])
    Z = np.zeros((ns, nb))
    E = np.zeros((ns, nb))
    Cv = np.zeros((ns, nb))
    check(load().qkrylov_ftlm_sweep_f64(hs, ns, _ptr(b), nb, int(n_random), int(n_steps),
                                        0 if batched else modes[plain_alpha],
                                        _ptr(Z), _ptr(E), _ptr(Cv)), "ftlm_sweep")
    return Z, E, Cv
