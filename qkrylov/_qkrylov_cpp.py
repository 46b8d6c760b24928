"""The reference's low-level module names (bindings/python/src/
binding_impl.hpp:72-277, registered once with "_FP32" and once with "_FP64"
by binding_fp32.cpp / binding_fp64.cpp) over this library's C ABI.

Only the HIP backend exists here: MatrixFreeHamiltonianHIP_FP32/_FP64 and the
solver functions with the "HIP" backend suffix.  The FP32 names compute in
FP64 and narrow their outputs to complex64 / float32."""
import numpy as np

import qkrylov_amd as _amd


class _Device:
    _backend = "hip"

    def __init__(self, device=0):
        self.id = device if isinstance(device, int) else 0

    @staticmethod
    def is_gpu_build():
        return True

    @staticmethod
    def backend_name():
        return "hip"

    @staticmethod
    def gpu_count():
        return _amd.gpu_count()


class _Sector:
    """Plain read-write fields (reference sector.hpp:10-41)."""

    def __init__(self):
        self.use_sz, self.sz2 = False, 0
        self.use_nup, self.use_ndn, self.nup, self.ndn = False, False, 0, 0
        self.use_n, self.n = False, 0
        self.use_nb, self.nb = False, 0

    def _native(self):
        s = _amd.Sector()
        if self.use_sz:
            s.set_sz(int(self.sz2))
        if self.use_n:
            s.set_n(int(self.n))
        if self.use_nb:
            s.set_nb(int(self.nb))
        if self.use_nup or self.use_ndn:
            # the C ABI sets both species together; a species left free is
            # the whole range, which the native basis reads from its flag
            s.set_hubbard_particles(int(self.nup), int(self.ndn))
            if not (self.use_nup and self.use_ndn):
                raise ValueError("only both particle numbers together are supported")
        return s


def _basis_class(native, narrow):
    class _B:
        def __init__(self, N, sector=None):
            sec = (sector or _Sector())._native()
            self._obj = native(N, sector=sec)

        def size(self):
            return self._obj.size

        def state(self, i):
            return self._obj.state(i)

        def index(self, s):
            return self._obj.index(s)

        def contains(self, s):
            return self._obj.contains(s)

        def nsites(self):
            return self._obj.nsites
    return _B


def _ham_class(dtype):
    class _H:
        def __init__(self, basis, site, ops, device=None):
            b = getattr(basis, "_obj", basis)
            self._obj = _amd.MatrixFreeHamiltonian(b, getattr(site, "_obj", site),
                                                   getattr(ops, "_obj", ops), dtype=dtype)
            self.dtype = dtype

        def apply(self, x):
            return self._obj.apply(x)

        def dimension(self):
            return self._obj.dimension

        def diagonal(self):
            return self._obj.diagonal()
    return _H


def _narrow(dtype, v):
    return v.astype(np.complex64) if dtype == np.float32 else v


def _solvers(dtype):
    def lanczos_ground_state(H, maxiter=200, tol=1e-12):
        r = _amd.lanczos_ground_state(H._obj, maxiter, tol)
        return r.energy, _narrow(dtype, r.eigenvector)

    def davidson_lowest(H, n_eig=1, max_subspace=20, tol=1e-8):
        r = _amd.davidson_lowest(H._obj, n_eig, max_subspace, tol)
        r.eigenvectors = [_narrow(dtype, v) for v in (r.eigenvectors or [])]
        return r

    def continued_fraction_coeffs(H, phi0, n_iter=100):
        r = _amd.continued_fraction_coeffs(H._obj, np.asarray(phi0, dtype=np.complex128), n_iter)
        return r.alphas, r.betas, r.norm_phi0

    def ftlm(H, beta, n_random=50, n_steps=100):
        return _amd.ftlm(H._obj, beta, n_random, n_steps)
    return lanczos_ground_state, davidson_lowest, continued_fraction_coeffs, ftlm


def evaluate_spectral_function(alphas, betas, norm_phi0, omega, E0, eta):
    return _amd.evaluate_spectral_function(alphas, betas, norm_phi0, omega, E0, eta)


for _sfx, _dt in (("_FP32", np.float32), ("_FP64", np.float64)):
    globals()["Device" + _sfx] = type("Device" + _sfx, (_Device,), {})
    globals()["Sector" + _sfx] = type("Sector" + _sfx, (_Sector,), {})
    for _name, _native in (("SpinHalfBasis", _amd.SpinHalfBasis), ("FermionBasis", _amd.FermionBasis),
                           ("HubbardBasis", _amd.HubbardBasis), ("TJBasis", _amd.TJBasis)):
        globals()[_name + _sfx] = type(_name + _sfx, (_basis_class(_native, _dt),), {})
    for _name in ("SpinHalfSite", "FermionSite", "HubbardSite", "TJSite"):
        globals()[_name + _sfx] = getattr(_amd, _name)
    globals()["OpSum" + _sfx] = _amd.OpSum
    globals()["MatrixFreeHamiltonianHIP" + _sfx] = type("MatrixFreeHamiltonianHIP" + _sfx,
                                                       (_ham_class(_dt),), {})
    _lgs, _dav, _cf, _ftlm = _solvers(_dt)
    globals()["lanczos_ground_state_HIP" + _sfx] = _lgs
    globals()["davidson_lowest_HIP" + _sfx] = _dav
    globals()["continued_fraction_coeffs_HIP" + _sfx] = _cf
    globals()["ftlm_HIP" + _sfx] = _ftlm
# the reference wrapper looks the continued-fraction binding up without the
# precision suffix (solvers.py:117)
continued_fraction_coeffs_HIP = globals()["continued_fraction_coeffs_HIP_FP64"]
