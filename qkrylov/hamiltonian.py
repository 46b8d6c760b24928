"""MatrixFreeHamiltonian with the reference's signature and defaults
(bindings/python/qkrylov/hamiltonian.py:25): device="cpu", dtype=np.float32.

The reference picks a Kokkos backend from ``device`` (CPU, Threads, Serial,
CUDA, HIP, SYCL).  This library has one backend, the GPU, and every device
string selects it; ``dtype=np.float32`` returns complex64 results as the
reference's FP32 build does, computed in FP64.
"""
import numpy as np

import qkrylov_amd as _amd


class MatrixFreeHamiltonian(_amd.MatrixFreeHamiltonian):
    def __init__(self, basis, site, ops, device: str = "cpu", dtype=np.float32):
        super().__init__(basis, site, ops, device="hip", dtype=dtype)
        self.device = str(device)
        self._backend_suffix = "HIP"
        self._cpp_obj = self

    def to(self, device=None, dtype=None):
        return MatrixFreeHamiltonian(self.basis, self.site, self.ops,
                                     device=self.device if device is None else device,
                                     dtype=self.dtype if dtype is None else dtype)
