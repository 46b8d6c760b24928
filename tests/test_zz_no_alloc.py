"""qkrylov_block_dot_noalloc_f64 (device buffers, scratch in a static device
array, no allocation) against qkrylov_block_dot_f64 (host buffers, allocates)
through the C ABI: the same kernels, so the projections must agree bit for
bit; two calls in a row, the second smaller, and strided vectors."""
import ctypes as ct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _host_ref(lib, V, W):
    m, n = V.shape
    nw = W.shape[0]
    G = np.zeros((m, nw), dtype=np.complex128)
    Vc, Wc = np.ascontiguousarray(V), np.ascontiguousarray(W)
    assert lib.qkrylov_block_dot_f64(Vc.ctypes.data, m, Wc.ctypes.data, nw, n, G.ctypes.data) == 0
    return G


def _noalloc(lib, torch, V, W, ldv=None, ldw=None):
    m, n = V.shape
    nw = W.shape[0]
    ldv = ldv or n
    ldw = ldw or n
    # device copies laid out with the requested strides
    Vd = torch.zeros((m, ldv), dtype=torch.complex128, device="cuda")
    Wd = torch.zeros((nw, ldw), dtype=torch.complex128, device="cuda")
    Vd[:, :n] = torch.from_numpy(V).to("cuda")
    Wd[:, :n] = torch.from_numpy(W).to("cuda")
    Gd = torch.full((m, nw), complex(np.nan, np.nan), dtype=torch.complex128, device="cuda")
    st = torch.cuda.current_stream()
    torch.cuda.synchronize()
    rc = lib.qkrylov_block_dot_noalloc_f64(Vd.data_ptr(), ldv, m, Wd.data_ptr(), ldw, nw, n,
                                           Gd.data_ptr(), ct.c_void_p(st.cuda_stream))
    assert rc == 0, lib.qkrylov_last_error()
    st.synchronize()
    return Gd.cpu().numpy()


def _rand(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_noalloc_equals_block_dot(qk):
    import torch
    lib = qk.load()
    # first call: the widest shape (m = 64, nw = 16), then a smaller one
    for (m, nw, n, seed) in ((64, 16, 300_001, 1), (5, 3, 1_234, 2)):
        print(f"[noalloc] m={m} nw={nw} n={n}", flush=True)
        V, W = _rand((m, n), seed), _rand((nw, n), seed + 10)
        G = _noalloc(lib, torch, V, W)
        G_ref = _host_ref(lib, V, W)
        assert np.array_equal(G.view(np.float64), G_ref.view(np.float64))
        exact = V.conj() @ W.T
        assert np.abs(G - exact).max() <= 1e-12 * np.abs(exact).max()


def test_noalloc_strided_and_empty(qk):
    import torch
    lib = qk.load()
    V, W = _rand((17, 4_000), 3), _rand((9, 4_000), 4)
    G = _noalloc(lib, torch, V, W, ldv=4_100, ldw=4_003)
    assert np.array_equal(G.view(np.float64), _host_ref(lib, V, W).view(np.float64))
    # n = 0: an empty projection is zero
    G0 = _noalloc(lib, torch, _rand((3, 0), 5), _rand((2, 0), 6), ldv=1, ldw=1)
    assert np.array_equal(G0, np.zeros((3, 2)))
    # sizes beyond the limits are refused without touching the GPU
    assert lib.qkrylov_block_dot_noalloc_f64(1, 10, 65, 1, 10, 1, 10, 1, None) == -1
    assert lib.qkrylov_block_dot_noalloc_f64(1, 10, 4, 1, 10, 17, 10, 1, None) == -1
    assert lib.qkrylov_block_dot_noalloc_f64(1, 5, 4, 1, 10, 2, 10, 1, None) == -1
