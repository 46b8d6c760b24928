"""H·x on the GPU against the oracle's CSR gather (reference
matrix_free_hamiltonian.cpp:45-91 and 141-168) on the same inputs.

Bar: y within 1e-12 relative (2-norm) and elementwise; diagonal bit-exact;
matrix columns bit-exact for models whose entries come from one term each."""
import ctypes as ct

import numpy as np
import pytest

from models import MODELS, build_pair

pytestmark = pytest.mark.gpu


def _rand(d, seed, nvec=None):
    rng = np.random.default_rng(seed)
    shape = (d,) if nvec is None else (nvec, d)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


EXCHANGE_MODELS = {"heis_chain10_sz0", "heis_obc4_sz0", "j1j2_4x3_sz0", "heis_chain20_sz0"}


@pytest.mark.parametrize("path", ["auto", "generic", "rows"])
@pytest.mark.parametrize("name", sorted(MODELS))
def test_apply_matches_oracle(qk, orc, name, path, monkeypatch):
    """auto: exchange kernel where it applies, else the row-program kernel;
    generic: every group in the group loop; rows: the row-program kernel
    also for exchange models."""
    print(f"[apply] {name} path={path}", flush=True)
    if path != "auto":
        monkeypatch.setenv("QKRYLOV_APPLY", path)
    H, oh, _ = build_pair(orc, name)
    info = H.info()
    if path == "auto" and name in EXCHANGE_MODELS:
        assert info["exchange_path"] == 1
    if path == "generic":
        assert info["exchange_path"] == 0 and info["row_path"] == 0
    if path == "rows":
        assert info["exchange_path"] == 0
        assert info["row_path"] == 1
    assert H.dimension == oh.dim
    x = _rand(H.dimension, 11)
    y = H.apply(x)
    y_ref = oh.apply(x)
    scale = max(np.abs(y_ref).max(), 1e-300)
    assert np.abs(y - y_ref).max() <= 1e-12 * scale
    assert np.linalg.norm(y - y_ref) <= 1e-12 * np.linalg.norm(y_ref) + 1e-300


@pytest.mark.parametrize("name", sorted(EXCHANGE_MODELS))
def test_exchange_symmetric_amplitudes_match_general(qk, orc, name, monkeypatch):
    """Bonds with one amplitude for both directions take it from the bond
    record (symmetric kernel); QKRYLOV_XSYM=0 selects it per lane (general
    kernel).  Both add the same terms in the same order: equal values, and
    both match the oracle."""
    print(f"[apply xsym] {name}", flush=True)
    H, oh, _ = build_pair(orc, name)
    monkeypatch.setenv("QKRYLOV_XSYM", "0")
    Hg, _, _ = build_pair(orc, name)
    monkeypatch.delenv("QKRYLOV_XSYM")
    assert H.info()["xsym"] == 1 and Hg.info()["xsym"] == 0
    x = _rand(H.dimension, 17)
    y, yg = H.apply(x), Hg.apply(x)
    assert np.array_equal(y, yg)
    y_ref = oh.apply(x)
    assert np.abs(y - y_ref).max() <= 1e-12 * max(np.abs(y_ref).max(), 1e-300)


@pytest.mark.parametrize("name", sorted(MODELS))
def test_diagonal_bit_exact(qk, orc, name):
    H, oh, _ = build_pair(orc, name)
    d = H.diagonal()
    d_ref = oh.diag()
    assert np.array_equal(d.view(np.float64), d_ref.view(np.float64)), \
        np.abs(d - d_ref).max()


@pytest.mark.parametrize("name", ["heis_obc4_sz0", "heis_chain10_sz0", "j1j2_4x3_sz0",
                                  "fermion_complex_11_n4", "hubbard_chain6_u4", "tj_3x2",
                                  "outside_bits_8_sz0", "xxz_fields_full9"])
def test_matrix_columns(qk, orc, name):
    """Dense matrix from unit vectors against the oracle CSR."""
    H, oh, _ = build_pair(orc, name)
    d = H.dimension
    assert d <= 1000, "the dense check is meant for small bases (largest here: 924)"
    M = H.apply(np.eye(d, dtype=np.complex128)).T        # column b = H e_b
    M_ref = oh.to_scipy().toarray()
    if name in ("heis_obc4_sz0", "heis_chain10_sz0", "j1j2_4x3_sz0", "hubbard_chain6_u4",
                "tj_3x2", "fermion_complex_11_n4", "outside_bits_8_sz0"):
        # one OpSum term per off-diagonal entry: values must be identical;
        # diagonal sums may be regrouped, compare them to 1 ulp scale
        off = ~np.eye(d, dtype=bool)
        assert np.array_equal(M[off], M_ref[off])
        assert np.allclose(np.diag(M), np.diag(M_ref), rtol=0, atol=1e-15 * np.abs(M_ref).max())
    else:
        assert np.abs(M - M_ref).max() <= 1e-15 * np.abs(M_ref).max()


@pytest.mark.parametrize("name", ["j1j2_4x3_sz0", "tj_4x3", "hubbard_chain8_half", "fermion_complex_11_n4"])
@pytest.mark.parametrize("nvec", [2, 3, 5, 8, 11])
def test_block_apply_equals_single(qk, orc, nvec, name):
    """Blocks of 8, 4, 2 and 1 vectors per launch add each state's terms in
    the same order as the single-vector kernel (exchange kernel for j1j2,
    row-program kernel for the others): identical bits."""
    H, oh, _ = build_pair(orc, name)
    X = _rand(H.dimension, 5, nvec)
    Y = H.apply(X)
    for v in range(nvec):
        assert np.array_equal(Y[v], H.apply(X[v]))


def test_float_abi_matches(qk, orc):
    """Reference float entry points: apply (planes, imag may be NULL),
    apply_complex and diagonal agree with the FP64 path to float rounding
    (reference tests/test_c_api.cpp:66-87)."""
    lib = qk.load()
    H, oh, _ = build_pair(orc, "heis_chain10_sz0")
    d = H.dimension
    x = _rand(d, 3)
    y64 = H.apply(x)
    xr = np.ascontiguousarray(x.real, dtype=np.float32)
    xi = np.ascontiguousarray(x.imag, dtype=np.float32)
    yr = np.zeros(d, np.float32)
    yi = np.zeros(d, np.float32)
    assert lib.qkrylov_hamiltonian_apply(H._h, xr.ctypes.data, xi.ctypes.data, yr.ctypes.data,
                                         yi.ctypes.data) == 0
    xc = np.empty(2 * d, np.float32)
    xc[0::2], xc[1::2] = xr, xi
    yc = np.zeros(2 * d, np.float32)
    assert lib.qkrylov_hamiltonian_apply_complex(H._h, xc.ctypes.data, yc.ctypes.data) == 0
    assert np.array_equal(yc[0::2], yr) and np.array_equal(yc[1::2], yi)
    assert np.abs(yr - y64.real).max() < 1e-5 * np.abs(y64).max()
    # imag plane may be NULL
    yr2 = np.zeros(d, np.float32)
    assert lib.qkrylov_hamiltonian_apply(H._h, xr.ctypes.data, None, yr2.ctypes.data, None) == 0
    y_real_only = H.apply(xr.astype(np.complex128))
    assert np.abs(yr2 - y_real_only.real).max() < 1e-5 * np.abs(y_real_only).max()
    dg = np.zeros(d, np.float32)
    assert lib.qkrylov_hamiltonian_diagonal(H._h, dg.ctypes.data) == 0
    assert np.array_equal(dg, oh.diag().real.astype(np.float32))
    assert lib.qkrylov_hamiltonian_apply(None, xr.ctypes.data, None, yr2.ctypes.data, None) == -1


def test_device_pointer_apply(qk, orc):
    """apply_device_f64 on torch-owned HBM buffers, on torch's stream."""
    torch = pytest.importorskip("torch")
    H, oh, _ = build_pair(orc, "heis_chain20_sz0")
    d = H.dimension
    x = _rand(d, 9, 3)
    xt = torch.from_numpy(x.copy()).to("cuda")
    yt = torch.empty_like(xt)
    stream = torch.cuda.current_stream().cuda_stream
    H.apply_device(xt.data_ptr(), yt.data_ptr(), 3, d, stream)
    torch.cuda.synchronize()
    y = yt.cpu().numpy()
    for v in range(3):
        ref = oh.apply(x[v])
        assert np.abs(y[v] - ref).max() <= 1e-12 * np.abs(ref).max()


def test_hermiticity_and_linearity_n24(qk, orc, lat):
    """Size-independent properties at the 6x4 J1-J2 size (d = 2,704,156):
    <x, H y> = conj(<y, H x>), H(ax + by) = a Hx + b Hy, and a bounded
    sample of oracle rows."""
    print("[apply] j1j2 6x4 properties", flush=True)
    os_ = lat.j1j2_square(6, 4, 1.0, 0.5)
    basis = qk.SpinHalfBasis(24, sz=0)
    H = qk.MatrixFreeHamiltonian(basis, qk.SpinHalfSite(), os_)
    d = H.dimension
    assert d == 2704156
    x, y = _rand(d, 1), _rand(d, 2)
    Hx, Hy = H.apply(x), H.apply(y)
    a, b = 0.3 - 1.1j, 2.0 + 0.5j
    Hxy = H.apply(a * x + b * y)
    assert np.linalg.norm(Hxy - (a * Hx + b * Hy)) <= 1e-12 * np.linalg.norm(Hxy)
    lhs = np.vdot(x, Hy)
    rhs = np.conj(np.vdot(y, Hx))
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs)
    ob = orc.Basis("spin", 24, {"use_sz": True, "sz2": 0})
    for r0 in (0, d // 2, d - 4096):
        part = orc.Ham(ob, "spin", os_, r0, r0 + 4096)
        ref = part.apply(x)
        got = Hx[r0:r0 + 4096]
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("name", ["tj_4x3", "hubbard_chain8_half", "heis_chain20_sz0"])
def test_autotune_keeps_results(qk, orc, name):
    """The autotuner only picks among kernels that compute the same y: after
    tuning for blocks of 4, single and block applies still match the oracle."""
    H, oh, _ = build_pair(orc, name)
    tun = H.autotune(4)
    print(f"[autotune] {name}: {tun}", flush=True)
    assert tun["tuned"] == 1
    X = _rand(H.dimension, 9, 5)
    Y = H.apply(X)
    for j in range(5):
        y_ref = oh.apply(X[j])
        assert np.abs(Y[j] - y_ref).max() <= 1e-12 * np.abs(y_ref).max()
    y1 = H.apply(X[2])
    assert np.abs(y1 - oh.apply(X[2])).max() <= 1e-12 * np.abs(y1).max()
