"""Lanczos with device-resident step indices (reference
src/solvers/lanczos.cpp:101-202): steps queued ahead of the host's stopping
test or replayed from one captured graph, checkpoints of the stored basis,
and the second Gram-Schmidt pass.  A resumed run must reproduce the
uninterrupted run bit for bit (energy, iteration count, Ritz vector), and
the graph and direct launch paths must agree bit for bit."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _chain(qk, lat, n, J=1.0):
    return qk.MatrixFreeHamiltonian(qk.SpinHalfBasis(n, sz=0), qk.SpinHalfSite(),
                                    lat.heisenberg_chain(n, j=J))


class _Mode:
    """QKRYLOV_LANCZOS_GRAPH=1 forces the captured-graph path, =0 the direct launches."""

    def __init__(self, value):
        self.value = value

    def __enter__(self):
        self.old = os.environ.get("QKRYLOV_LANCZOS_GRAPH")
        os.environ["QKRYLOV_LANCZOS_GRAPH"] = self.value

    def __exit__(self, *a):
        if self.old is None:
            del os.environ["QKRYLOV_LANCZOS_GRAPH"]
        else:
            os.environ["QKRYLOV_LANCZOS_GRAPH"] = self.old


@pytest.mark.parametrize("mode,n", [("1", 14), ("0", 18)])
def test_checkpoint_resume_reproduces_run(qk, lat, tmp_path, mode, n):
    H = _chain(qk, lat, n)
    with _Mode(mode):
        print(f"[lanczos ckpt] mode={mode} N={n} d={H.dimension}", flush=True)
        r0 = qk.lanczos_ground_state(H, 200, 1e-12)
        assert r0.converged and r0.iterations > 30
        a = tmp_path / "a.ckpt"
        b = tmp_path / "b.ckpt"
        r1 = qk.lanczos_checkpointed(H, 200, 1e-12, save_path=a, stop_after=10)
        assert r1.iterations == 10 and not r1.converged and r1.eigenvector is None
        # resume, checkpoint every 7 steps, stop again at 25
        r2 = qk.lanczos_checkpointed(H, 200, 1e-12, save_path=b, save_every=7, stop_after=25,
                                     resume_path=a)
        assert r2.iterations == 25 and not r2.converged
        r3 = qk.lanczos_checkpointed(H, 200, 1e-12, resume_path=b)
    print(f"[lanczos ckpt] E0={r0.energy:.15f} it={r0.iterations} resumed E0={r3.energy:.15f} "
          f"it={r3.iterations}", flush=True)
    assert r3.energy == r0.energy
    assert r3.iterations == r0.iterations and r3.converged == r0.converged
    assert np.array_equal(r3.eigenvector, r0.eigenvector)


def test_graph_and_direct_paths_agree(qk, lat):
    H = _chain(qk, lat, 16)
    with _Mode("1"):
        rg = qk.lanczos_ground_state(H, 200, 1e-12)
    with _Mode("0"):
        rd = qk.lanczos_ground_state(H, 200, 1e-12)
    assert rg.energy == rd.energy and rg.iterations == rd.iterations
    assert np.array_equal(rg.eigenvector, rd.eigenvector)
    # a forced step count (tol 0): the queue runs to maxiter without a host test per step
    with _Mode("1"):
        fg = qk.lanczos_ground_state(H, 120, 0.0)
    with _Mode("0"):
        fd = qk.lanczos_ground_state(H, 120, 0.0)
    assert fg.iterations == fd.iterations == 120 and fg.energy == fd.energy


def test_resume_rejects_other_hamiltonian(qk, lat, tmp_path):
    H = _chain(qk, lat, 12)
    p = tmp_path / "c.ckpt"
    qk.lanczos_checkpointed(H, 200, 1e-12, save_path=p, stop_after=5)
    H2 = _chain(qk, lat, 12, J=1.5)
    with pytest.raises(RuntimeError, match="different Hamiltonian"):
        qk.lanczos_checkpointed(H2, 200, 1e-12, resume_path=p)
    # a damaged checkpoint is refused
    raw = bytearray(p.read_bytes())
    raw[len(raw) // 2] ^= 1
    (tmp_path / "bad.ckpt").write_bytes(bytes(raw))
    with pytest.raises(RuntimeError, match="checksum"):
        qk.lanczos_checkpointed(H, 200, 1e-12, resume_path=tmp_path / "bad.ckpt")


@pytest.mark.parametrize("mode", ["1", "0"])
def test_second_gram_schmidt_pass(qk, orc, lat, mode):
    """N=8 chain, Sz=0 (d = 70), tol 0 and maxiter = d: the random start
    reaches every eigenvalue, so the Krylov space is exhausted after 68
    steps.  There the three-term residual is rounding noise lying mostly in
    the stored basis, the first projection removes more than half of it and
    the second pass (the DGKS test) must run; then beta falls below the
    reference's 1e-15 breakdown threshold (lanczos.cpp:155-158).  A CPU
    restatement of this algorithm on the oracle's matrix shows the same
    (second pass at step 67, breakdown at step 67, for three start
    vectors).  The energy must be the dense ground energy and the oracle's."""
    os_ = lat.heisenberg_chain(8)
    H = qk.MatrixFreeHamiltonian(qk.SpinHalfBasis(8, sz=0), qk.SpinHalfSite(), os_)
    oh = orc.Ham(orc.Basis("spin", 8, {"use_sz": True, "sz2": 0}), "spin", os_)
    assert H.dimension == 70
    e_dense = np.linalg.eigvalsh(oh.to_scipy().toarray())[0]
    with _Mode(mode):
        r = qk.lanczos_ground_state(H, 70, 0.0)
    e_orc = oh.lanczos(70, 0.0)[0]
    print(f"[lanczos dgks] mode={mode} it={r.iterations} converged={r.converged} "
          f"second passes={r.second_passes} E0={r.energy:.15f} dense={e_dense:.15f}", flush=True)
    assert r.second_passes >= 1
    assert abs(r.energy - e_dense) < 1e-10 and abs(r.energy - e_orc) < 1e-10
    assert np.linalg.norm(H.apply(r.eigenvector) - r.energy * r.eigenvector) < 1e-8
