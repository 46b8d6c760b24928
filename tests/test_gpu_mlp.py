"""Learned Davidson preconditioner on the GPU: BF16-MFMA forward and backward
kernels against the CPU oracle (oracle/mlp_oracle.py), the features kernel,
the correction terms and the pair loss against the oracle, the MINRES
correction against a dense solve, training, and Davidson with the trained
network against the Jacobi method."""
import time

import numpy as np
import pytest

from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def _net(qk, seed=3):
    net = qk.MLPPreconditioner(11)
    net.set_params(M.init_params(seed))
    return net


def test_forward_matches_oracle(qk):
    net = _net(qk)
    p = net.get_params()
    rng = np.random.default_rng(0)
    for n in (1, 63, 64, 1000):          # ragged tiles of 64 rows
        X = rng.standard_normal((n, 32)).astype(np.float32)
        o = net.eval(X)
        ref, _ = M.forward(p, X)
        scale = np.max(np.abs(ref)) + 1e-3
        print(f"[mlp forward] n={n} max|diff|/scale={np.max(np.abs(o - ref)) / scale:.2e}", flush=True)
        # FP32 accumulation order and rare one-ulp BF16 flips of activations
        assert np.max(np.abs(o - ref)) <= 2e-2 * scale
        assert np.median(np.abs(o - ref)) <= 1e-4 * scale


def test_gradients_match_oracle(qk):
    net = _net(qk, 4)
    p = net.get_params()
    rng = np.random.default_rng(1)
    for n in (50, 777):
        X = rng.standard_normal((n, 32)).astype(np.float32)
        go = rng.standard_normal((n, 2)).astype(np.float32)
        _, g = net.eval(X, go)
        ref = M.backward(p, X, go)
        for name, a, b in (("W1", M.W1, M.B1), ("b1", M.B1, M.W2), ("W2", M.W2, M.B2),
                           ("b2", M.B2, M.W3), ("W3", M.W3, M.B3), ("b3", M.B3, M.NPARAMS)):
            rel = np.linalg.norm(g[a:b] - ref[a:b]) / (np.linalg.norm(ref[a:b]) + 1e-12)
            print(f"[mlp grad] n={n} {name} rel={rel:.2e}", flush=True)
            assert rel < 2e-2, (name, rel)


def test_untrained_network_is_damped_neumann(qk):
    net = qk.MLPPreconditioner(5)
    X = np.random.default_rng(2).standard_normal((300, 32)).astype(np.float32)
    o = net.eval(X)
    assert o.shape == (300, 2) and np.all(o == 0.0)


def _tj_small(qk, orc, lat, lx=3, ly=2, k=2):
    n = lx * ly
    ops = lat.tj_model(lat.square_bonds(lx, ly), 1.0, 0.4)
    sec = qk.Sector().set_hubbard_particles(k, k)
    H = qk.MatrixFreeHamiltonian(qk.TJBasis(n, sector=sec), qk.TJSite(), ops)
    ob = orc.Basis("tj", n, {"use_nup": True, "nup": k, "use_ndn": True, "ndn": k})
    A = orc.Ham(ob, "tj", ops).to_scipy().tocsr()
    return H, A


def test_correction_matches_oracle(qk, orc, lat):
    for lx, ly, k in ((3, 2, 2), (4, 3, 3)):
        H, A = _tj_small(qk, orc, lat, lx, ly, k)
        d = A.shape[0]
        diag = A.diagonal().real
        rng = np.random.default_rng(lx * ly)
        net = _net(qk, 6)
        r = rng.standard_normal(d) + 1j * rng.standard_normal(d)
        for theta in (float(diag.min()) - 0.7, float(np.median(diag))):
            t, z, w, w2 = qk.mlp_correction(net, H, r, theta)
            z_ref, w_ref, w2_ref = M.neumann_terms(diag, theta, r, lambda x: A @ x)
            ez = np.linalg.norm(z - z_ref) / np.linalg.norm(z_ref)
            ew = np.linalg.norm(w - w_ref) / np.linalg.norm(w_ref)
            ew2 = np.linalg.norm(w2 - w2_ref) / np.linalg.norm(w2_ref)
            # the combine on the network's own output is exact arithmetic
            o = net.eval(qk.mlp_features(H, theta))
            et = np.linalg.norm(t - M.combine(o, z, w, w2)) / np.linalg.norm(t)
            # and the network output agrees with the oracle's forward pass
            o_ref, _ = M.forward(net.get_params(), qk.mlp_features(H, theta))
            eo = np.linalg.norm(t - M.combine(o_ref, z_ref, w_ref, w2_ref)) / np.linalg.norm(t)
            print(f"[mlp correction] {lx}x{ly} d={d} theta={theta:.3f} z {ez:.1e} w {ew:.1e} "
                  f"w2 {ew2:.1e} t {et:.1e} t(oracle net) {eo:.1e}", flush=True)
            assert ez < 1e-12 and ew < 1e-12 and ew2 < 1e-12 and et < 1e-12
            assert eo < 1e-2


def test_pair_loss_matches_oracle(qk, orc, lat):
    for lx, ly, k in ((3, 2, 2), (4, 3, 3)):
        H, A = _tj_small(qk, orc, lat, lx, ly, k)
        n = A.shape[0]
        rng = np.random.default_rng(12)
        P = 3
        c = lambda: rng.standard_normal((P, n)) + 1j * rng.standard_normal((P, n))
        z, w, w2, u, b = c(), c(), c(), c(), c()
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        b /= np.linalg.norm(b, axis=1, keepdims=True)
        theta = A.diagonal().real.min() - np.array([0.3, 1.1, 0.05])
        o = (rng.standard_normal((P, n, 2)) * 0.7).astype(np.float32)
        L, g = qk.mlp_pair_loss(H, o, z, w, w2, b, u, theta)
        Lref, gref = 0.0, []
        for p in range(P):
            lp, gp = M.pair_loss(o[p], z[p], w[p], w2[p], b[p], u[p], theta[p], lambda x: A @ x)
            Lref += lp
            gref.append(gp / P)
        gref = np.array(gref)
        eg = np.max(np.abs(g - gref)) / np.max(np.abs(gref))
        print(f"[mlp loss] {lx}x{ly} d={n} L={L:.12f} ref={Lref:.12f} grad max rel={eg:.1e}", flush=True)
        assert abs(L - Lref) <= 1e-10 * P
        assert eg < 1e-6      # the gradient is stored in FP32
        assert qk.mlp_pair_loss(H, o, z, w, w2, b, u, theta, grad=False) == L


def _spin_chain(qk, lat, n):
    b = qk.SpinHalfBasis(n, sz=0)
    return qk.MatrixFreeHamiltonian(b, qk.SpinHalfSite(), lat.heisenberg_chain(n)), b


def test_features_match_oracle(qk, orc, lat):
    n = 10
    H, b = _spin_chain(qk, lat, n)
    ob = orc.Basis("spin", n, {"use_sz": True, "sz2": 0})
    oh = orc.Ham(ob, "spin", lat.heisenberg_chain(n))
    A = oh.to_scipy().tocsr()
    diag = A.diagonal().real
    off = A.copy()
    off.setdiag(0)
    off.eliminate_zeros()
    ngroups = H.info()["lo"] + H.info()["hi"] + H.info()["mix"]
    # one group per bond here, so the row's non-zero off-diagonals are its groups
    conn = (np.diff(off.indptr) / ngroups).astype(np.float32)
    occ = M.occupations(ob.states(), n, False)
    for theta in (-4.5, float(diag[3])):
        f = qk.mlp_features(H, theta)
        ref = M.features(diag, occ, conn, theta, n)
        ulp = np.abs(f.view(np.uint32).astype(np.int64) - ref.view(np.uint32).astype(np.int64))
        assert np.max(ulp) <= 1 << 16, np.max(ulp)   # at most one BF16 ulp (tanh rounding)
        assert np.mean(f == ref) > 0.99


def test_jd_correction_matches_dense_solve(qk, orc, lat):
    bonds = lat.square_bonds(3, 2)
    ops = lat.tj_model(bonds, 1.0, 0.4)
    sec = qk.Sector().set_hubbard_particles(2, 2)
    H = qk.MatrixFreeHamiltonian(qk.TJBasis(6, sector=sec), qk.TJSite(), ops)
    ob = orc.Basis("tj", 6, {"use_nup": True, "nup": 2, "use_ndn": True, "ndn": 2})
    A = orc.Ham(ob, "tj", ops).to_scipy().toarray()
    d = A.shape[0]
    rng = np.random.default_rng(3)
    Q, _ = np.linalg.qr(rng.standard_normal((d, 4)) + 1j * rng.standard_normal((d, 4)))
    w, S = np.linalg.eigh(Q.conj().T @ A @ Q)
    u = Q @ S[:, 0]
    theta = float(w[0])
    r = A @ u - theta * u
    t, rr = qk.jd_correction(H, r, u, theta, max_iter=600, rtol=1e-13)
    P = np.eye(d) - np.outer(u, u.conj())
    Md = P @ (A - theta * np.eye(d)) @ P + np.outer(u, u.conj())
    t_ref = np.linalg.solve(Md, -r)
    print(f"[jd] d={d} minres rel res={rr:.1e} |t-t_ref|/|t_ref|="
          f"{np.linalg.norm(t - t_ref) / np.linalg.norm(t_ref):.2e}", flush=True)
    assert abs(np.vdot(u, t)) < 1e-10 * np.linalg.norm(t)
    assert np.linalg.norm(t - t_ref) <= 1e-8 * np.linalg.norm(t_ref)


def test_training_lowers_loss_and_davidson_agrees(qk, lat):
    def tj(lx, ly, k):
        ops = lat.tj_model(lat.square_bonds(lx, ly), 1.0, 0.4)
        sec = qk.Sector().set_hubbard_particles(k, k)
        return qk.MatrixFreeHamiltonian(qk.TJBasis(lx * ly, sector=sec), qk.TJSite(), ops)

    Hs = tj(2, 4, 3)     # d = 560
    net = qk.MLPPreconditioner(7)
    t = time.time()
    hist = net.train(Hs, n_eig=4, max_subspace=24, max_iter=20, max_pairs=48, epochs=15,
                     batch_pairs=8, lr=1e-3)
    print(f"[mlp train] {net.last_train} loss {hist[0]:.4f} -> {hist[-1]:.4f} "
          f"in {time.time() - t:.1f} s", flush=True)
    assert net.last_train["pairs"] > 0
    assert net.last_train["solve_rel_res"] < 1e-3   # MINRES, at most 500 iterations per pair
    assert hist[-1] < hist[0]

    # Davidson on the training lattice and on 2x5 (d = 3,150).  Every variant
    # converges and every converged Ritz value is an eigenvalue of H (dense
    # spectrum) to within the residual tolerance.  Which eigenvalues are found can differ: the reference's
    # start (unit vectors on the lowest diagonal entries) spans no component
    # of eigenvectors in some symmetry sectors, H and every diagonal
    # preconditioner keep that so, and such an eigenvector enters only through
    # accumulated rounding (asked for 4, the reference's Jacobi method misses
    # the fourth eigenvalue of 2x5 that way, and which variant misses the
    # fourth of 2x4 depends on rounding).  The lowest three are reached by
    # every variant on both lattices, so that is what is asked for; where the
    # trained network and Jacobi return the same eigenvalues, the network
    # must take fewer iterations.  The 4x4 target lattice is measured by
    # bench.py.
    from oracle import oracle as orc
    compared = compared_untrained = 0
    for (lx, ly, k), tag in (((2, 4, 3), "2x4"), ((2, 5, 4), "2x5")):
        H = tj(lx, ly, k)
        ob = orc.Basis("tj", lx * ly, {"use_nup": True, "nup": k, "use_ndn": True, "ndn": k})
        A = orc.Ham(ob, "tj", lat.tj_model(lat.square_bonds(lx, ly), 1.0, 0.4)).to_scipy().toarray()
        spec = np.linalg.eigvalsh(A)
        res = {}
        for name, pre in (("jacobi", None), ("neumann", qk.MLPPreconditioner(3)), ("mlp", net)):
            t = time.time()
            r = qk.davidson_lowest_ex(H, 3, 24, 1e-8, 400, pre)
            res[name] = r
            ev = np.asarray(r.eigenvalues)
            dist = np.max(np.min(np.abs(ev[:, None] - spec[None, :]), axis=1))
            print(f"[davidson {tag}] {name}: iterations={r.iterations} converged={r.converged} "
                  f"time={time.time() - t:.2f} s ev={ev} lowest={spec[:3]} "
                  f"max distance to the spectrum={dist:.1e}", flush=True)
            assert r.converged, name
            # a unit Ritz vector with residual r has an eigenvalue of the
            # Hermitian H within |r| < tol of its Ritz value
            assert dist <= 1e-8, name
        same = np.allclose(res["mlp"].eigenvalues, res["jacobi"].eigenvalues, rtol=0, atol=2e-8)
        if same:
            compared += 1
            assert res["mlp"].iterations < res["jacobi"].iterations, tag
        # training must pay: fewer iterations than the untrained network
        # (the damped first-order correction z - w/2)
        if np.allclose(res["mlp"].eigenvalues, res["neumann"].eigenvalues, rtol=0, atol=2e-8):
            compared_untrained += 1
            assert res["mlp"].iterations < res["neumann"].iterations, tag
    assert compared >= 1 and compared_untrained >= 1


def test_block_dot_matches_numpy(qk):
    import ctypes as ct
    lib = qk.load()
    rng = np.random.default_rng(5)
    for m, nw, n in ((1, 1, 1), (5, 3, 1001), (37, 16, 20000), (64, 7, 99)):
        V = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
        W = rng.standard_normal((nw, n)) + 1j * rng.standard_normal((nw, n))
        G = np.empty((m, nw), dtype=np.complex128)
        assert lib.qkrylov_block_dot_f64(V.ctypes.data, m, W.ctypes.data, nw, n, G.ctypes.data) == 0
        ref = V.conj() @ W.T
        print(f"[block_dot] m={m} nw={nw} n={n} max rel={np.max(np.abs(G - ref)) / np.max(np.abs(ref)):.1e}",
              flush=True)
        np.testing.assert_allclose(G, ref, rtol=0, atol=1e-13 * np.max(np.abs(ref)) * max(1, np.sqrt(n)))
