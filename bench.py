"""Benchmark: y = H·x for the periodic N=28 spin-1/2 Heisenberg chain in the
Sz=0 sector (d = 40,116,600, complex FP64), one GPU.

A step is one application of H to one vector already resident in HBM; the
library's autotuner picks the kernel variant first (untimed).  The JSON line
also carries:
  block     28 vectors per call;
  ftlm      N=20 XXZ chain, all 21 Sz sectors, R=64, M=100, 200 temperatures:
            the batched sweep against the plain per-vector loop (samples/s);
  lanczos   J1-J2 6x4 (d=2,704,156), 300 steps with full reorthogonalisation,
            and the N=16 chain to tol 1e-12 (time to ground state);
  davidson  t-J 4x4 (d=411,840), lowest 8 to 1e-8: Jacobi against the damped
            first-order correction (untrained network) and the learned
            second-order correction trained on 3x4 (iterations, wall time:
            the faster of two runs of each; training timed apart);
  dynamics  N=28 ground state, 28 S^z(q) start vectors as one block, 200
            continued-fraction steps, S(q, w) on 2000 points;
  roofline  of the apply kernel (traffic from the committed rocprofv3 summary);
  cpu_baseline  the oracle's CSR gather SpMV (the reference algorithm) on a
            row sample of the N=28 chain, plus the same CPU path for Lanczos
            (N=16) and FTLM (2 vectors per sector), on the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--skip-ftlm] [--skip-cpu]
                    [--skip-lanczos] [--skip-davidson] [--skip-dynamics]
                    [--apply-only | --ftlm-only | --davidson-only]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
BYTES_PER_STATE = 32           # read x (16 B) + write y (16 B), complex FP64
N_SITES = 28


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profiles(kernel: str, bsize: int):
    """Bytes per launch between the L2 caches and memory, measured by
    rocprofv3 --pmc (L2 read and write requests times their sizes, the sizes
    checked on a copy kernel), if a summary is committed.  Reads served by
    the memory-side cache (MALL) are included: an upper bound on HBM bytes."""
    p = ROOT / "profiles" / "hx_heisenberg_n28_pmc.md"
    if not p.exists():
        return None
    m = re.search(rf"^traffic_bytes_per_launch\[{kernel},b={bsize}\]:\s*([0-9.eE+]+)", p.read_text(),
                  re.M)
    return float(m.group(1)) if m else None


def bench_apply(qk, torch, H, d, nvec, steps, warmup):
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    x = torch.randn((nvec, d), dtype=torch.complex128, device="cuda", generator=g)
    y = torch.empty_like(x)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    for _ in range(warmup):
        H.apply_device(x.data_ptr(), y.data_ptr(), nvec, d, sp)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        H.apply_device(x.data_ptr(), y.data_ptr(), nvec, d, sp)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    # one launch per 8 vectors (block kernel); average launch duration
    launches = (nvec + 7) // 8
    del x, y
    torch.cuda.empty_cache()
    return ms, launches


def cpu_baseline(lat, n_rows=4_000_000, reps=15):
    """Oracle CSR gather SpMV (the reference algorithm, matrix_free_hamiltonian.cpp
    :45-91 and :156-167) on rows [0, n_rows) of the same H, OpenMP on the host."""
    from oracle import oracle as orc
    threads = orc.lib().orc_threads()
    os_ = lat.heisenberg_chain(N_SITES)
    t0 = time.time()
    ob = orc.Basis("spin", N_SITES, {"use_sz": True, "sz2": 0})
    d = ob.size
    part = orc.Ham(ob, "spin", os_, 0, n_rows)
    setup = time.time() - t0
    rng = np.random.default_rng(7)
    x = rng.standard_normal(d) + 1j * rng.standard_normal(d)
    for _ in range(2):
        part.apply(x)
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        part.apply(x)
        ts.append(time.perf_counter() - t)
    med = float(np.median(ts))
    rows_per_s = n_rows / med
    del part
    out = {
        "value": rows_per_s / d,
        "unit": "applies/s",
        "cores": int(threads),
        "kind": "port",
        "sample": (f"rows [0, {n_rows}) of the N=28 chain (d={d}); CSR built by the oracle in "
                   f"{setup:.1f} s; median of {reps} gather SpMVs after 2 warm-ups = "
                   f"{med * 1e3:.1f} ms, scaled to whole applies"),
    }
    # the same CPU path for the solver legs: Lanczos on the N=16 chain (whole
    # run) and FTLM on the N=20 XXZ sectors with 2 of the 64 vectors per sector
    n16 = orc.Ham(orc.Basis("spin", 16, {"use_sz": True, "sz2": 0}), "spin", lat.heisenberg_chain(16))
    t = time.perf_counter()
    e16, it16, _, _ = n16.lanczos(200, 1e-12)
    out["lanczos_n16"] = {"time_to_ground_state_s": time.perf_counter() - t, "E0": e16,
                          "iterations": it16}
    temps = 0.05 * 200.0 ** (np.arange(200) / 199.0)
    t = time.perf_counter()
    nsamp = 0
    for k in range(21):
        oh = orc.Ham(orc.Basis("spin", 20, {"use_sz": True, "sz2": 2 * k - 20}), "spin",
                     lat.xxz_chain(20, 1.0, 1.5))
        oh.ftlm(1.0 / temps, 2, 100)
        nsamp += 2
    dt = time.perf_counter() - t
    out["ftlm"] = {"samples_per_s": nsamp / dt, "sample": "21 sectors x 2 vectors, M=100, 200 "
                   "temperatures, CSR build included", "time_s": dt}
    log(f"cpu: lanczos N=16 {out['lanczos_n16']}, ftlm {out['ftlm']}")
    return out


def bench_ftlm(qk, lat, batched_only=False):
    """FTLM sweep over all 21 Sz sectors of the N=20 XXZ chain: the batched
    sweep against the plain per-vector loops (batched_only: the batched sweep
    alone, for profiling its kernels)."""
    n = 20
    os_ = lat.xxz_chain(n, 1.0, 1.5)
    hs = []
    for k in range(n + 1):
        sec = qk.Sector().set_sz(2 * k - n)
        hs.append(qk.MatrixFreeHamiltonian(qk.SpinHalfBasis(n, sector=sec), qk.SpinHalfSite(), os_))
    temps = 0.05 * 200.0 ** (np.arange(200) / 199.0)
    betas = 1.0 / temps
    R, M = 64, 100
    nsamp = R * len(hs)
    # untimed: the library caches the seed-42 normal stream (R * max dim values,
    # drawn once on the host) and both modes read that same cache
    t = time.perf_counter()
    qk.ftlm_sweep(hs, betas[:1], R, 2, batched=True)
    if batched_only:
        log(f"ftlm warm-up (random stream, kernels): {time.perf_counter() - t:.3f} s")
        t = time.perf_counter()
        qk.ftlm_sweep(hs, betas, R, M, batched=True)
        dt = time.perf_counter() - t
        log(f"ftlm batched: {dt:.3f} s")
        return {"workload": "XXZ chain N=20 (Jxy=1, Jz=1.5), all 21 Sz sectors, R=64, M=100, "
                            "200 temperatures; batched sweep only",
                "samples": nsamp, "samples_per_s_batched": nsamp / dt}
    for how in ("fused", "dot", "workgroup"):
        qk.ftlm_sweep(hs, betas[:1], 1, 2, batched=False, plain_alpha=how)
    log(f"ftlm warm-up (random stream, kernels): {time.perf_counter() - t:.3f} s")
    out = {}
    # plain loops, one vector at a time: plain_tile through the batched
    # sweep's kernel with alpha from its partial sums per tile (the batched
    # sweep's arithmetic per vector); plain_dot with alpha from a separate
    # dot, as the reference loop takes it; plain_workgroup through the
    # single-vector kernel with alpha from its partial sums per workgroup
    for name, batched, how in (("batched", True, "fused"), ("plain_tile", False, "fused"),
                               ("plain_dot", False, "dot"), ("plain_workgroup", False, "workgroup")):
        t = time.perf_counter()
        Z, E, Cv = qk.ftlm_sweep(hs, betas, R, M, batched=batched, plain_alpha=how)
        dt = time.perf_counter() - t
        out[name] = (dt, Z, E, Cv)
        log(f"ftlm {name}: {dt:.3f} s")
    tb, Zb, Eb, Cb = out["batched"]
    fastest = min(("plain_tile", "plain_dot", "plain_workgroup"), key=lambda k: out[k][0])
    tp = out[fastest][0]
    _, Zt_, Et_, Ct_ = out["plain_tile"]

    def reldiff(a, b):
        return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))

    def diff_cv(a, b):
        # Cv = beta^2 (<E^2> - <E>^2) is the difference of two terms of size
        # up to beta^2 <E^2> = Cv + beta^2 <E>^2; rounding moves it by a few
        # ulps of that, so differences are scaled by it
        return float(np.max(np.abs(a - b) / np.maximum(Cb + betas**2 * Eb**2, 1e-300)))

    # combine sectors (thermodynamics of the whole chain)
    Zt = Zb.sum(0)
    Et = (Zb * Eb).sum(0) / Zt
    return {
        "workload": "XXZ chain N=20 (Jxy=1, Jz=1.5), all 21 Sz sectors, R=64, M=100, 200 temperatures",
        "samples": nsamp,
        "samples_per_s_batched": nsamp / tb,
        "samples_per_s_plain": nsamp / tp,
        "plain_fastest": fastest,
        "samples_per_s_plain_tile": nsamp / out["plain_tile"][0],
        "samples_per_s_plain_dot": nsamp / out["plain_dot"][0],
        "samples_per_s_plain_workgroup": nsamp / out["plain_workgroup"][0],
        "speedup": tp / tb,
        "speedup_note": "against the fastest of the three plain loops",
        "max_rel_diff_Z": reldiff(Zb, Zt_),
        "max_rel_diff_E": reldiff(Eb, Et_),
        "max_rel_diff_Cv": reldiff(Cb, Ct_),
        "diff_note": "batched against plain_tile (the same kernel, tiles and sums per vector); the "
                     "other plain loops sum alpha in another order",
        "max_diff_Cv_plain_dot": diff_cv(Cb, out["plain_dot"][3]),
        "max_diff_Cv_plain_workgroup": diff_cv(Cb, out["plain_workgroup"][3]),
        "cv_diff_note": "batched against the other plain loops: |dCv| / (beta^2 <E^2>), the scale of "
                        "the two terms whose difference Cv is",
        "E_per_site_T0.05": float(Et[0] / n),
    }


def bench_lanczos_j1j2(qk, lat):
    """J1-J2 Heisenberg on the 6x4 periodic square lattice (J1=1, J2=0.5, Sz=0,
    d=2,704,156): 300 Lanczos steps with full reorthogonalisation (tol 0),
    ground energy and the residual of the Ritz vector.  Also the N=16 chain
    (d=12,870) to tol 1e-12, the reference's CPU-runnable case."""
    H16 = qk.MatrixFreeHamiltonian(qk.SpinHalfBasis(16, sz=0), qk.SpinHalfSite(),
                                   lat.heisenberg_chain(16))
    qk.lanczos_ground_state(H16, maxiter=5)       # warm-up
    t = time.perf_counter()
    r16 = qk.lanczos_ground_state(H16, maxiter=200, tol=1e-12)
    n16 = {"time_to_ground_state_s": time.perf_counter() - t, "E0": r16.energy,
           "iterations": r16.iterations}
    log(f"lanczos N=16: {n16}")
    t = time.perf_counter()
    H = qk.MatrixFreeHamiltonian(qk.SpinHalfBasis(24, sz=0), qk.SpinHalfSite(),
                                 lat.j1j2_square(6, 4, 1.0, 0.5))
    setup = time.perf_counter() - t
    t = time.perf_counter()
    r = qk.lanczos_ground_state(H, maxiter=300, tol=0.0)
    dt = time.perf_counter() - t
    v = r.eigenvector
    res = float(np.linalg.norm(H.apply(v) - r.energy * v) / np.linalg.norm(v))
    log(f"lanczos J1-J2 6x4: E0={r.energy:.12f} iterations={r.iterations} time={dt:.2f} s "
        f"|H psi - E psi|={res:.2e}")
    return {"workload": "J1-J2 Heisenberg 6x4 periodic, J1=1, J2=0.5, Sz=0, d=2,704,156; "
                        "300 Lanczos steps, full reorthogonalisation, reference start vector",
            "dim": H.dimension, "setup_s": setup, "time_to_ground_state_s": dt,
            "iterations": r.iterations, "E0": r.energy, "E0_per_site": r.energy / 24,
            "ritz_residual": res, "chain_n16": n16}


def bench_davidson_tj(qk, lat, msub=48, max_iter=400, phase="all"):
    """t-J on the 4x4 periodic square lattice (t=1, J=0.4, 7 up, 7 down,
    d=411,840), lowest 8 eigenpairs to residual 1e-8: Jacobi Davidson against
    the damped first-order correction (untrained network) and the learned
    second-order correction trained on 3x4 (d=16,632).  phase "train": only the
    training on 3x4; "solve": only the 4x4 solves, with the untrained network
    (each lattice profiled on its own)."""
    def tj(lx, ly, k):
        ops = lat.tj_model(lat.square_bonds(lx, ly), 1.0, 0.4)
        sec = qk.Sector().set_hubbard_particles(k, k)
        return qk.MatrixFreeHamiltonian(qk.TJBasis(lx * ly, sector=sec), qk.TJSite(), ops)

    Hs = tj(3, 4, 5) if phase != "solve" else None
    H = tj(4, 4, 7) if phase != "train" else None
    # untimed: kernel choice (row-program or group loop, vectors per launch)
    tuning = {}
    if Hs is not None:
        tuning["3x4"] = Hs.autotune(8)
    if H is not None:
        tuning["4x4"] = H.autotune(8)
    log(f"davidson autotune: {tuning}")
    net = qk.MLPPreconditioner(7)
    t_train, hist = 0.0, []
    if Hs is not None:
        t = time.perf_counter()
        hist = net.train(Hs, n_eig=8, max_subspace=48, max_iter=30, max_pairs=192, epochs=20,
                         batch_pairs=8, lr=1e-3)
        t_train = time.perf_counter() - t
        log(f"mlp training on 3x4: {t_train:.1f} s {net.last_train}")
    if phase == "train":
        return {"workload": "training on t-J 3x4 (d=16,632)", "train_s": t_train, "train": net.last_train,
                "loss_history": [float(x) for x in hist]}
    out = {"workload": "t-J 4x4 periodic, t=1, J=0.4, 7 up 7 down, d=411,840; lowest 8, "
                       f"residual 1e-8, max_subspace {msub}, at most {max_iter} iterations",
           "dim": H.dimension, "train_s": t_train, "train": net.last_train, "autotune": tuning,
           "loss_history": [float(x) for x in hist]}
    res = {}
    # each variant twice (deterministic: the same iterations), the faster
    # time reported, so that first-use allocations do not favour one variant
    for name, pre in (("jacobi", None), ("neumann", qk.MLPPreconditioner(7)), ("mlp", net)):
        times = []
        for _ in range(2):
            t = time.perf_counter()
            r = qk.davidson_lowest_ex(H, 8, msub, 1e-8, max_iter, pre)
            times.append(time.perf_counter() - t)
        dt = min(times)
        res[name] = r
        out[name] = {"iterations": r.iterations, "converged": r.converged, "time_s": dt,
                     "times_s": times, "eigenvalues": [float(x) for x in r.eigenvalues]}
        log(f"davidson 4x4 {name}: iterations={r.iterations} converged={r.converged} {dt:.2f} s "
            f"(runs {times[0]:.2f}, {times[1]:.2f} s)")
    out["max_eig_diff"] = max(float(np.max(np.abs(np.asarray(res[k].eigenvalues)
                                                  - np.asarray(res["jacobi"].eigenvalues))))
                              for k in ("neumann", "mlp"))
    out["speedup_time"] = out["jacobi"]["time_s"] / out["mlp"]["time_s"]
    out["speedup_iterations"] = out["jacobi"]["iterations"] / max(1, out["mlp"]["iterations"])
    out["trained_vs_untrained_iterations"] = out["neumann"]["iterations"] / max(1, out["mlp"]["iterations"])
    out["settings"] = (f"lowest 8, residual tol 1e-8, max_subspace {msub}, at most {max_iter} iterations "
                       "(the reference's max(20, 2k) = 20 and 100 iterations leave Jacobi Davidson "
                       "unconverged here); same settings for every variant")
    out["timing_note"] = ("time_s = solve time, the faster of two runs; training on 3x4 is not counted "
                          "(train_s); the learned preconditioner's per-Hamiltonian tables are built on "
                          "first use and kept with the Hamiltonian (the neumann variant's first run "
                          "builds them)")
    return out


def bench_tj_apply(qk, torch, lat, reps=200):
    """y = H x for the t-J model on the 4x4 periodic lattice (t=1, J=0.4,
    7 up, 7 down, d=411,840), one vector and blocks of 8, after the
    autotuner's choice; times from HIP events on the apply's stream."""
    ops = lat.tj_model(lat.square_bonds(4, 4), 1.0, 0.4)
    sec = qk.Sector().set_hubbard_particles(7, 7)
    H = qk.MatrixFreeHamiltonian(qk.TJBasis(16, sector=sec), qk.TJSite(), ops)
    d = H.dimension
    tuning = H.autotune(8)
    out = {"workload": "t-J 4x4 periodic, t=1, J=0.4, 7 up 7 down, d=411,840; y = H x",
           "dim": d, "autotune": tuning}
    for nvec in (1, 8):
        ms, _ = bench_apply(qk, torch, H, d, nvec, reps, 10)
        gbs = BYTES_PER_STATE * d * nvec / (ms * 1e-3) / 1e9
        out[f"b{nvec}"] = {"ms_per_launch": ms, "ms_per_vector": ms / nvec, "achieved_GBs": gbs,
                           "frac": gbs / HBM_PEAK_GBS}
        log(f"t-J 4x4 apply b={nvec}: {ms:.4f} ms ({ms / nvec:.4f} ms per vector, {gbs:.0f} GB/s)")
    return out


def bench_dynamics(qk, torch, H, d):
    """N=28 chain: ground state by Lanczos (reference start, tol 1e-12), then
    S^z(q)|0> for the 28 momenta q = 2 pi k / 28 as one block of start vectors,
    200 continued-fraction steps each, S(q, w) on 2000 points in [0, 5],
    eta = 0.05."""
    t = time.perf_counter()
    gs = qk.lanczos_ground_state(H, maxiter=200, tol=1e-12)
    t_gs = time.perf_counter() - t
    log(f"N=28 ground state: E0={gs.energy:.12f} iterations={gs.iterations} {t_gs:.2f} s")
    psi = torch.from_numpy(gs.eigenvector).to("cuda")
    nq, M = N_SITES, 200
    phi = torch.empty((nq, d), dtype=torch.complex128, device="cuda")
    torch.cuda.synchronize()
    t = time.perf_counter()
    H.sz_q_device(psi.data_ptr(), phi.data_ptr(), nq, d, torch.cuda.current_stream().cuda_stream)
    H.sync()
    torch.cuda.synchronize()
    t_phi = time.perf_counter() - t
    del psi
    gs.eigenvector = None
    t = time.perf_counter()
    cf = qk.continued_fraction_block_device(H, phi.data_ptr(), nq, M)
    t_cf = time.perf_counter() - t
    del phi
    torch.cuda.empty_cache()
    omegas = np.linspace(0.0, 5.0, 2000)
    t = time.perf_counter()
    S = np.array([[qk.evaluate_spectral_function(c.alphas, c.betas, c.norm_phi0, w, gs.energy, 0.05)
                   for w in omegas] for c in cf])
    t_s = time.perf_counter() - t
    # sum rule: the weight inside [0, 5] against |phi_q|^2
    dw = omegas[1] - omegas[0]
    weight = S.sum(1) * dw
    norms2 = np.array([c.norm_phi0 ** 2 for c in cf])
    kpi = nq // 2
    peak = float(omegas[int(np.argmax(S[kpi]))])
    log(f"dynamics: phi_q {t_phi:.2f} s, CF block {nq}x{M} {t_cf:.2f} s, S(q,w) {t_s:.2f} s, "
        f"q=pi peak at w={peak:.3f}")
    return {"workload": "N=28 chain, Sz=0, d=40,116,600: Lanczos ground state (tol 1e-12), "
                        "S^z(q)|0> for 28 q as one block, 200 CF steps, S(q,w) 2000 points, eta 0.05",
            "E0": gs.energy, "E0_per_site": gs.energy / N_SITES, "gs_iterations": gs.iterations,
            "time_ground_state_s": t_gs, "time_phi_s": t_phi, "time_cf_block_s": t_cf,
            "time_spectra_s": t_s, "cf_steps": [len(c.alphas) for c in cf],
            "q_pi_peak_omega": peak,
            "sum_rule_ratio_q_pi": float(weight[kpi] / norms2[kpi]) if norms2[kpi] > 0 else None,
            "q0_norm": float(cf[0].norm_phi0)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--block", type=int, default=28)
    ap.add_argument("--skip-ftlm", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--apply-only", action="store_true", help="time the b=1 apply only")
    ap.add_argument("--ftlm-only", action="store_true", help="time the FTLM sweep only")
    ap.add_argument("--ftlm-batched-only", action="store_true",
                    help="with --ftlm-only: the batched sweep alone (for profiling)")
    ap.add_argument("--davidson-only", action="store_true",
                    help="run the t-J 4x4 Davidson leg only (training included)")
    ap.add_argument("--dav-train-only", action="store_true",
                    help="with --davidson-only: only the training on 3x4")
    ap.add_argument("--dav-solve-only", action="store_true",
                    help="with --davidson-only: only the 4x4 solves (untrained network)")
    ap.add_argument("--skip-lanczos", action="store_true", help="skip the J1-J2 6x4 Lanczos run")
    ap.add_argument("--skip-davidson", action="store_true", help="skip the t-J 4x4 Davidson runs")
    ap.add_argument("--skip-dynamics", action="store_true", help="skip the N=28 dynamics run")
    ap.add_argument("--tj-apply", action="store_true",
                    help="time the t-J 4x4 apply only (one vector and blocks of 8)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 1 GiB device to device 5 times (calibrates PMC byte counters), "
                         "then time copies and a read-only reduction of the same 1 GiB")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        rates = {}
        for name, f, nbytes in (("copy_GBs", lambda: y.copy_(x), 2 << 30),
                                ("read_GBs", lambda: x.sum(), 1 << 30)):
            f()
            e0.record()
            for _ in range(10):
                f()
            e1.record()
            torch.cuda.synchronize()
            rates[name] = nbytes / (e0.elapsed_time(e1) / 10 * 1e-3) / 1e9
        print(json.dumps({"copy_probe_bytes_read": 1 << 30, "copy_probe_bytes_written": 1 << 30, **rates}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the H·x path runs on one device (sharding across GPUs "
              "is described, not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import qkrylov_amd as qk
    from qkrylov_amd import lattices as lat

    log(f"device: {qk.device_info()}")
    if args.ftlm_only:
        print(json.dumps(bench_ftlm(qk, lat, batched_only=args.ftlm_batched_only)))
        return
    if args.davidson_only:
        phase = "train" if args.dav_train_only else ("solve" if args.dav_solve_only else "all")
        print(json.dumps(bench_davidson_tj(qk, lat, phase=phase)))
        return
    if args.tj_apply:
        print(json.dumps(bench_tj_apply(qk, torch, lat)))
        return
    t = time.time()
    basis = qk.SpinHalfBasis(N_SITES, sz=0)
    H = qk.MatrixFreeHamiltonian(basis, qk.SpinHalfSite(), lat.heisenberg_chain(N_SITES))
    d = H.dimension
    log(f"setup N={N_SITES} d={d} in {time.time() - t:.2f} s")
    # untimed: the library's autotuner picks the kernel variant for b=1 and
    # the block grouping for b=8 on this Hamiltonian
    t = time.time()
    tuning = H.autotune(min(8, max(1, args.block)))
    log(f"autotune in {time.time() - t:.2f} s: {tuning}")
    log(f"info={H.info()}")

    ms, launches = bench_apply(qk, torch, H, d, 1, args.steps, args.warmup)
    achieved = BYTES_PER_STATE * d / (ms * 1e-3) / 1e9
    log(f"b=1: {ms:.4f} ms/apply  {1e3 / ms:.1f} applies/s  {achieved:.1f} GB/s")
    extra = {}
    if not args.apply_only and args.block > 1:
        nb = args.block
        msb, lb = bench_apply(qk, torch, H, d, nb, max(3, args.steps // 10), 1)
        ach_b = BYTES_PER_STATE * d * nb / (msb * 1e-3) / 1e9
        extra["block"] = {"vectors": nb, "ms_per_block": msb, "vector_applies_per_s": nb * 1e3 / msb,
                          "achieved_GBs": ach_b, "frac": ach_b / HBM_PEAK_GBS}
        log(f"b={nb}: {msb:.3f} ms/block  {nb * 1e3 / msb:.1f} vector-applies/s  {ach_b:.1f} GB/s")
    def leg(name, fn):
        # a failing side run is reported in the line, it does not lose the headline
        try:
            extra[name] = fn()
        except Exception as e:  # noqa: BLE001
            log(f"{name} failed: {e!r}")
            extra[name] = {"error": repr(e)}
        torch.cuda.empty_cache()

    if not args.apply_only:
        if not args.skip_ftlm:
            leg("ftlm", lambda: bench_ftlm(qk, lat))
        if not args.skip_lanczos:
            leg("lanczos", lambda: bench_lanczos_j1j2(qk, lat))
        if not args.skip_davidson:
            leg("davidson", lambda: bench_davidson_tj(qk, lat))
        if not args.skip_dynamics:
            leg("dynamics", lambda: bench_dynamics(qk, torch, H, d))
    cpu = None
    if not args.apply_only and not args.skip_cpu:
        cpu = cpu_baseline(lat)
        log(f"cpu baseline: {cpu}")

    line = {
        "metric": "H·x applies/s (% HBM roofline); FTLM samples/s batched vs plain at 1e-10",
        "value": 1e3 / ms,
        "unit": "applies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "periodic Heisenberg chain N=28, J=1, Sz=0 sector, d=40,116,600; "
                               "y = H x, one complex FP64 vector (seeded normal) resident in HBM",
                   "nsites": N_SITES, "dim": d, "vectors_per_step": 1, "parallelism": "1 GPU",
                   "kernel": {k: H.info()[k] for k in ("states_per_lane", "bond_batch", "vec_group")}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_from_profiles("k_apply", 1),
                     "traffic_scope": "L2<->memory bytes per apply (MALL hits included; "
                                      "upper bound on HBM bytes)"},
        "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
